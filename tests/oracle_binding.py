"""ctypes binding of the CPU oracle (oracle/pf2_oracle.cpp) — test infrastructure.

Only tests, __graft_entry__.smoke() and bench.py's CPU-baseline leg use it.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_PATH = os.path.join(ROOT, "oracle", "build", "libpf2_oracle.so")

_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(ORACLE_PATH):
        raise RuntimeError(f"{ORACLE_PATH} missing: run `make -C oracle`")
    lib = ctypes.CDLL(ORACLE_PATH)
    vp, ip = ctypes.c_void_p, ctypes.c_int
    lib.orc_create.restype = vp
    lib.orc_create.argtypes = [ip, ip, ip, vp, vp, ip, vp, ip]
    lib.orc_destroy.argtypes = [vp]
    lib.orc_set_params.argtypes = [vp, ctypes.c_double, ip, ip, vp]
    lib.orc_set_filter.argtypes = [vp, ip]
    lib.orc_set_fields.argtypes = [vp, vp, vp]
    lib.orc_get_field.argtypes = [vp, vp]
    for f in ("orc_seed", "orc_correct", "orc_sort", "orc_check_empty_cells", "orc_move", "orc_initialize_maps"):
        getattr(lib, f).argtypes = [vp]
    lib.orc_insert.argtypes = [vp, ip, vp, vp, vp]
    lib.orc_project.argtypes = [vp, vp]
    lib.orc_step.argtypes = [vp, vp]
    lib.orc_set_walls.argtypes = [vp, ip, vp, vp, vp, vp]
    lib.orc_update_vertices.argtypes = [vp, vp, ip]
    lib.orc_count.restype = ctypes.c_longlong
    lib.orc_count.argtypes = [vp]
    lib.orc_particle_counter.restype = ctypes.c_uint32
    lib.orc_particle_counter.argtypes = [vp]
    lib.orc_download.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.orc_cell_counts.argtypes = [vp, vp]
    lib.orc_stats.argtypes = [vp, vp]
    lib.orc_inverse_map.argtypes = [ip, vp, vp, ip, vp]
    lib.orc_inverse_newton.argtypes = [ip, vp, vp, vp]
    lib.orc_inverse_closed_form.argtypes = [vp, vp, vp]
    lib.orc_clearly_outside.argtypes = [ip, vp, vp]
    lib.orc_forward_map.argtypes = [ip, vp, vp, vp]
    lib.orc_shape_values.argtypes = [ip, vp, vp]
    lib.orc_field_shape_values.argtypes = [ip, ip, vp, vp]
    lib.orc_star.argtypes = [vp, ip, vp, vp, ip]
    lib.orc_time_steps.restype = ctypes.c_double
    lib.orc_time_steps.argtypes = [vp, ip]
    lib.orc_set_reference_mode.restype = ctypes.c_int
    lib.orc_set_reference_mode.argtypes = [vp, ip]
    lib.orc_ref_order_check.argtypes = [vp, vp]
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """CPU restatement of the particle stage; same call surface as ParticleStage."""

    def __init__(self, m, dt, substeps, max_per_part, quantities, filter=True):
        self.lib = load()
        self.mesh = m
        self.dim = m.dim
        self._v = _f(m.vertices)
        self._c = np.ascontiguousarray(m.cells, dtype=np.int32)
        deg = getattr(m, "degree", 1)
        self._cd = np.ascontiguousarray(m.cell_dofs, dtype=np.int32) if deg == 2 else None
        self.h = self.lib.orc_create(m.dim, m.n_cells, m.n_vertices, _p(self._v), _p(self._c), m.n_dofs,
                                     _p(self._cd), deg)
        if not self.h:
            raise RuntimeError("orc_create rejected the mesh")
        q = np.ascontiguousarray((list(quantities) + [1, 1, 1])[:3], dtype=np.int32)
        self.lib.orc_set_params(self.h, dt, substeps, max_per_part, _p(q))
        self.lib.orc_set_filter(self.h, 1 if filter else 0)

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, dt, substeps, max_per_part, quantities):
        q = np.ascontiguousarray((list(quantities) + [1, 1, 1])[:3], dtype=np.int32)
        self.lib.orc_set_params(self.h, dt, substeps, max_per_part, _p(q))

    def upload_fields(self, V=None, V_old=None):
        a = None if V is None else _f(V)
        b = None if V_old is None else _f(V_old)
        self.lib.orc_set_fields(self.h, _p(a), _p(b))

    def field(self):
        out = np.empty((self.dim, self.mesh.n_dofs))
        self.lib.orc_get_field(self.h, _p(out))
        return out

    def seed_particles(self):
        self.lib.orc_seed(self.h)

    def insert_particles(self, cell, xi, v=None):
        cell = np.ascontiguousarray(cell, dtype=np.int32)
        xi = _f(xi)
        vv = None if v is None else _f(v)
        self.lib.orc_insert(self.h, len(cell), _p(cell), _p(xi), _p(vv))

    def correct_particles_velocities(self):
        self.lib.orc_correct(self.h)

    def move_particles(self):
        self.lib.orc_move(self.h)

    def sort_particles_into_subdomains_and_cells(self):
        self.lib.orc_sort(self.h)

    def check_empty_cells(self):
        self.lib.orc_check_empty_cells(self.h)

    def distribute_particle_velocities_to_grid(self):
        out = np.empty((self.dim, self.mesh.n_dofs))
        self.lib.orc_project(self.h, _p(out))
        return out

    def step(self):
        out = np.empty((self.dim, self.mesh.n_dofs))
        self.lib.orc_step(self.h, _p(out))
        return out

    def set_walls(self, vertex, dof, bc, mask):
        vertex = np.ascontiguousarray(vertex, dtype=np.int32)
        dof = np.ascontiguousarray(dof, dtype=np.int32)
        bc = _f(bc)
        mask = np.ascontiguousarray(mask, dtype=np.int32)
        self.lib.orc_set_walls(self.h, len(vertex), _p(vertex), _p(dof), _p(bc), _p(mask))

    def update_vertices(self, vertices, recompute_maps=False):
        self._v = _f(vertices)
        self.lib.orc_update_vertices(self.h, _p(self._v), 1 if recompute_maps else 0)

    def initialize_maps(self):
        self.lib.orc_initialize_maps(self.h)

    def set_reference_mode(self, on=True):
        """Keep the particles in a std::unordered_multimap keyed by cell and
        visit them in its order, as the reference's particle handler does
        (before any particle exists)."""
        if self.lib.orc_set_reference_mode(self.h, 1 if on else 0) != 0:
            raise RuntimeError("reference container mode is set before seeding")

    def ref_order_check(self):
        """Reference container mode: (cells whose equal_range is not in
        descending insertion order, particles visited)."""
        out = np.zeros(2, dtype=np.int64)
        self.lib.orc_ref_order_check(self.h, _p(out))
        return int(out[0]), int(out[1])

    def n_global_particles(self):
        return int(self.lib.orc_count(self.h))

    def particle_counter(self):
        return int(self.lib.orc_particle_counter(self.h))

    def particles(self):
        n = self.n_global_particles()
        cell = np.empty(n, dtype=np.int32)
        pid = np.empty(n, dtype=np.uint32)
        x = np.empty((n, self.dim))
        xi = np.empty((n, self.dim))
        v = np.empty((n, self.dim))
        self.lib.orc_download(self.h, _p(cell), _p(pid), _p(x), _p(xi), _p(v))
        return {"cell": cell, "id": pid, "x": x, "xi": xi, "v": v}

    def cell_counts(self):
        out = np.empty(self.mesh.n_cells, dtype=np.int32)
        self.lib.orc_cell_counts(self.h, _p(out))
        return out

    def stats(self):
        out = np.zeros(5, dtype=np.uint64)
        self.lib.orc_stats(self.h, _p(out))
        return dict(zip(("lost", "deleted", "spawned", "searches", "newton"), out.tolist()))

    def time_steps(self, steps):
        return float(self.lib.orc_time_steps(self.h, steps))


# ---- geometry probes ------------------------------------------------------

def inverse_map(P, p, use_filter=True):
    lib = load()
    P = _f(P)
    p = _f(p)
    out = np.zeros(P.shape[1])
    ok = lib.orc_inverse_map(P.shape[1], _p(P), _p(p), 1 if use_filter else 0, _p(out))
    return bool(ok), out


def inverse_newton(P, p):
    lib = load()
    P = _f(P)
    p = _f(p)
    out = np.zeros(P.shape[1])
    ok = lib.orc_inverse_newton(P.shape[1], _p(P), _p(p), _p(out))
    return bool(ok), out


def inverse_closed_form(P, p):
    lib = load()
    P = _f(P)
    p = _f(p)
    out = np.zeros(2)
    ok = lib.orc_inverse_closed_form(_p(P), _p(p), _p(out))
    return bool(ok), out


def clearly_outside(P, p):
    lib = load()
    P = _f(P)
    p = _f(p)
    return bool(lib.orc_clearly_outside(P.shape[1], _p(P), _p(p)))


def forward_map(P, xi):
    lib = load()
    P = _f(P)
    xi = _f(xi)
    out = np.zeros(P.shape[1])
    lib.orc_forward_map(P.shape[1], _p(P), _p(xi), _p(out))
    return out


def shape_values(dim, xi):
    lib = load()
    xi = _f(xi)
    out = np.zeros(1 << dim)
    lib.orc_shape_values(dim, _p(xi), _p(out))
    return out


def field_shape_values(dim, degree, xi):
    """Shape values of a cell's velocity DoFs (degree 1 or 2) at xi, DoF order."""
    lib = load()
    xi = _f(xi)
    out = np.zeros((1 << dim) if degree == 1 else 3 ** dim)
    lib.orc_field_shape_values(dim, degree, _p(xi), _p(out))
    return out
