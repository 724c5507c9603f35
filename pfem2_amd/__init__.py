"""pfem2_amd — the PFEM2 particle stage on MI355X (gfx950), driven through a C ABI.

The native library `libpfem2_amd.so` (hand-written HIP, built in-tree by
`python -m pfem2_amd.build`) holds one simulation per handle.  This module is
the Python mirror of the reference's particle interface
(pfem2Solver / pfem2ParticleHandler in pfem2particle.h): the method names
are the reference's, the mesh and the nodal vectors are numpy arrays.

There is no fallback: if the library is missing or no GPU is visible the
calls raise.
"""
import ctypes
import os

import numpy as np

from . import mesh  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpfem2_amd.so")

FP64 = 0
FP32_MIXED = 1

_lib = None


class Pf2Error(RuntimeError):
    pass


class _Mesh(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int), ("n_cells", ctypes.c_int), ("n_vertices", ctypes.c_int),
                ("n_dofs", ctypes.c_int), ("vertices", ctypes.c_void_p), ("cell_vertices", ctypes.c_void_p),
                ("cell_dofs", ctypes.c_void_p), ("degree", ctypes.c_int)]


def _mesh_arrays(m):
    """Contiguous vertex, cell-vertex and (degree 2) cell-DoF arrays of a
    mesh object; the caller keeps them alive while the struct is in use."""
    vertices = _f64(m.vertices)
    cells = np.ascontiguousarray(m.cells, dtype=np.int32)
    deg = getattr(m, "degree", 1)
    cd = np.ascontiguousarray(m.cell_dofs, dtype=np.int32) if deg == 2 else None
    return vertices, cells, cd


def _mesh_struct(m, arrays):
    vertices, cells, cd = arrays
    return _Mesh(m.dim, m.n_cells, m.n_vertices, m.n_dofs, _ptr(vertices), _ptr(cells), _ptr(cd),
                 getattr(m, "degree", 1))


class _Params(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_double), ("substeps", ctypes.c_int), ("max_per_part", ctypes.c_int),
                ("quantities", ctypes.c_int * 3), ("precision", ctypes.c_int),
                ("keep_velocity_ext", ctypes.c_int)]


# every entry point of include/pfem2_amd.h
EXPORTS = [
    "pf2_last_error", "pf2_create", "pf2_destroy", "pf2_set_params", "pf2_upload_fields",
    "pf2_upload_fields_device", "pf2_download_field", "pf2_seed_particles", "pf2_insert_particles",
    "pf2_correct_particles_velocities", "pf2_move_particles", "pf2_distribute_particle_velocities_to_grid",
    "pf2_distribute_particle_velocities_to_grid_static", "pf2_step", "pf2_set_walls",
    "pf2_sort_particles_into_subdomains_and_cells", "pf2_check_empty_cells", "pf2_reinterpolate_fields",
    "pf2_calculate_loads", "pf2_batch_create", "pf2_batch_step", "pf2_batch_destroy",
    "pf2_update_vertices", "pf2_initialize_maps", "pf2_n_global_particles", "pf2_download_particles",
    "pf2_cell_counts", "pf2_particle_counter", "pf2_stats", "pf2_synchronize", "pf2_set_profiling",
    "pf2_kernel_times", "pf2_kernel_name", "pf2_step_info", "pf2_create_in", "pf2_estimate_bytes",
    "pf2_memory_bytes", "pf2_step_with_field", "pf2_set_graphs", "pf2_move_mesh", "pf2_set_motion",
    "pf2_move_mesh_scaled", "pf2_server_submit_moving",
    "pf2_server_last_error", "pf2_server_create", "pf2_server_destroy", "pf2_server_submit", "pf2_server_poll",
    "pf2_server_drain", "pf2_server_result", "pf2_server_stats",
]


class _LoadsParams(ctypes.Structure):
    _fields_ = [("mu", ctypes.c_double), ("rho", ctypes.c_double), ("u_mean", ctypes.c_double),
                ("diam", ctypes.c_double), ("dof_a", ctypes.c_int), ("dof_e", ctypes.c_int)]


class _ServerConfig(ctypes.Structure):
    _fields_ = [("hbm_budget", ctypes.c_ulonglong), ("max_active", ctypes.c_int),
                ("steps_per_chunk", ctypes.c_int), ("chunks_ahead", ctypes.c_int), ("use_graphs", ctypes.c_int),
                ("batched", ctypes.c_int), ("batch_max_cells", ctypes.c_int)]


def load_library(path=LIB_PATH):
    """Load the HIP library (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise Pf2Error(f"{path} is missing: build it with `python -m pfem2_amd.build`")
    lib = ctypes.CDLL(path)
    vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.pf2_last_error.restype = ctypes.c_char_p
    lib.pf2_kernel_name.restype = ctypes.c_char_p
    lib.pf2_kernel_name.argtypes = [ip]
    lib.pf2_create.argtypes = [ctypes.POINTER(_Mesh), ctypes.POINTER(_Params), ctypes.c_longlong,
                               ctypes.POINTER(vp)]
    lib.pf2_destroy.argtypes = [vp]
    lib.pf2_set_params.argtypes = [vp, ctypes.POINTER(_Params)]
    lib.pf2_upload_fields.argtypes = [vp, vp, vp]
    lib.pf2_upload_fields_device.argtypes = [vp, vp, vp]
    lib.pf2_download_field.argtypes = [vp, vp]
    lib.pf2_seed_particles.argtypes = [vp]
    lib.pf2_insert_particles.argtypes = [vp, ip, vp, vp, vp]
    for f in ("pf2_correct_particles_velocities", "pf2_move_particles",
              "pf2_sort_particles_into_subdomains_and_cells", "pf2_check_empty_cells", "pf2_initialize_maps",
              "pf2_synchronize"):
        getattr(lib, f).argtypes = [vp]
    lib.pf2_distribute_particle_velocities_to_grid.argtypes = [vp, vp]
    lib.pf2_distribute_particle_velocities_to_grid_static.argtypes = [vp, vp]
    lib.pf2_step.argtypes = [vp, vp]
    lib.pf2_set_walls.argtypes = [vp, ip, vp, vp, vp, vp]
    lib.pf2_update_vertices.argtypes = [vp, vp, ip]
    lib.pf2_reinterpolate_fields.argtypes = [vp, vp, ip, vp]
    lib.pf2_batch_create.argtypes = [ip, vp, ctypes.POINTER(vp)]
    lib.pf2_batch_step.argtypes = [vp, ip, vp, vp]
    lib.pf2_batch_destroy.argtypes = [vp]
    lib.pf2_calculate_loads.argtypes = [vp, vp, vp, ip, vp, ctypes.POINTER(_LoadsParams), vp]
    lib.pf2_move_mesh.argtypes = [vp, vp, ip]
    lib.pf2_set_motion.argtypes = [vp, vp]
    lib.pf2_move_mesh_scaled.argtypes = [vp, dp]
    lib.pf2_n_global_particles.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong)]
    lib.pf2_download_particles.argtypes = [vp, vp, vp, vp, vp, vp]
    lib.pf2_cell_counts.argtypes = [vp, vp]
    lib.pf2_particle_counter.argtypes = [vp, ctypes.POINTER(ctypes.c_uint32)]
    lib.pf2_stats.argtypes = [vp, vp]
    lib.pf2_set_profiling.argtypes = [vp, ip]
    lib.pf2_kernel_times.argtypes = [vp, vp, vp, ip]
    lib.pf2_step_info.argtypes = [vp, vp, ip]
    ull, ll = ctypes.c_ulonglong, ctypes.c_longlong
    lib.pf2_create_in.argtypes = [ctypes.POINTER(_Mesh), ctypes.POINTER(_Params), ll, vp, ull, vp,
                                  ctypes.POINTER(vp)]
    lib.pf2_estimate_bytes.argtypes = [ctypes.POINTER(_Mesh), ctypes.POINTER(_Params), ll, ctypes.POINTER(ull)]
    lib.pf2_memory_bytes.argtypes = [vp, ctypes.POINTER(ull)]
    lib.pf2_step_with_field.argtypes = [vp, vp]
    lib.pf2_set_graphs.argtypes = [vp, ip]
    lib.pf2_server_last_error.restype = ctypes.c_char_p
    lib.pf2_server_create.argtypes = [ctypes.POINTER(_ServerConfig), ctypes.POINTER(vp)]
    lib.pf2_server_destroy.argtypes = [vp]
    lib.pf2_server_submit.argtypes = [vp, ctypes.POINTER(_Mesh), ctypes.POINTER(_Params), vp, ip,
                                      ctypes.POINTER(ll)]
    lib.pf2_server_submit_moving.argtypes = [vp, ctypes.POINTER(_Mesh), ctypes.POINTER(_Params), vp, ip, vp, dp, dp,
                                             dp, ctypes.POINTER(ll)]
    lib.pf2_server_poll.argtypes = [vp, ctypes.POINTER(ip), ctypes.POINTER(ip)]
    lib.pf2_server_drain.argtypes = [vp]
    lib.pf2_server_result.argtypes = [vp, ll, vp, ctypes.POINTER(ll), ctypes.POINTER(dp)]
    lib.pf2_server_stats.argtypes = [vp, vp, ip]
    _lib = lib
    return lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != shape:
        a = a.reshape(shape)
    return a


def _params_of(dim, dt, substeps, max_per_part, quantities, precision=FP64, keep_velocity_ext=False):
    q = list(quantities) + [1] * (3 - len(quantities))
    return _Params(dt, substeps, max_per_part, (ctypes.c_int * 3)(*q[:3]), precision, 1 if keep_velocity_ext else 0)


def estimate_bytes(m, dt, substeps, max_per_part, quantities, precision=FP64, keep_velocity_ext=False, capacity=0):
    """Device bytes a simulation of this mesh and parameters takes (upper bound)."""
    lib = load_library()
    p = _params_of(m.dim, dt, substeps, max_per_part, quantities, precision, keep_velocity_ext)
    ms = _Mesh(m.dim, m.n_cells, m.n_vertices, m.n_dofs, None, None, None, getattr(m, "degree", 1))
    out = ctypes.c_ulonglong()
    rc = lib.pf2_estimate_bytes(ctypes.byref(ms), ctypes.byref(p), capacity, ctypes.byref(out))
    if rc != 0:
        raise Pf2Error(f"pf2 error {rc}: {lib.pf2_last_error().decode()}")
    return out.value


class ParticleStage:
    """One simulation's particle stage (one pf2 handle).

    Methods carry the reference's names: seed_particles,
    correct_particles_velocities, move_particles,
    distribute_particle_velocities_to_grid (pfem2Solver), and
    sort_particles_into_subdomains_and_cells, initialize_maps,
    n_global_particles (pfem2ParticleHandler).
    """

    def __init__(self, m, dt, substeps, max_per_part, quantities, precision=FP64, keep_velocity_ext=False,
                 capacity=0):
        self._lib = load_library()
        self.mesh = m
        self.dim = m.dim
        self._arrays = _mesh_arrays(m)
        q = list(quantities) + [1] * (3 - len(quantities))
        self._params = _Params(dt, substeps, max_per_part, (ctypes.c_int * 3)(*q[:3]), precision,
                               1 if keep_velocity_ext else 0)
        ms = _mesh_struct(m, self._arrays)
        h = ctypes.c_void_p()
        self._check(self._lib.pf2_create(ctypes.byref(ms), ctypes.byref(self._params), capacity, ctypes.byref(h)))
        self._h = h

    # -- plumbing --
    def _check(self, rc):
        if rc != 0:
            raise Pf2Error(f"pf2 error {rc}: {self._lib.pf2_last_error().decode()}")

    def close(self):
        if getattr(self, "_h", None):
            self._lib.pf2_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- parameters / fields --
    def set_params(self, dt=None, substeps=None, max_per_part=None, quantities=None):
        p = self._params
        if dt is not None:
            p.dt = dt
        if substeps is not None:
            p.substeps = substeps
        if max_per_part is not None:
            p.max_per_part = max_per_part
        if quantities is not None:
            q = list(quantities) + [1] * (3 - len(quantities))
            p.quantities = (ctypes.c_int * 3)(*q[:3])
        self._check(self._lib.pf2_set_params(self._h, ctypes.byref(p)))

    def upload_fields(self, V=None, V_old=None):
        """Nodal velocity V (solutionVx/Vy[/Vz]) and V_old, shape (dim, n_dofs)."""
        shape = (self.dim, self.mesh.n_dofs)
        a = None if V is None else _f64(V, shape)
        b = None if V_old is None else _f64(V_old, shape)
        self._check(self._lib.pf2_upload_fields(self._h, _ptr(a), _ptr(b)))

    def download_field(self):
        out = np.empty((self.dim, self.mesh.n_dofs))
        self._check(self._lib.pf2_download_field(self._h, _ptr(out)))
        return out

    # -- pfem2Solver particle methods --
    def seed_particles(self):
        self._check(self._lib.pf2_seed_particles(self._h))

    def insert_particles(self, cell, xi, v=None):
        cell = np.ascontiguousarray(cell, dtype=np.int32)
        xi = _f64(xi, (len(cell), self.dim))
        vv = None if v is None else _f64(v, (len(cell), self.dim))
        self._check(self._lib.pf2_insert_particles(self._h, len(cell), _ptr(cell), _ptr(xi), _ptr(vv)))

    def correct_particles_velocities(self):
        self._check(self._lib.pf2_correct_particles_velocities(self._h))

    def move_particles(self):
        self._check(self._lib.pf2_move_particles(self._h))

    def distribute_particle_velocities_to_grid(self, download=True):
        out = np.empty((self.dim, self.mesh.n_dofs)) if download else None
        self._check(self._lib.pf2_distribute_particle_velocities_to_grid(self._h, _ptr(out)))
        return out

    def distribute_particle_velocities_to_grid_static(self, V_device_ptr=0):
        """The projection with its per-cell sums in the library's static
        device array (no allocation); the field goes to device memory at
        V_device_ptr (dim x n_dofs, component-major; 0: kept resident only).
        No host synchronisation."""
        self._check(self._lib.pf2_distribute_particle_velocities_to_grid_static(
            self._h, ctypes.c_void_p(V_device_ptr) if V_device_ptr else None))

    def step(self, download=False):
        out = np.empty((self.dim, self.mesh.n_dofs)) if download else None
        self._check(self._lib.pf2_step(self._h, _ptr(out)))
        return out

    def step_with_field(self, V_device_ptr):
        """One step refreshing V from a device field (component-major) given
        by its address, e.g. a torch tensor's data_ptr(); no host sync."""
        self._check(self._lib.pf2_step_with_field(self._h, ctypes.c_void_p(V_device_ptr)))

    def set_graphs(self, on=True):
        self._check(self._lib.pf2_set_graphs(self._h, 1 if on else 0))

    def memory_bytes(self):
        out = ctypes.c_ulonglong()
        self._check(self._lib.pf2_memory_bytes(self._h, ctypes.byref(out)))
        return out.value

    def check_empty_cells(self):
        self._check(self._lib.pf2_check_empty_cells(self._h))

    def set_walls(self, vertex, dof, bc, mask):
        vertex = np.ascontiguousarray(vertex, dtype=np.int32)
        dof = np.ascontiguousarray(dof, dtype=np.int32)
        bc = _f64(bc, (len(vertex), self.dim))
        mask = np.ascontiguousarray(mask, dtype=np.int32)
        self._check(self._lib.pf2_set_walls(self._h, len(vertex), _ptr(vertex), _ptr(dof), _ptr(bc), _ptr(mask)))

    def update_vertices(self, vertices, recompute_maps=False):
        self._vertices = _f64(vertices, (self.mesh.n_vertices, self.dim))
        self._check(self._lib.pf2_update_vertices(self._h, _ptr(self._vertices), 1 if recompute_maps else 0))

    def move_mesh(self, vertices, recompute_maps=False):
        """transform_grid + sort_particles_into_subdomains_and_cells +
        check_empty_cells in one pass."""
        self._vertices = _f64(vertices, (self.mesh.n_vertices, self.dim))
        self._check(self._lib.pf2_move_mesh(self._h, _ptr(self._vertices), 1 if recompute_maps else 0))

    def set_motion(self, W):
        """Prescribed motion X0 + a W: the current vertices become X0."""
        self._W = _f64(W, (self.mesh.n_vertices, self.dim))
        self._check(self._lib.pf2_set_motion(self._h, _ptr(self._W)))

    def move_mesh_scaled(self, a):
        """move_mesh with the vertices X0 + a W formed on the device."""
        self._check(self._lib.pf2_move_mesh_scaled(self._h, float(a)))

    # -- pfem2ParticleHandler methods --
    def sort_particles_into_subdomains_and_cells(self):
        self._check(self._lib.pf2_sort_particles_into_subdomains_and_cells(self._h))

    def reinterpolate_fields(self, U, extra=None):
        """reinterpolate_fields: V (resident) and the extra nodal scalar
        fields (n_extra x n_dofs, returned updated) at the vertices displaced
        by U (n_vertices x dim), on the mesh before it moves."""
        U = _f64(U, (self.mesh.n_vertices, self.dim))
        if extra is None:
            self._check(self._lib.pf2_reinterpolate_fields(self._h, _ptr(U), 0, None))
            return None
        E = np.array(extra, dtype=np.float64, copy=True).reshape(-1, self.mesh.n_dofs)
        self._check(self._lib.pf2_reinterpolate_fields(self._h, _ptr(U), E.shape[0], _ptr(E)))
        return E

    def calculate_loads(self, P, faces, mu, rho, u_mean, diam, dof_a=-1, dof_e=-1, V=None):
        """calculate_loads over boundary faces (cell, face number):
        {Cx_nu, Cx_p, Cy_nu, Cy_p, P[dof_a] - P[dof_e]}; V None: the resident V."""
        P = _f64(P, (self.mesh.n_dofs,))
        faces = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 2)
        Vv = None if V is None else _f64(V, (self.dim, self.mesh.n_dofs))
        lp = _LoadsParams(mu, rho, u_mean, diam, dof_a, dof_e)
        out = np.zeros(5)
        self._check(self._lib.pf2_calculate_loads(self._h, _ptr(Vv), _ptr(P), len(faces), _ptr(faces),
                                                  ctypes.byref(lp), _ptr(out)))
        return out

    def initialize_maps(self):
        self._check(self._lib.pf2_initialize_maps(self._h))

    def n_global_particles(self):
        n = ctypes.c_longlong()
        self._check(self._lib.pf2_n_global_particles(self._h, ctypes.byref(n)))
        return n.value

    def particles(self):
        """cell, id, x, xi, v of all particles (cell-major, ascending id inside a cell)."""
        n = self.n_global_particles()
        cell = np.empty(n, dtype=np.int32)
        pid = np.empty(n, dtype=np.uint32)
        x = np.empty((n, self.dim))
        xi = np.empty((n, self.dim))
        v = np.empty((n, self.dim))
        self._check(self._lib.pf2_download_particles(self._h, _ptr(cell), _ptr(pid), _ptr(x), _ptr(xi), _ptr(v)))
        return {"cell": cell, "id": pid, "x": x, "xi": xi, "v": v}

    def cell_counts(self):
        out = np.empty(self.mesh.n_cells, dtype=np.int32)
        self._check(self._lib.pf2_cell_counts(self._h, _ptr(out)))
        return out

    def particle_counter(self):
        c = ctypes.c_uint32()
        self._check(self._lib.pf2_particle_counter(self._h, ctypes.byref(c)))
        return c.value

    def stats(self):
        out = np.zeros(5, dtype=np.uint64)
        self._check(self._lib.pf2_stats(self._h, _ptr(out)))
        return dict(zip(("lost", "deleted", "spawned", "searches", "zero_weight"), out.tolist()))

    def synchronize(self):
        self._check(self._lib.pf2_synchronize(self._h))

    def step_info(self):
        """Work counts of the last move: particles before/after, particles queued
        for the vertex-star search per relocation round (queued[0], queued[1];
        queued[2]: the last re-sort after a mesh move), particles that changed
        cell."""
        out = np.zeros(8, dtype=np.int64)
        k = self._lib.pf2_step_info(self._h, _ptr(out), 8)
        if k < 0:
            self._check(k)
        return {"n_before": int(out[0]), "n_after": int(out[1]),
                "queued": [int(v) for v in out[2:5]], "arrivals": int(out[5])}

    # -- per-kernel HIP-event timing --
    def set_profiling(self, on=True):
        self._check(self._lib.pf2_set_profiling(self._h, 1 if on else 0))

    def kernel_times(self):
        ms = np.zeros(32)
        n = np.zeros(32, dtype=np.int64)
        k = self._lib.pf2_kernel_times(self._h, _ptr(ms), _ptr(n), 32)
        if k < 0:
            self._check(k)
        return {self._lib.pf2_kernel_name(i).decode(): (float(ms[i]), int(n[i])) for i in range(k) if n[i] > 0}


class Batch:
    """Steps of several ParticleStage simulations (one dimension and
    precision, Q1 fields, no walls) with one launch per kernel for all of
    them (pf2_batch_step); each one's results equal its own step's."""

    def __init__(self, max_sims):
        self._lib = load_library()
        b = ctypes.c_void_p()
        rc = self._lib.pf2_batch_create(max_sims, None, ctypes.byref(b))
        if rc != 0:
            raise Pf2Error(f"pf2 error {rc}: {self._lib.pf2_last_error().decode()}")
        self._b = b

    def step(self, stages, field_ptrs):
        """One step of every stage, stage i refreshing V from the device
        field at field_ptrs[i] (component-major, e.g. a tensor's data_ptr())."""
        n = len(stages)
        hs = (ctypes.c_void_p * n)(*[st._h for st in stages])
        fs = (ctypes.c_void_p * n)(*field_ptrs)
        rc = self._lib.pf2_batch_step(self._b, n, hs, fs)
        if rc != 0:
            raise Pf2Error(f"pf2 error {rc}: {self._lib.pf2_last_error().decode()}")

    def close(self):
        if getattr(self, "_b", None):
            self._lib.pf2_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Pfem2Solver:
    """Mirror of the reference's pfem2Solver particle interface
    (cylinder2D/pfem2particle.h:115-192): the driver sets `solution` (nodal V,
    shape (dim, n_dofs)) and calls the four particle methods in the
    reference's order; `old_solution` is kept on the device and equals the
    last projected field."""

    def __init__(self, m, time_step, substeps=3, max_per_part=3, precision=FP64, keep_velocity_ext=False):
        self.mesh = m
        self.time_step = time_step
        self._substeps = substeps
        self._max = max_per_part
        self._precision = precision
        self._keep = keep_velocity_ext
        self.stage = None
        self.particle_handler = None

    def seed_particles(self, quantities, V):
        self.stage = ParticleStage(self.mesh, self.time_step, self._substeps, self._max, quantities,
                                   self._precision, self._keep)
        self.stage.upload_fields(V, V)
        self.stage.seed_particles()
        self.particle_handler = self.stage

    def set_solution(self, V):
        self.stage.upload_fields(V, None)

    def correct_particles_velocities(self):
        self.stage.correct_particles_velocities()

    def move_particles(self):
        self.stage.move_particles()

    def distribute_particle_velocities_to_grid(self):
        return self.stage.distribute_particle_velocities_to_grid(download=True)


class Server:
    """Serving loop for independent simulations on one GPU (pf2_server_*).

    submit() copies a request (mesh, parameters, fixed field V, step count);
    poll() admits, issues and retires without waiting for the GPU; result()
    returns the last projected field, the final particle count and the
    latency from submission to completion.  batched: all active simulations
    of one dimension and precision (of at most batch_max_cells cells, when
    given) advance together, one launch per kernel of the step
    (pf2_batch_step); otherwise each replays its own step graph on its own
    stream."""

    def __init__(self, hbm_budget=0, max_active=0, steps_per_chunk=0, chunks_ahead=0, use_graphs=True, batched=False,
                 batch_max_cells=0):
        self._lib = load_library()
        cfg = _ServerConfig(hbm_budget, max_active, steps_per_chunk, chunks_ahead, 1 if use_graphs else 0,
                            1 if batched else 0, int(batch_max_cells))
        h = ctypes.c_void_p()
        self._check(self._lib.pf2_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._s = h
        self._dims = {}

    def _check(self, rc):
        if rc < 0:
            raise Pf2Error(f"pf2 server error {rc}: {self._lib.pf2_server_last_error().decode()}")

    def close(self):
        if getattr(self, "_s", None):
            self._lib.pf2_server_destroy(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, m, V, n_steps, dt, substeps, max_per_part, quantities, precision=FP64):
        arrays = _mesh_arrays(m)
        V = _f64(V, (m.dim, m.n_dofs))
        ms = _mesh_struct(m, arrays)
        p = _params_of(m.dim, dt, substeps, max_per_part, quantities, precision)
        rid = ctypes.c_longlong()
        self._check(self._lib.pf2_server_submit(self._s, ctypes.byref(ms), ctypes.byref(p), _ptr(V), n_steps,
                                                ctypes.byref(rid)))
        self._dims[rid.value] = (m.dim, m.n_dofs)
        return rid.value

    def submit_moving(self, m, V, n_steps, dt, substeps, max_per_part, quantities, W, amplitude, omega, phase=0.0,
                      precision=FP64):
        """A request whose mesh moves to X0 + amplitude sin(omega k dt + phase) W
        after each step k."""
        arrays = _mesh_arrays(m)
        V = _f64(V, (m.dim, m.n_dofs))
        W = _f64(W, (m.n_vertices, m.dim))
        ms = _mesh_struct(m, arrays)
        p = _params_of(m.dim, dt, substeps, max_per_part, quantities, precision, True)
        rid = ctypes.c_longlong()
        self._check(self._lib.pf2_server_submit_moving(self._s, ctypes.byref(ms), ctypes.byref(p), _ptr(V), n_steps,
                                                       _ptr(W), float(amplitude), float(omega), float(phase),
                                                       ctypes.byref(rid)))
        self._dims[rid.value] = (m.dim, m.n_dofs)
        return rid.value

    def poll(self):
        npend, nact = ctypes.c_int(), ctypes.c_int()
        self._check(self._lib.pf2_server_poll(self._s, ctypes.byref(npend), ctypes.byref(nact)))
        return npend.value, nact.value

    def drain(self):
        self._check(self._lib.pf2_server_drain(self._s))

    def result(self, rid):
        """(V, n_particles, latency_ms), or None while the request runs."""
        dim, nd = self._dims[rid]
        V = np.empty((dim, nd))
        n = ctypes.c_longlong()
        lat = ctypes.c_double()
        rc = self._lib.pf2_server_result(self._s, rid, _ptr(V), ctypes.byref(n), ctypes.byref(lat))
        if rc == -3:
            return None
        self._check(rc)
        return V, n.value, lat.value

    def stats(self):
        out = np.zeros(10)
        self._check(self._lib.pf2_server_stats(self._s, _ptr(out), 10))
        keys = ("finished", "failed", "p50_ms", "p99_ms", "particle_updates", "span_s", "updates_per_s",
                "peak_bytes", "max_active", "mean_ms")
        return dict(zip(keys, out.tolist()))
