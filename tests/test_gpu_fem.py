"""reinterpolate_fields and calculate_loads on the GPU against the oracle:
bitwise equal nodal fields and force coefficients."""
import ctypes

import numpy as np
import pytest

from oracle_binding import Oracle, load
from parity_util import bitwise_equal
from pfem2_amd import FP64, FP32_MIXED, Pf2Error, ParticleStage
from pfem2_amd import mesh as M

pytestmark = pytest.mark.gpu


def _oracle_reinterp(o, U, extra):
    lib = load()
    lib.orc_reinterpolate_fields.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    U = np.ascontiguousarray(U, dtype=np.float64)
    E = np.ascontiguousarray(extra, dtype=np.float64).copy()
    lib.orc_reinterpolate_fields(o.h, U.ctypes.data, E.shape[0], E.ctypes.data)
    return o.field(), E


def _oracle_loads(o, V, P, faces, mu, rho, u, d, a, e):
    lib = load()
    dp, ip, vp = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
    lib.orc_calculate_loads.argtypes = [vp, vp, vp, ip, vp, dp, dp, dp, dp, ip, ip, vp]
    V = np.ascontiguousarray(V, dtype=np.float64)
    P = np.ascontiguousarray(P, dtype=np.float64)
    F = np.ascontiguousarray(faces, dtype=np.int32)
    out = np.zeros(5)
    lib.orc_calculate_loads(o.h, V.ctypes.data, P.ctypes.data, len(F), F.ctypes.data, mu, rho, u, d, a, e,
                            out.ctypes.data)
    return out


@pytest.mark.parametrize("case", ["cylinder", "large", "box3d"])
@pytest.mark.parametrize("precision", [FP64, FP32_MIXED])
def test_reinterpolation_equals_oracle(case, precision):
    rng = np.random.default_rng(61)
    if case == "box3d":
        m = M.box_3d(5, 4, 4, (0, 0, 0), (1.0, 0.8, 0.8), jitter=0.15, seed=62)
        V = M.fourier_perturbation(m.vertices, [1.0, 0.8, 0.8], seed=63)
        U = rng.uniform(-0.05, 0.05, size=m.vertices.shape)
    else:
        m = M.cylinder_channel(0.12, seed=64)
        V = M.cylinder_flow(m, 1.0, eps=0.02, seed=65)
        r = np.hypot(m.vertices[:, 0] - 0.2, m.vertices[:, 1] - 0.2)
        w = 0.5 * (1.0 + np.cos(np.pi * np.clip((r - 0.05) / 0.1, 0.0, 1.0)))
        U = np.zeros_like(m.vertices)
        U[:, 1] = 0.01 * w
        if case == "large":
            # displacements of about a cell: many vertices land in a neighbour
            U = U + rng.uniform(-0.004, 0.004, size=U.shape)
    P = rng.normal(size=(1, m.n_dofs))
    g = ParticleStage(m, 0.01, 2, 3, [2] * m.dim, precision)
    o = Oracle(m, 0.01, 2, 3, [2] * m.dim)
    for s in (g, o):
        s.upload_fields(V, V)
    Eg = g.reinterpolate_fields(U, P)
    Vo, Eo = _oracle_reinterp(o, U, P)
    assert bitwise_equal(g.download_field(), Vo)
    assert bitwise_equal(Eg, Eo)
    print(case, "reinterpolated fields bitwise equal; max change", float(np.max(np.abs(Vo - V))), flush=True)
    # the next step uses the reinterpolated V (V_old unchanged)
    cell, xi = M.jittered_lattice(m, [2] * m.dim, 0.05, 66)
    for s in (g, o):
        s.insert_particles(cell, xi)
    Vg1, Vo1 = g.step(download=True), o.step()
    if precision == FP64:
        assert bitwise_equal(Vg1, Vo1)
    else:
        assert np.linalg.norm(Vg1 - Vo1) <= 1e-5 * np.linalg.norm(Vo1)


def test_loads_equal_oracle():
    m = M.cylinder_channel(0.2, seed=67)
    V = M.cylinder_flow(m, 1.0, eps=0.02, seed=68)
    rng = np.random.default_rng(69)
    P = 0.3 * rng.normal(size=m.n_dofs) + 1.0
    faces = M.boundary_faces(m, "cylinder")
    assert len(faces) > 10
    g = ParticleStage(m, 0.01, 2, 3, [2, 2], FP64)
    o = Oracle(m, 0.01, 2, 3, [2, 2])
    args = (0.001, 1.0, 0.2, 0.1, 3, 17)
    a = g.calculate_loads(P, faces, *args, V=V)
    b = _oracle_loads(o, V, P, faces, *args)
    assert bitwise_equal(a, b), (a, b)
    g.upload_fields(V, V)
    c = g.calculate_loads(P, faces, *args)          # the resident V
    assert bitwise_equal(c, b)
    print("loads", a.tolist(), flush=True)


def test_fem_routines_refuse_what_they_do_not_cover():
    m = M.q2(M.quad_2d(3, 3))
    g = ParticleStage(m, 0.01, 2, 3, [2, 2], FP64)
    with pytest.raises(Pf2Error, match="Q1"):
        g.reinterpolate_fields(np.zeros((m.n_vertices, 2)))
    with pytest.raises(Pf2Error, match="Q1"):
        g.calculate_loads(np.zeros(m.n_dofs), np.zeros((0, 2)), 1.0, 1.0, 1.0, 1.0)
    g.close()
    m = M.quad_2d(3, 3)
    g = ParticleStage(m, 0.01, 2, 3, [2, 2], FP64)
    with pytest.raises(Pf2Error, match="out of range"):
        g.calculate_loads(np.zeros(m.n_dofs), np.array([[0, 4]]), 1.0, 1.0, 1.0, 1.0)
    g.close()
