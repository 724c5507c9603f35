"""Degree-2 velocity fields on the FP32-mixed path against the FP64 oracle,
per field (definitions in test_gpu_fp32_mixed.field_errors): the fused and
the separate projection sums (2-D and 3-D), the 3-D partials with two
accumulators per lane, the moving-mesh re-sort through the FP64 regroup, and
the closed form (a quadratic field is reproduced at every seeded particle to
FP32 rounding).

2-D runs ten steps with the FP32-mixed path's bounds (1e-5 per field).  In
3-D the projection is ill-conditioned with a Q2 basis: the vertex basis
functions are negative over most of the cell, so a node's weight sum
sum_p N_k(xi_p) cancels to a small fraction of its terms (down to 1e-3 of
sum_p |N_k| in the box case, 1.3e-4 in the pipe), and the FP32 rounding of
xi and N_k (~6e-8) is amplified by that ratio: FP32 terms summed exactly
already put the pipe's projected field 6.9e-5 from the FP64 one.  Through
the corrections V - V_old the difference then grows step by step.  The 3-D
cases therefore check one step: the particle fields to 1e-5 as in 2-D, the
projected field to 5e-5 (box) and 1e-3 (pipe)."""
import numpy as np
import pytest

from oracle_binding import Oracle
from pfem2_amd import FP32_MIXED, ParticleStage
from pfem2_amd import mesh as M
from test_gpu_fp32_mixed import field_errors

pytestmark = pytest.mark.gpu


def _case(name):
    if name == "channel":       # q 4x4, MAX 3: bound 75 > 64, separate partials
        m = M.q2(M.channel_2d(60, 10, 3.0, 1.0))
        V = M.poiseuille_2d(m, 1.0, 1.0, 0.02, seed=60)
        return m, V, 0.01, 10, 3, [4, 4], 10, 1e-5
    if name == "cylinder":      # q 3x3, MAX 3: bound 48, sums fused into the write
        m = M.q2(M.cylinder_channel(0.12, seed=61))
        V = M.cylinder_flow(m, 1.0, eps=0.01, seed=62)
        return m, V, 0.002, 10, 3, [3, 3], 10, 1e-5
    if name == "box3d":         # q 2x2x2, MAX 2: bound 54, fused, 7 sums per lane
        m = M.q2(M.box_3d(6, 5, 5, (0, 0, 0), (1.0, 0.8, 0.8), jitter=0.15, seed=63))
        V = M.uniform_field(m, [0.5, 0.2, -0.1]) + 0.3 * M.fourier_perturbation(m.dof_points, [1.0, 0.8, 0.8], seed=64)
        return m, V, 0.02, 5, 2, [2, 2, 2], 1, 5e-5
    # pipe, q 2x2x2, MAX 3: bound 81 > 64, separate partials with 2 sums per lane
    m = M.q2(M.pipe_3d(10, 10, 14, 0.5, 1.0))
    V = M.pipe_flow(m, 1.0, 0.5, eps=0.05, seed=65)
    return m, V, 0.01, 10, 3, [2, 2, 2], 1, 1e-3


@pytest.mark.parametrize("name", ["channel", "cylinder", "box3d", "pipe3d"])
def test_fp32_q2_error_vs_fp64_oracle(name):
    m, V, dt, S, MAX, q, steps, v_bound = _case(name)
    g = ParticleStage(m, dt, S, MAX, q, FP32_MIXED)
    o = Oracle(m, dt, S, MAX, q)
    cell, xi = M.jittered_lattice(m, q, 0.05, 66)
    for s in (g, o):
        s.upload_fields(V, V)
        s.insert_particles(cell, xi)
    n0 = len(cell)
    for _ in range(steps):
        Vg = g.step(download=True)
        Vo = o.step()
        g.upload_fields(V, None)
        o.upload_fields(V, None)
    err = field_errors(g, o, m, Vg, Vo, n0)
    print("q2 fp32", name, err, flush=True)
    assert err["V_rel_l2"] <= v_bound
    assert err["x_rel"] <= 1e-5
    assert err["v_p999"] <= 1e-5
    assert err["v_rel"] <= 1e-3
    assert err["owner_mismatch"] <= 1e-3
    assert err["count_mismatch"] <= 1e-2


def test_fp32_q2_seeded_particles_take_the_quadratic_field():
    m = M.q2(M.quad_2d(8, 6, (0.0, 0.0), (2.0, 1.5)))
    P = m.dof_points
    V = np.stack([1.0 + 0.3 * P[:, 0] ** 2 - 0.5 * P[:, 0] * P[:, 1], 0.2 * P[:, 1] ** 2 - P[:, 0]])
    g = ParticleStage(m, 0.01, 2, 3, [3, 3], FP32_MIXED)
    g.upload_fields(V, V)
    g.seed_particles()
    p = g.particles()
    x = p["x"]
    exact = np.stack([1.0 + 0.3 * x[:, 0] ** 2 - 0.5 * x[:, 0] * x[:, 1], 0.2 * x[:, 1] ** 2 - x[:, 0]], axis=1)
    assert np.max(np.abs(p["v"] - exact)) < 1e-6 * np.max(np.abs(exact))


def test_fp32_q2_correction_and_advection_follow_the_quadratic_field():
    """One step in a steady quadratic field with V_old = 0: the correction
    adds V at each particle (the FP32 monomial form of the Q2 increment), and
    the advection moves particles by dt * u(x) to FP32 accuracy."""
    m = M.q2(M.quad_2d(10, 10, (0.0, 0.0), (1.0, 1.0)))
    P = m.dof_points
    V = np.stack([0.2 + 0.1 * P[:, 1] ** 2, 0.05 * P[:, 0] * (1.0 - P[:, 0])])
    g = ParticleStage(m, 0.01, 4, 3, [3, 3], FP32_MIXED)
    g.upload_fields(np.zeros_like(V), np.zeros_like(V))
    g.seed_particles()
    p0 = g.particles()
    g.upload_fields(V, None)
    g.correct_particles_velocities()
    p1 = g.particles()
    x = p0["x"]
    u = np.stack([0.2 + 0.1 * x[:, 1] ** 2, 0.05 * x[:, 0] * (1.0 - x[:, 0])], axis=1)
    assert np.max(np.abs(p1["v"] - u)) < 1e-6
    g.move_particles()
    p2 = g.particles()
    o = np.argsort(p2["id"])
    i0 = np.argsort(p0["id"])
    moved = p2["x"][o] - p0["x"][i0]
    # forward Euler over 4 substeps in a field that varies slowly: within 1e-5 of dt * u(x0)
    assert np.max(np.abs(moved - 0.01 * u[i0])) < 1e-5


def test_fp32_q2_moving_mesh_through_the_fp64_regroup():
    """Mesh move by the three separate calls on an FP32-mixed Q2 handle: the
    re-sort relocates in FP64 and the census regroups the FP64 copy,
    interpolating new particles from V through the DoF map (9 DoFs)."""
    m = M.q2(M.cylinder_channel(0.1, seed=67))
    V = M.cylinder_flow(m, 1.0, eps=0.01, seed=68)
    g = ParticleStage(m, 0.002, 8, 2, [3, 3], FP32_MIXED, keep_velocity_ext=True)
    o = Oracle(m, 0.002, 8, 2, [3, 3])
    cell, xi = M.jittered_lattice(m, [3, 3], 0.05, 69)
    for st in (g, o):
        st.upload_fields(V, V)
        st.insert_particles(cell, xi)
    n0 = len(cell)
    X0 = m.vertices.copy()
    d = np.hypot(X0[:, 0] - 0.2, X0[:, 1] - 0.2)
    w = np.clip(1.0 - (d - 0.05) / 0.15, 0.0, 1.0)
    for s in range(4):
        g.step()
        o.step()
        X = X0.copy()
        X[:, 1] += 0.003 * np.sin(0.5 * (s + 1)) * w
        g.update_vertices(X)
        g.sort_particles_into_subdomains_and_cells()
        g.check_empty_cells()
        o.update_vertices(X)
        o.sort_particles_into_subdomains_and_cells()
        o.check_empty_cells()
        g.upload_fields(V, None)
        o.upload_fields(V, None)
    Vg = g.step(download=True)
    Vo = o.step()
    err = field_errors(g, o, m, Vg, Vo, n0)
    print("q2 moving fp32", err, flush=True)
    assert err["V_rel_l2"] <= 1e-5
    assert err["x_rel"] <= 1e-5
    assert err["owner_mismatch"] <= 1e-3
    assert err["count_mismatch"] <= 1e-2
