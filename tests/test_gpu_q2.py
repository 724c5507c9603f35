"""Degree-2 velocity fields (FE_Q<dim>(2) on the Q1 geometry) on the FP64
path against the oracle: bitwise equal cells, ids, counts, x, xi, v and
projected fields, in 2-D and 3-D, through the fused step, the separate
calls and the moving-mesh re-sort; plus the closed form (a quadratic field
is reproduced exactly at every particle)."""
import numpy as np
import pytest

from oracle_binding import Oracle
from parity_util import assert_particles_match, bitwise_equal
from pfem2_amd import FP64, Pf2Error, ParticleStage
from pfem2_amd import mesh as M

pytestmark = pytest.mark.gpu


def _pair(m, V, dt, S, MAX, q, seed, keep_vext=False):
    g = ParticleStage(m, dt, S, MAX, q, FP64, keep_velocity_ext=keep_vext)
    o = Oracle(m, dt, S, MAX, q)
    cell, xi = M.jittered_lattice(m, q, 0.05, seed)
    for s in (g, o):
        s.upload_fields(V, V)
        s.insert_particles(cell, xi)
    return g, o


def _steps(g, o, steps, where, V_next=None):
    for s in range(steps):
        Vg = g.step(download=True)
        Vo = o.step()
        assert_particles_match(g.particles(), o.particles(), where=f"{where} step {s}", bitwise=True)
        assert np.array_equal(g.cell_counts(), o.cell_counts()), f"{where} step {s}"
        assert g.particle_counter() == o.particle_counter()
        assert bitwise_equal(Vg, Vo), f"{where} step {s}: projected field"
        if V_next is not None:
            g.upload_fields(V_next(s), None)
            o.upload_fields(V_next(s), None)
    print(where, "bitwise equal,", g.n_global_particles(), "particles", flush=True)


def test_q2_channel_2d():
    m = M.q2(M.channel_2d(60, 10, 3.0, 1.0))
    V = M.poiseuille_2d(m, 1.0, 1.0, 0.02, seed=40)
    V2 = M.poiseuille_2d(m, 1.0, 1.0, 0.03, seed=41)
    g, o = _pair(m, V, 0.01, 10, 3, [4, 4], seed=42)
    _steps(g, o, 5, "q2 channel", lambda s: V2 if s % 2 == 0 else V)


def test_q2_cylinder_2d():
    m = M.q2(M.cylinder_channel(0.12, seed=43))
    V = M.cylinder_flow(m, 1.0, eps=0.02, seed=44)
    g, o = _pair(m, V, 0.004, 10, 3, [4, 4], seed=45)
    _steps(g, o, 4, "q2 cylinder", lambda s: V)


def test_q2_box_and_pipe_3d():
    m = M.q2(M.box_3d(5, 4, 4, (0, 0, 0), (1.0, 0.8, 0.8), jitter=0.15, seed=46))
    V = M.uniform_field(m, [0.5, 0.2, -0.1]) + 0.3 * M.fourier_perturbation(m.dof_points, [1.0, 0.8, 0.8], seed=47)
    g, o = _pair(m, V, 0.05, 4, 2, [2, 2, 2], seed=48)
    _steps(g, o, 3, "q2 box", lambda s: V)
    m = M.q2(M.pipe_3d(8, 8, 10, 0.5, 1.0))
    V = M.pipe_flow(m, 1.0, 0.5, eps=0.05, seed=49)
    g, o = _pair(m, V, 0.02, 5, 2, [2, 2, 2], seed=50)
    _steps(g, o, 2, "q2 pipe", lambda s: V)


def test_q2_seeded_particles_take_the_quadratic_field_exactly():
    m = M.q2(M.quad_2d(8, 6, (0.0, 0.0), (2.0, 1.5)))
    P = m.dof_points
    V = np.stack([1.0 + 0.3 * P[:, 0] ** 2 - 0.5 * P[:, 0] * P[:, 1], 0.2 * P[:, 1] ** 2 - P[:, 0]])
    g = ParticleStage(m, 0.01, 2, 3, [3, 3], FP64)
    o = Oracle(m, 0.01, 2, 3, [3, 3])
    for s in (g, o):
        s.upload_fields(V, V)
        s.seed_particles()
    pg, po = g.particles(), o.particles()
    assert_particles_match(pg, po, where="q2 seed", bitwise=True)
    x = pg["x"]
    exact = np.stack([1.0 + 0.3 * x[:, 0] ** 2 - 0.5 * x[:, 0] * x[:, 1], 0.2 * x[:, 1] ** 2 - x[:, 0]], axis=1)
    assert np.max(np.abs(pg["v"] - exact)) < 1e-13


@pytest.mark.parametrize("kind", ["2d", "3d"])
def test_q2_separate_calls_equal_fused_step(kind):
    if kind == "2d":
        m = M.q2(M.cylinder_channel(0.08, seed=51))
        V = M.cylinder_flow(m, 1.0, eps=0.02, seed=52)
        q, dt, MAX = [3, 3], 0.004, 3
    else:
        m = M.q2(M.pipe_3d(8, 8, 10, 0.5, 1.0))
        V = M.pipe_flow(m, 1.0, 0.5, eps=0.05, seed=53)
        q, dt, MAX = [2, 2, 2], 0.02, 2
    cell, xi = M.jittered_lattice(m, q, 0.05, 54)
    g1 = ParticleStage(m, dt, 5, MAX, q, FP64)
    g2 = ParticleStage(m, dt, 5, MAX, q, FP64)
    for g in (g1, g2):
        g.upload_fields(V, V)
        g.insert_particles(cell, xi)
    for _ in range(2):
        a = g1.step(download=True)
        g2.correct_particles_velocities()
        g2.move_particles()
        b = g2.distribute_particle_velocities_to_grid()
        assert bitwise_equal(a, b)
        pa, pb = g1.particles(), g2.particles()
        for k in ("cell", "id", "xi", "v"):
            assert np.array_equal(pa[k], pb[k])
        g1.upload_fields(V, None)
        g2.upload_fields(V, None)


def test_q2_moving_mesh():
    m = M.q2(M.cylinder_channel(0.08, seed=55))
    V = M.cylinder_flow(m, 1.0, eps=0.02, seed=56)
    g, o = _pair(m, V, 0.004, 5, 3, [3, 3], seed=57, keep_vext=True)
    X0 = m.vertices.copy()
    d = np.hypot(X0[:, 0] - 0.2, X0[:, 1] - 0.2)
    w = np.clip(1.0 - (d - 0.05) / 0.15, 0.0, 1.0)
    for s in range(2):
        g.step()
        o.step()
        X = X0.copy()
        X[:, 1] += 0.004 * np.sin(0.7 * (s + 1)) * w
        g.move_mesh(X)
        o.update_vertices(X)
        o.sort_particles_into_subdomains_and_cells()
        o.check_empty_cells()
        assert_particles_match(g.particles(), o.particles(), where=f"q2 moving {s}", bitwise=True)
        assert np.array_equal(g.cell_counts(), o.cell_counts())


def test_q2_limits():
    """Degree-2 handles take no walls: the wall fix-up is defined on Q1
    vertex DoFs."""
    m = M.q2(M.quad_2d(3, 3))
    g = ParticleStage(m, 0.01, 2, 3, [2, 2], FP64)
    with pytest.raises(Pf2Error, match="Q1"):
        g.set_walls(np.array([0]), np.array([0]), np.zeros((1, 2)), np.array([3]))
    g.close()
