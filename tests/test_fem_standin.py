"""Host FEM stand-in of the channel driver (pfem2_amd/fem.py), CPU only.

Pinned by what the mathematics fixes: the operators reproduce exact
integrals of polynomials; the scheme's steady state is the Poiseuille
profile with its linear pressure away from the inlet and outlet; the
correction reduces the discrete divergence; the Dirichlet values hold; the
step is deterministic."""
import numpy as np

from pfem2_amd import fem
from pfem2_amd import mesh as M


def test_operators_integrate_polynomials():
    m = M.channel_2d(8, 4, 2.0, 1.0)
    Mm, K, Gx, Gy = fem.assemble(m)
    one = np.ones(m.n_dofs)
    x, y = m.vertices[:, 0], m.vertices[:, 1]
    assert np.isclose(one @ Mm @ one, 2.0)                       # area
    assert np.allclose(K @ one, 0.0)                              # constants are harmonic
    assert np.isclose(one @ Gx @ x, 2.0) and np.isclose(one @ Gy @ y, 2.0)   # int d x/dx = area
    assert np.isclose(one @ K @ x, 0.0) and np.isclose(x @ K @ x, 2.0)       # int |grad x|^2 = area


def test_steady_state_is_poiseuille_away_from_the_ends():
    """The fractional step (non-incremental pressure, as in the reference's
    cylinder2D driver) settles to a steady state; mid-channel it is the
    Poiseuille profile with the pressure gradient -8 mu umax / H^2."""
    m = M.channel_2d(60, 10, 5.0, 1.0)
    f = fem.ChannelFEM(m, dt=0.01, umax=1.0, mu=0.01)
    V0 = M.poiseuille_2d(m, 1.0, 1.0, 0.0)
    V = V0.copy()
    for _ in range(300):
        Vn = f.step(V)
        change = np.max(np.abs(Vn - V))
        V = Vn
    assert change < 1e-3
    mid = np.abs(m.vertices[:, 0] - 2.5) < 0.05
    assert np.max(np.abs(V - V0)[:, mid]) < 2e-3
    x = m.vertices[:, 0]
    inner = (x > 1.0) & (x < 4.0)
    slope = np.polyfit(x[inner], f.p[inner], 1)[0]
    assert abs(slope + 8 * 0.01) < 0.1 * 8 * 0.01


def test_correction_reduces_divergence_and_keeps_boundary_values():
    m = M.channel_2d(60, 10, 5.0, 1.0)
    f = fem.ChannelFEM(m, dt=0.01, umax=1.0, mu=0.01)
    V = M.poiseuille_2d(m, 1.0, 1.0, 0.1, seed=5)
    Vn = f.step(V)
    inner = np.setdiff1d(np.arange(m.n_dofs), np.concatenate([f.walls, f.inlet, f.outlet]))
    assert np.linalg.norm(f.divergence(Vn)[inner]) < 0.8 * np.linalg.norm(f.divergence(V)[inner])
    assert np.all(Vn[:, f.walls] == 0.0)
    assert np.array_equal(Vn[0, f.inlet], f.u_in) and np.all(Vn[1, f.inlet] == 0.0)


def test_step_is_deterministic():
    m = M.channel_2d(30, 6, 3.0, 1.0)
    V = M.poiseuille_2d(m, 1.0, 1.0, 0.05, seed=6)
    a = fem.ChannelFEM(m, 0.01).step(V)
    b = fem.ChannelFEM(m, 0.01).step(V)
    assert np.array_equal(a, b)
