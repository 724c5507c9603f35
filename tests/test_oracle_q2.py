"""Degree-2 velocity fields in the oracle, pinned by what the mathematics
fixes (the reference's fields are FE_Q(1) only, so nothing of the reference
pins them):
  * the basis equals the Lagrange basis obtained by inverting the monomial
    Vandermonde matrix of Q2 at the support points (an independent
    construction), and N_k(node_j) = delta_kj at the support points that the
    mesh's DoF numbering (deal.II line/face order) assigns to local DoF j;
  * interpolation reproduces quadratics exactly: the velocity a particle
    receives equals the closed-form field at its position;
  * a Q2 Poiseuille channel: particles keep y and advance by dt * u(y) per
    step (exact with Q2), and the projection of a constant velocity is that
    constant.
"""
import itertools

import numpy as np
import pytest

from oracle_binding import Oracle, field_shape_values
from pfem2_amd import mesh as M


def _support_unit(dim):
    """Unit-cell coordinates of local DoF k from the mesh's numbering on one
    unit cell (identity map)."""
    if dim == 2:
        m = M.quad_2d(1, 1, (0.0, 0.0), (1.0, 1.0))
    else:
        m = M.box_3d(1, 1, 1)
    mq = M.q2(m)
    return mq.dof_xyz[mq.cell_dofs[0]]


def _vandermonde_basis(nodes, pts):
    """Lagrange basis of span{x^a y^b [z^c]: a, b, c <= 2} at pts."""
    dim = nodes.shape[1]
    exps = list(itertools.product(range(3), repeat=dim))

    def mono(P):
        return np.stack([np.prod([P[:, d] ** e[d] for d in range(dim)], axis=0) for e in exps], axis=1)

    A = mono(nodes)                   # (nd, nd): row j = monomials at node j
    C = np.linalg.solve(A, np.eye(len(exps)))   # columns: coefficients of basis k
    return mono(pts) @ C              # (npts, nd)


@pytest.mark.parametrize("dim", [2, 3])
def test_q2_basis_equals_vandermonde_lagrange_basis(dim):
    nodes = _support_unit(dim)
    rng = np.random.default_rng(dim)
    pts = rng.uniform(0.0, 1.0, size=(50, dim))
    ref = _vandermonde_basis(nodes, pts)
    got = np.array([field_shape_values(dim, 2, p) for p in pts])
    assert np.max(np.abs(got - ref)) < 1e-12
    assert np.max(np.abs(got.sum(axis=1) - 1.0)) < 1e-14


@pytest.mark.parametrize("dim", [2, 3])
def test_q2_kronecker_at_mesh_support_points(dim):
    nodes = _support_unit(dim)
    N = np.array([field_shape_values(dim, 2, p) for p in nodes])
    assert np.array_equal(N, np.eye(len(nodes)))


def test_q1_field_values_are_the_vertex_shape_values():
    rng = np.random.default_rng(5)
    for dim in (2, 3):
        for _ in range(10):
            xi = rng.uniform(0.0, 1.0, size=dim)
            a = field_shape_values(dim, 1, xi)
            b = np.array([np.prod([xi[d] if (k >> d) & 1 else 1.0 - xi[d] for d in range(dim)])
                          for k in range(1 << dim)])
            assert np.max(np.abs(a - b)) < 1e-15


def _quadratic_2d(P):
    x, y = P[:, 0], P[:, 1]
    return np.stack([1.0 + 0.5 * x - 0.3 * y + 0.2 * x * x - 0.7 * x * y + 0.4 * y * y,
                     -0.2 + 0.1 * x + 0.9 * y - 0.5 * x * x + 0.3 * x * y + 0.6 * y * y])


def _quadratic_3d(P):
    x, y, z = P[:, 0], P[:, 1], P[:, 2]
    return np.stack([0.3 + x * y - 0.5 * z * z + 0.2 * x,
                     -0.1 + 0.4 * y * z + 0.3 * x * x - y,
                     1.0 - 0.2 * x * z + 0.5 * y * y + 0.1 * z])


@pytest.mark.parametrize("dim", [2, 3])
def test_q2_interpolation_reproduces_quadratics(dim):
    """On parallelogram / parallelepiped cells (affine maps) a quadratic
    field is in the Q2 space: the velocity interpolated at a particle equals
    the field at the particle's location."""
    if dim == 2:
        m = M.q2(M.quad_2d(5, 4, (0.0, 0.0), (2.0, 1.5)))
        f = _quadratic_2d
    else:
        m = M.q2(M.box_3d(3, 3, 2, (0, 0, 0), (1.5, 1.2, 0.8)))
        f = _quadratic_3d
    V = f(m.dof_points)
    o = Oracle(m, 0.01, 1, 3, [2] * dim)
    o.upload_fields(V, V)
    cell, xi = M.jittered_lattice(m, [3] * dim, 0.15, 9)
    o.insert_particles(cell, xi)
    p = o.particles()
    exact = f(p["x"]).T
    assert np.max(np.abs(p["v"] - exact)) < 1e-13


def test_q2_poiseuille_advection_is_exact():
    """Pure Poiseuille flow u = 4 y (1 - y) on a structured channel with a Q2
    field: every particle keeps y and moves by dt * u(y) in x per step (the
    field is reproduced exactly, so only rounding remains)."""
    m = M.q2(M.channel_2d(40, 8, 4.0, 1.0))
    V = M.poiseuille_2d(m, 1.0, 1.0, eps=0.0)
    dt, S = 0.01, 10
    o = Oracle(m, dt, S, 3, [3, 3])
    o.upload_fields(V, V)
    cell, xi = M.jittered_lattice(m, [3, 3], 0.05, 3)
    o.insert_particles(cell, xi)
    p0 = o.particles()
    o.correct_particles_velocities()
    o.move_particles()
    p1 = o.particles()
    pos0 = {int(i): x for i, x in zip(p0["id"], p0["x"])}
    n = 0
    for i, x in zip(p1["id"], p1["x"]):
        if int(i) in pos0:
            x0 = pos0[int(i)]
            u = 4.0 * x0[1] * (1.0 - x0[1])
            assert abs(x[1] - x0[1]) < 1e-14
            assert abs(x[0] - (x0[0] + dt * u)) < 1e-13
            n += 1
    assert n > 0.9 * len(p0["id"])


def test_q2_projection_of_a_constant_is_that_constant():
    m = M.q2(M.quad_2d(6, 5, (0.0, 0.0), (1.2, 1.0), jitter=0.2, seed=4))
    V = M.uniform_field(m, [0.7, -0.3])
    o = Oracle(m, 0.01, 2, 3, [3, 3])
    o.upload_fields(V, V)
    cell, xi = M.jittered_lattice(m, [3, 3], 0.1, 6)
    o.insert_particles(cell, xi)
    Vp = o.distribute_particle_velocities_to_grid()
    assert np.max(np.abs(Vp - V)) < 1e-12
