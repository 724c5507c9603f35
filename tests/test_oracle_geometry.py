"""Pins of the oracle's restatement of the deal.II Q1 primitives.

The reference has no tests and deal.II is absent, so these check the
mathematics: shape functions (partition of unity, Kronecker property),
the forward map (vertices, affine exactness), the inverse map (round trips
on distorted cells, closed form == Newton, rectangles in closed form), and
the clearly-outside shortcut (never changes a decision).
"""
import numpy as np
import pytest

import oracle_binding as ob

RNG = np.random.default_rng(1234)
UNIT2 = np.array([[0, 0], [1, 0], [0, 1], [1, 1]], dtype=float)
UNIT3 = np.array([[(k >> 0) & 1, (k >> 1) & 1, (k >> 2) & 1] for k in range(8)], dtype=float)


def random_quad(rng, jitter=0.25, scale=1.0, offset=(0.0, 0.0)):
    P = UNIT2 + rng.uniform(-jitter, jitter, size=(4, 2))
    return P * scale + np.array(offset)


def random_hex(rng, jitter=0.15):
    return UNIT3 + rng.uniform(-jitter, jitter, size=(8, 3))


def convex(P):
    cyc = [0, 1, 3, 2]
    for i in range(4):
        a, b, c = cyc[i], cyc[(i + 1) % 4], cyc[(i + 2) % 4]
        e0, e1 = P[b] - P[a], P[c] - P[b]
        if e0[0] * e1[1] - e0[1] * e1[0] <= 0:
            return False
    return True


@pytest.mark.parametrize("dim", [2, 3])
def test_shape_partition_of_unity_and_kronecker(dim):
    for _ in range(200):
        xi = RNG.uniform(-0.5, 1.5, size=dim)
        N = ob.shape_values(dim, xi)
        assert abs(N.sum() - 1.0) < 1e-14
    verts = UNIT2 if dim == 2 else UNIT3
    for b, vb in enumerate(verts):
        N = ob.shape_values(dim, vb)
        expect = np.zeros(len(verts))
        expect[b] = 1.0
        assert np.array_equal(N, expect)


def test_shape_values_closed_form_2d():
    for _ in range(100):
        x, y = RNG.uniform(0, 1, 2)
        N = ob.shape_values(2, np.array([x, y]))
        ref = [(1 - x) * (1 - y), x * (1 - y), (1 - x) * y, x * y]
        assert np.array_equal(N, ref)


@pytest.mark.parametrize("dim", [2, 3])
def test_forward_map_vertices_and_affine(dim):
    verts = UNIT2 if dim == 2 else UNIT3
    P = (random_quad(RNG) if dim == 2 else random_hex(RNG))
    for k, vk in enumerate(verts):
        assert np.allclose(ob.forward_map(P, vk), P[k], atol=1e-15, rtol=0)
    # parallelogram / parallelepiped: F(xi) = x0 + A xi exactly (to rounding)
    A = np.eye(dim) + RNG.uniform(-0.3, 0.3, size=(dim, dim))
    x0 = RNG.uniform(-2, 2, size=dim)
    Pa = np.array([x0 + A @ v for v in verts])
    for _ in range(50):
        xi = RNG.uniform(0, 1, size=dim)
        assert np.allclose(ob.forward_map(Pa, xi), x0 + A @ xi, atol=1e-14)


def test_inverse_round_trip_2d_distorted():
    n_ok = 0
    for _ in range(400):
        P = random_quad(RNG, 0.3, scale=RNG.uniform(1e-3, 10), offset=RNG.uniform(-5, 5, 2))
        if not convex(P):
            continue
        xi = RNG.uniform(0, 1, size=2)
        x = ob.forward_map(P, xi)
        ok, r = ob.inverse_map(P, x)
        assert ok
        assert np.max(np.abs(r - xi)) < 1e-10
        n_ok += 1
    assert n_ok > 200


def test_inverse_round_trip_3d_distorted():
    for _ in range(200):
        P = random_hex(RNG, 0.15) * RNG.uniform(1e-2, 5) + RNG.uniform(-3, 3, 3)
        xi = RNG.uniform(0, 1, size=3)
        x = ob.forward_map(P, xi)
        ok, r = ob.inverse_map(P, x)
        assert ok
        assert np.max(np.abs(r - xi)) < 1e-9


def test_closed_form_equals_newton_inside():
    for _ in range(300):
        P = random_quad(RNG, 0.3)
        if not convex(P):
            continue
        xi = RNG.uniform(0.01, 0.99, size=2)
        x = ob.forward_map(P, xi)
        ok1, r1 = ob.inverse_closed_form(P, x)
        ok2, r2 = ob.inverse_newton(P, x)
        assert ok1 and ok2
        assert np.max(np.abs(r1 - r2)) < 1e-10
        assert np.max(np.abs(r1 - xi)) < 1e-10


def test_inverse_rectangle_matches_division():
    for _ in range(200):
        x0, y0 = RNG.uniform(-3, 3, 2)
        hx, hy = RNG.uniform(0.01, 2, 2)
        P = np.array([[x0, y0], [x0 + hx, y0], [x0, y0 + hy], [x0 + hx, y0 + hy]])
        p = np.array([x0 + RNG.uniform(0, hx), y0 + RNG.uniform(0, hy)])
        ok, r = ob.inverse_map(P, p)
        assert ok
        ref = np.array([(p[0] - x0) / hx, (p[1] - y0) / hy])
        assert np.max(np.abs(r - ref)) < 4e-15 * max(1.0, abs(x0) / hx, abs(y0) / hy)


def test_points_outside_are_not_inside():
    """Outside points never map into the unit cell (closed form, Newton, filter)."""
    for _ in range(300):
        P = random_quad(RNG, 0.25)
        if not convex(P):
            continue
        xi = RNG.uniform(-1.0, 2.0, size=2)
        if np.all((xi >= -0.02) & (xi <= 1.02)):
            continue
        x = ob.forward_map(P, xi)
        for flt in (True, False):
            ok, r = ob.inverse_map(P, x, use_filter=flt)
            inside = ok and np.all((r >= 0) & (r <= 1))
            assert not inside


def test_filter_never_changes_the_decision_2d():
    """The clearly-outside shortcut agrees with the full inverse map,
    including points within 1e-6 of the cell boundary."""
    n_filtered = 0
    for _ in range(2000):
        P = random_quad(RNG, 0.3)
        if not convex(P):
            continue
        if RNG.uniform() < 0.5:
            xi = RNG.uniform(-0.6, 1.6, size=2)
        else:  # near an edge
            xi = RNG.uniform(0, 1, size=2)
            d = RNG.integers(2)
            xi[d] = RNG.choice([0.0, 1.0]) + RNG.uniform(-1e-6, 1e-6)
        x = ob.forward_map(P, xi)
        ok_f, r_f = ob.inverse_map(P, x, use_filter=True)
        ok_n, r_n = ob.inverse_map(P, x, use_filter=False)
        in_f = ok_f and np.all((r_f >= 0) & (r_f <= 1))
        in_n = ok_n and np.all((r_n >= 0) & (r_n <= 1))
        assert in_f == in_n
        if in_f:
            assert np.array_equal(r_f, r_n)
        if ob.clearly_outside(P, x):
            n_filtered += 1
            assert not in_n
    assert n_filtered > 100


def test_filter_never_changes_the_decision_3d():
    for _ in range(500):
        P = random_hex(RNG, 0.12)
        xi = RNG.uniform(-0.5, 1.5, size=3)
        x = ob.forward_map(P, xi)
        ok_f, r_f = ob.inverse_map(P, x, use_filter=True)
        ok_n, r_n = ob.inverse_map(P, x, use_filter=False)
        in_f = ok_f and np.all((r_f >= 0) & (r_f <= 1))
        in_n = ok_n and np.all((r_n >= 0) & (r_n <= 1))
        assert in_f == in_n


def test_degenerate_cell_fails_cleanly():
    P = np.array([[0, 0], [1, 0], [0, 0], [1, 0]], dtype=float)   # zero area
    ok, r = ob.inverse_map(P, np.array([0.5, 0.0]))
    inside = ok and np.all((r >= 0) & (r <= 1))
    # either the mapping fails or the answer is a point of the degenerate cell; no crash
    assert isinstance(inside, (bool, np.bool_))
