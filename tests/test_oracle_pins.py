"""Pins of the oracle where the first pins left it open: the boundary-aware
wall fix-up of cylinderDyM (cylinderDyM/pfem2particle.cpp:960-992) and the
3-D seed, reseed and relocation (tubePoiseuille3D/pfem2particle.cpp:489-567).

Every expectation here comes from a closed form or a brute-force search that
does not restate the oracle's code path: uniform fields make the projection
reproduce the particle velocity, so the fix-up's outcome reduces to a sum of
shape functions known in closed form on a structured grid; seeding and
reseeding on unit cubes have exact lattice positions; relocation is checked
against exhaustive point location.
"""
import itertools

import numpy as np
import pytest

import oracle_binding as ob
from oracle_binding import Oracle
from pfem2_amd import mesh as M


# ---------------------------------------------------------------------------
# wall fix-up with boundary values and component masks

def _uniform_run(nx, ny, u, q=(2, 2), walls=None):
    """Particles on a lattice with the uniform velocity u; one projection
    with the given walls (vertex list, bc (n x 2), mask (n)).  Returns the
    oracle, the particles before and after, and the projected field."""
    m = M.channel_2d(nx, ny, float(nx), float(ny))        # unit squares
    o = Oracle(m, 0.1, 1, 3, list(q))
    V = M.uniform_field(m, u)
    o.upload_fields(V, V)
    o.seed_particles()
    p0 = o.particles()
    # particle velocities exactly u (seeding interpolates, which may round)
    o2 = Oracle(m, 0.1, 1, 3, list(q))
    o2.upload_fields(V, V)
    o2.insert_particles(p0["cell"], p0["xi"], np.tile(np.asarray(u, dtype=float), (len(p0["id"]), 1)))
    before = o2.particles()
    if walls is not None:
        vert, bc, mask = walls
        o2.set_walls(vert, vert, bc, mask)
    Vp = o2.distribute_particle_velocities_to_grid()
    return m, o2, before, o2.particles(), Vp


def _bottom_eta(m, p):
    """1 - eta: the sum N_0 + N_1 of the two bottom vertices of a cell of
    unit squares (N_0 = (1-xi)(1-eta), N_1 = xi (1-eta))."""
    return 1.0 - p["xi"][:, 1]


def test_wall_bc_equal_to_flow_is_a_noop():
    """Projection of a uniform particle velocity u reproduces u at every
    node (to rounding); a wall whose boundary value is u then deviates by
    less than the 1e-5 trigger: no particle changes, the wall DoFs take bc."""
    u = np.array([0.7, 0.01])
    m0 = M.channel_2d(6, 3, 6.0, 3.0)
    wall = np.asarray(m0.boundary_vertices["bottom"], dtype=np.int32)
    body = np.asarray(m0.boundary_vertices["top"], dtype=np.int32)
    vert = np.concatenate([wall, body])
    bc = np.concatenate([np.tile(u, (len(wall), 1)), np.tile([123.0, u[1]], (len(body), 1))])
    mask = np.concatenate([np.full(len(wall), 3), np.full(len(body), 2)]).astype(np.int32)
    m, o, before, after, Vp = _uniform_run(6, 3, u, walls=(vert, bc, mask))
    assert np.array_equal(after["v"], before["v"])
    assert np.array_equal(Vp[:, wall], np.tile(u[:, None], (1, len(wall))))
    # mask 2 enforces only the y component: x keeps the projected value u_x
    assert np.array_equal(Vp[1, body], np.full(len(body), u[1]))
    assert np.allclose(Vp[0, body], u[0], rtol=0, atol=1e-15)


@pytest.mark.parametrize("mask", [3, 2, 1])
def test_wall_correction_closed_form(mask):
    """Uniform particle velocity u, bottom-wall nodes with bc != u: every
    particle in a bottom-row cell gets v = u + (bc - u)(1 - eta) in the
    enforced components (the two wall vertices' shape functions sum to
    1 - eta) and keeps its other components bit for bit; particles of other
    cells do not change at all."""
    u = np.array([0.6, -0.25])
    bcv = np.array([0.0, 0.01])                               # cylinderDyM body: (0, body_velVy)
    m0 = M.channel_2d(5, 3, 5.0, 3.0)
    wall = np.asarray(m0.boundary_vertices["bottom"], dtype=np.int32)
    m, o, before, after, Vp = _uniform_run(5, 3, u, walls=(wall, np.tile(bcv, (len(wall), 1)),
                                                            np.full(len(wall), mask, dtype=np.int32)))
    assert np.array_equal(after["id"], before["id"])
    row0 = after["cell"] < 5
    s = _bottom_eta(m, after)
    for d in range(2):
        if (mask >> d) & 1:
            expect = u[d] + (bcv[d] - u[d]) * s[row0]
            assert np.allclose(after["v"][row0, d], expect, rtol=0, atol=1e-14)
            assert np.array_equal(Vp[d, wall], np.full(len(wall), bcv[d]))
        else:
            assert np.array_equal(after["v"][row0, d], before["v"][row0, d])
    assert np.array_equal(after["v"][~row0], before["v"][~row0])


def test_wall_full_mask_cell_takes_bc():
    """A cell whose four vertices are all wall nodes with a common bc and
    full mask: the shape functions sum to one, so every particle in it ends
    with v = bc (to rounding), whatever its velocity before."""
    u = np.array([1.3, 0.4])
    bcv = np.array([-0.2, 0.05])
    m0 = M.channel_2d(3, 3, 3.0, 3.0)
    vert = np.arange(m0.n_vertices, dtype=np.int32)          # every vertex is a wall node
    m, o, before, after, Vp = _uniform_run(3, 3, u, walls=(vert, np.tile(bcv, (len(vert), 1)),
                                                            np.full(len(vert), 3, dtype=np.int32)))
    assert np.allclose(after["v"], np.tile(bcv, (len(after["id"]), 1)), rtol=0, atol=2e-15)
    assert np.array_equal(Vp, np.tile(bcv[:, None], (1, m.n_vertices)))


@pytest.mark.parametrize("delta,fires", [(0.5e-5, False), (2e-5, True)])
def test_wall_trigger_threshold(delta, fires):
    """|V - bc| > 1e-5 in an enforced component triggers the correction.
    With bc = u - delta: below the threshold nothing changes; above it the
    enforced components change by -delta (1 - eta) and the other one keeps
    its bits."""
    u = np.array([0.45, 0.3])
    m0 = M.channel_2d(4, 2, 4.0, 2.0)
    wall = np.asarray(m0.boundary_vertices["bottom"], dtype=np.int32)
    bc = np.tile([0.0, u[1] - delta], (len(wall), 1))        # x not enforced (mask 2)
    m, o, before, after, Vp = _uniform_run(4, 2, u, walls=(wall, bc, np.full(len(wall), 2, dtype=np.int32)))
    assert np.array_equal(after["v"][:, 0], before["v"][:, 0])
    row0 = after["cell"] < 4
    if fires:
        expect = u[1] - delta * _bottom_eta(m, after)[row0]
        assert np.allclose(after["v"][row0, 1], expect, rtol=0, atol=1e-15)
        assert not np.array_equal(after["v"][row0, 1], before["v"][row0, 1])
    else:
        assert np.array_equal(after["v"], before["v"])
    assert np.array_equal(Vp[1, wall], np.full(len(wall), u[1] - delta))


def test_wall_unenforced_deviation_does_not_trigger():
    """A large deviation in a component the mask does not enforce is not a
    trigger (the reference tests x only on the body, boundary id 3)."""
    u = np.array([0.45, 0.3])
    m0 = M.channel_2d(4, 2, 4.0, 2.0)
    wall = np.asarray(m0.boundary_vertices["bottom"], dtype=np.int32)
    bc = np.tile([u[0] + 1.0, u[1]], (len(wall), 1))
    m, o, before, after, Vp = _uniform_run(4, 2, u, walls=(wall, bc, np.full(len(wall), 2, dtype=np.int32)))
    assert np.array_equal(after["v"], before["v"])


def test_wall_corrections_accumulate_in_vertex_order():
    """A particle next to two wall nodes is corrected by both; with a
    uniform flow the two corrections add up to the closed form whatever
    their order, and each wall DoF is compared with its own projected value
    (set to bc only after its particles are corrected)."""
    u = np.array([0.2, 0.9])
    m0 = M.channel_2d(2, 2, 2.0, 2.0)
    # the two left vertices of the bottom-left cell (0 and 3: x = 0, y = 0, 1)
    vert = np.array([3, 0], dtype=np.int32)
    bcv = np.array([[0.0, 0.0], [0.0, 0.5]])
    m, o, before, after, Vp = _uniform_run(2, 2, u, walls=(vert, bcv, np.full(2, 3, dtype=np.int32)))
    xi = after["xi"]
    # N_0 = (1-xi)(1-eta) for vertex 0 of cell 0; vertex 3 is local 2 of cell 0
    # (N_2 = (1-xi) eta) and local 0 of cell 2 (N_0 = (1-xi)(1-eta))
    expect = before["v"].copy()
    c0 = after["cell"] == 0
    c2 = after["cell"] == 2
    N0 = (1 - xi[:, 0]) * (1 - xi[:, 1])
    N2 = (1 - xi[:, 0]) * xi[:, 1]
    for d in range(2):
        expect[c0, d] = u[d] + (bcv[1, d] - u[d]) * N0[c0] + (bcv[0, d] - u[d]) * N2[c0]
        expect[c2, d] = u[d] + (bcv[0, d] - u[d]) * N0[c2]
    assert np.allclose(after["v"], expect, rtol=0, atol=1e-15)


# ---------------------------------------------------------------------------
# 3-D: seed order, part keys with the phantom index, relocation

def _unit_box(nx, ny, nz):
    return M.box_3d(nx, ny, nz, (0.0, 0.0, 0.0), (float(nx), float(ny), float(nz)))


def _corner3(m, cell, nx, ny):
    return np.stack([cell % nx, (cell // nx) % ny, cell // (nx * ny)], axis=1).astype(float)


def test_seed_3d_order_ids_and_positions():
    """tubePoiseuille3D/pfem2particle.cpp:497-499: i outer, j, k inner;
    xi = ((i+1/2)/qx, (j+1/2)/qy, (k+1/2)/qz); ids in cell, then lattice order."""
    nx, ny, nz = 3, 2, 2
    m = _unit_box(nx, ny, nz)
    q = [2, 3, 2]
    o = Oracle(m, 0.1, 1, 3, q)
    V = M.uniform_field(m, [0.0, 0.0, 0.0])
    o.upload_fields(V, V)
    o.seed_particles()
    p = o.particles()
    per = 12
    assert len(p["id"]) == m.n_cells * per
    assert np.array_equal(p["id"], np.arange(1, m.n_cells * per + 1))
    # the reference's expression: (i + 1.0/2) * hx with hx = 1.0/quantities[0]
    lat = np.array([[(i + 0.5) * (1.0 / 2), (j + 0.5) * (1.0 / 3), (k + 0.5) * (1.0 / 2)]
                    for i in range(2) for j in range(3) for k in range(2)])
    for c in range(m.n_cells):
        assert np.array_equal(p["xi"][c * per:(c + 1) * per], lat)
    assert np.array_equal(p["cell"], np.repeat(np.arange(m.n_cells), per))
    assert np.allclose(p["x"], _corner3(m, p["cell"], nx, ny) + p["xi"], rtol=0, atol=1e-15)


def test_seed_3d_velocity_reproduces_linear_field():
    m = M.box_3d(3, 3, 2, (0, 0, 0), (1.5, 1.2, 0.8), jitter=0.1, seed=21)
    a = np.array([0.3, -0.2, 0.1])
    B = np.array([[0.5, 1.0, -0.3], [-0.7, 0.25, 0.4], [0.2, 0.0, 0.9]])
    V = a[:, None] + B @ m.vertices.T
    o = Oracle(m, 0.1, 1, 3, [2, 2, 2])
    o.upload_fields(V, V)
    o.seed_particles()
    p = o.particles()
    # trilinear interpolation reproduces affine fields exactly on any hexahedron
    assert np.allclose(p["v"], a + p["x"] @ B.T, rtol=0, atol=1e-14)


def _expected_reseed_3d(pre, n_cells, q, MAX):
    """Per-cell counts after check_cell_for_empty_parts and the (cell, i, j, k)
    of every new particle in id order, from the sub-part counts: part index
    per direction = trunc(xi_d / (1/q_d)) with the phantom value q_d."""
    counts = np.zeros(n_cells, dtype=int)
    spawns = []
    h = 1.0 / np.asarray(q, dtype=float)
    for c in range(n_cells):
        sel = pre["cell"] == c
        cnt = {}
        for x in pre["xi"][sel]:
            key = tuple(int(v) for v in (x / h))
            cnt[key] = cnt.get(key, 0) + 1
        total = sum(min(v, MAX) for v in cnt.values())
        for i, j, k in itertools.product(range(q[0]), range(q[1]), range(q[2])):
            if cnt.get((i, j, k), 0) == 0:
                total += 1
                spawns.append((c, i, j, k))
        counts[c] = total
    return counts, spawns


def test_reseed_3d_counts_order_and_phantom():
    nx, ny, nz = 2, 2, 2
    m = _unit_box(nx, ny, nz)
    q, MAX = [2, 3, 2], 2
    o = Oracle(m, 0.1, 1, MAX, q)
    V = M.uniform_field(m, [0.1, 0.2, 0.3])
    o.upload_fields(V, V)
    rng = np.random.default_rng(22)
    n = 150
    cell = rng.integers(0, m.n_cells, n).astype(np.int32)
    xi = rng.uniform(0, 1, size=(n, 3))
    # crowd one part of cell 0 and put particles exactly on the far faces
    # (xi_d = 1 gives the phantom index q_d, a part never reseeded)
    cell[:8] = 0
    xi[:8] = [0.1, 0.1, 0.1]
    xi[8:14] = np.array([[1.0, 0.5, 0.5], [0.5, 1.0, 0.5], [0.5, 0.5, 1.0],
                         [1.0, 1.0, 1.0], [1.0, 0.2, 0.9], [0.3, 1.0, 1.0]])
    o.insert_particles(cell, xi)
    pre = o.particles()
    counter0 = o.particle_counter()
    o.check_empty_cells()
    post = o.particles()
    counts, spawns = _expected_reseed_3d(pre, m.n_cells, q, MAX)
    assert np.array_equal(o.cell_counts(), counts)
    # phantom-part particles survive (their part holds at most MAX here)
    for pid in pre["id"][8:14]:
        assert pid in post["id"]
    # crowded part: the MAX smallest ids survive
    crowd = np.sort(pre["id"][(pre["cell"] == 0) & np.all(pre["xi"] == 0.1, axis=1)])
    assert all(k in post["id"] for k in crowd[:MAX])
    assert not any(k in post["id"] for k in crowd[MAX:])
    # new particles: ids continue the counter in cell, then i, j, k order,
    # at the part centres
    new = post["id"] > counter0
    order = np.argsort(post["id"][new])
    got_cell = post["cell"][new][order]
    got_xi = post["xi"][new][order]
    assert np.array_equal(np.sort(post["id"][new]), np.arange(counter0 + 1, counter0 + 1 + len(spawns)))
    assert got_cell.tolist() == [s[0] for s in spawns]
    centres = np.array([[(s[1] + 0.5) * (1.0 / q[0]), (s[2] + 0.5) * (1.0 / q[1]), (s[3] + 0.5) * (1.0 / q[2])]
                        for s in spawns])
    assert np.array_equal(got_xi, centres)
    assert o.stats()["spawned"] == len(spawns)


def _brute_cells_3d(m, x):
    out = []
    for c in range(m.n_cells):
        ok, r = ob.inverse_map(m.vertices[m.cells[c]], x, use_filter=False)
        if ok and np.all((r >= 0) & (r <= 1)):
            out.append(c)
    return out


def test_relocation_3d_matches_brute_force():
    """After a move on a jittered hexahedral box every surviving particle's
    cell contains its position (exhaustive search over all cells), and its
    reference point maps back to its position."""
    m = M.box_3d(4, 4, 3, (0, 0, 0), (1.0, 1.0, 0.75), jitter=0.2, seed=23)
    V = M.uniform_field(m, [0.2, -0.1, 0.05]) + 0.5 * M.fourier_perturbation(m.vertices, [1.0, 1.0, 0.75], seed=24)
    o = Oracle(m, 0.08, 2, 10, [2, 2, 2])
    o.upload_fields(V, V)
    rng = np.random.default_rng(25)
    n = 160
    cell = rng.integers(0, m.n_cells, n).astype(np.int32)
    xi = rng.uniform(0.05, 0.95, size=(n, 3))
    o.insert_particles(cell, xi)
    o.sort_particles_into_subdomains_and_cells()     # no motion: nothing changes
    p0 = o.particles()
    assert np.array_equal(np.sort(p0["id"]), np.arange(1, n + 1))
    o.move_particles()
    p = o.particles()
    moved = 0
    i0 = {k: j for j, k in enumerate(p0["id"])}
    for j in range(len(p["id"])):
        if p["id"][j] > n:
            continue                                   # reseeded
        cands = _brute_cells_3d(m, p["x"][j])
        assert p["cell"][j] in cands
        P = m.vertices[m.cells[p["cell"][j]]]
        # Newton stops at a residual of 1e-11 of the cell size (deal.II's tolerance)
        assert np.allclose(ob.forward_map(P, p["xi"][j]), p["x"][j], rtol=0, atol=1e-11)
        moved += p["cell"][j] != p0["cell"][i0[p["id"][j]]]
    assert moved > 10


def _single3(m, cell, xi, U, dt):
    o = Oracle(m, dt, 1, 10, [1, 1, 1])
    V = M.uniform_field(m, U)
    o.upload_fields(V, V)
    o.insert_particles(np.array([cell], dtype=np.int32), np.array([xi]))
    o.move_particles()
    return o


def test_relocation_3d_face_edge_corner_and_lost():
    nx = ny = nz = 3
    m = _unit_box(nx, ny, nz)
    centre = 1 + 3 * 1 + 9 * 1                       # cell (1, 1, 1)

    def cid(i, j, k):
        return i + 3 * j + 9 * k

    def where(o):
        p = o.particles()
        sel = p["id"] == 1
        return (int(p["cell"][sel][0]), p["xi"][sel][0]) if sel.any() else (None, None)

    # face crossing in +x
    c, r = where(_single3(m, centre, [0.9, 0.5, 0.5], [1.0, 0.0, 0.0], 0.2))
    assert c == cid(2, 1, 1) and np.allclose(r, [0.1, 0.5, 0.5], atol=1e-14)
    # edge crossing in +x +y: the star of the selected vertex holds the diagonal cell
    c, r = where(_single3(m, centre, [0.9, 0.9, 0.5], [1.0, 1.0, 0.0], 0.2))
    assert c == cid(2, 2, 1) and np.allclose(r, [0.1, 0.1, 0.5], atol=1e-14)
    # corner crossing in +x +y +z
    c, r = where(_single3(m, centre, [0.9, 0.9, 0.9], [1.0, 1.0, 1.0], 0.2))
    assert c == cid(2, 2, 2) and np.allclose(r, [0.1, 0.1, 0.1], atol=1e-14)
    # a jump of two cells leaves the vertex star: deleted, then its cell is reseeded
    o = _single3(m, cid(0, 1, 1), [0.5, 0.5, 0.5], [1.0, 0.0, 0.0], 1.7)
    assert 1 not in o.particles()["id"] and o.stats()["lost"] == 1
    # leaving the mesh through a boundary face: deleted
    o = _single3(m, cid(2, 1, 1), [0.9, 0.5, 0.5], [1.0, 0.0, 0.0], 0.5)
    assert 1 not in o.particles()["id"] and o.stats()["lost"] == 1
