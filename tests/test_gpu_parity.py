"""GPU (HIP, FP64 path) versus the CPU oracle on identical seeded inputs.

Bar: cell ownership, ids and per-cell counts bit-exact; x, xi, v and the
projected nodal field bitwise equal to the oracle's (the kernels evaluate the
oracle's expressions in its order, without contraction).  The bar named for
this path is 1e-12 relative; bitwise equality is stricter and is what the
tests assert.
"""
import numpy as np
import pytest

from oracle_binding import Oracle
from parity_util import assert_particles_match, bitwise_equal, compare_particles, field_maxrel
from pfem2_amd import FP32_MIXED, FP64, ParticleStage
from pfem2_amd import mesh as M

pytestmark = pytest.mark.gpu

REL = 1e-12


def _pair(m, V, dt, S, MAX, q, jitter=0.05, seed=0, keep_vext=False, seed_lattice=False):
    g = ParticleStage(m, dt, S, MAX, q, FP64, keep_velocity_ext=keep_vext)
    o = Oracle(m, dt, S, MAX, q)
    for s in (g, o):
        s.upload_fields(V, V)
    if seed_lattice:
        g.seed_particles()
        o.seed_particles()
    else:
        cell, xi = M.jittered_lattice(m, q, jitter, seed)
        g.insert_particles(cell, xi)
        o.insert_particles(cell, xi)
    return g, o


def _run_steps(g, o, V_of_step, steps, where):
    reps = []
    for s in range(steps):
        Vg = g.step(download=True)
        Vo = o.step()
        pg, po = g.particles(), o.particles()
        rep = assert_particles_match(pg, po, REL, f"{where} step {s}", bitwise=True)
        assert np.array_equal(g.cell_counts(), o.cell_counts()), f"{where} step {s}: cell counts"
        fr = field_maxrel(Vg, Vo)
        rep["field_maxrel"] = fr
        rep["field_bitwise"] = bitwise_equal(Vg, Vo)
        assert rep["field_bitwise"], f"{where} step {s}: projected field not bitwise equal (max rel {fr})"
        reps.append(rep)
        Vn = V_of_step(s)
        g.upload_fields(Vn, None)
        o.upload_fields(Vn, None)
    print(where, reps[-1], flush=True)
    return reps


def test_seed_lattice_bitwise():
    m = M.channel_2d(40, 8, 2.0, 1.0)
    V = M.poiseuille_2d(m, 1.0, 1.0, 0.01, seed=3)
    g, o = _pair(m, V, 0.01, 10, 3, [4, 4], seed_lattice=True)
    rep = compare_particles(g.particles(), o.particles())
    print("seed", rep, flush=True)
    assert rep["same_count"] and rep["cell_equal"] and rep["id_equal"]
    assert rep["x_bitwise"] and rep["xi_bitwise"] and rep["v_bitwise"]
    assert g.particle_counter() == o.particle_counter() == m.n_cells * 16


def test_channel_steps_2d():
    """Structured channel (rectangles): closed-form inverse, 10 substeps."""
    m = M.channel_2d(100, 10, 5.0, 1.0)
    V = M.poiseuille_2d(m, 1.0, 1.0, 0.01, seed=4)
    g, o = _pair(m, V, 0.01, 10, 3, [4, 4], seed=5)
    V2 = M.poiseuille_2d(m, 1.0, 1.0, 0.02, seed=6)
    _run_steps(g, o, lambda s: V2 if s % 2 == 0 else V, 6, "channel")
    assert g.particle_counter() == o.particle_counter()
    sg, so = g.stats(), o.stats()
    assert sg["lost"] == so["lost"] and sg["spawned"] == so["spawned"] and sg["deleted"] == so["deleted"]


def test_cylinder_steps_2d():
    """Unstructured-style quads around a cylinder: general bilinear cells."""
    m = M.cylinder_channel(0.15, seed=7)
    V = M.cylinder_flow(m, 1.0, eps=0.02, seed=8)
    g, o = _pair(m, V, 0.004, 10, 3, [4, 4], seed=9)
    _run_steps(g, o, lambda s: V, 5, "cylinder")


def test_high_cfl_relocation_and_losses():
    """Large time step: particles cross several cells per step, many leave
    through the outlet and some are not found by the vertex-star rule."""
    m = M.quad_2d(12, 6, (0.0, 0.0), (3.0, 1.5), jitter=0.2, seed=10)
    V = M.uniform_field(m, [1.0, 0.3]) + 0.2 * M.fourier_perturbation(m.vertices, [3.0, 1.5], seed=11)
    g, o = _pair(m, V, 0.25, 5, 2, [3, 3], seed=12)
    _run_steps(g, o, lambda s: V, 6, "high-cfl")
    assert o.stats()["lost"] > 0


def test_box_steps_3d():
    m = M.box_3d(6, 5, 4, (0, 0, 0), (1.2, 1.0, 0.8), jitter=0.15, seed=13)
    V = M.uniform_field(m, [0.5, 0.2, -0.1]) + 0.3 * M.fourier_perturbation(m.vertices, [1.2, 1.0, 0.8], seed=14)
    g, o = _pair(m, V, 0.05, 4, 2, [2, 2, 2], seed=15)
    _run_steps(g, o, lambda s: V, 4, "box3d")


def test_pipe_steps_3d():
    m = M.pipe_3d(10, 10, 12, 0.5, 1.0)
    V = M.pipe_flow(m, 1.0, 0.5, eps=0.05, seed=16)
    g, o = _pair(m, V, 0.02, 5, 3, [2, 2, 2], seed=17)
    _run_steps(g, o, lambda s: V, 3, "pipe3d")


@pytest.mark.parametrize("kind", ["cyl", "pipe3d"])
@pytest.mark.parametrize("precision", [FP64, FP32_MIXED])
def test_separate_calls_equal_fused_step(precision, kind):
    """pf2_step (which computes the projection sums while regrouping) equals
    the three separate calls bit for bit."""
    if kind == "cyl":
        m = M.cylinder_channel(0.08, seed=18)
        V = M.cylinder_flow(m, 1.0, eps=0.02, seed=19)
        q, dt, MAX = [3, 3], 0.004, 3
    else:
        m = M.pipe_3d(8, 8, 10, 0.5, 1.0)
        V = M.pipe_flow(m, 1.0, 0.5, eps=0.05, seed=19)
        q, dt, MAX = [2, 2, 2], 0.02, 2
    cell, xi = M.jittered_lattice(m, q, 0.05, 20)
    g1 = ParticleStage(m, dt, 5, MAX, q, precision)
    g2 = ParticleStage(m, dt, 5, MAX, q, precision)
    for g in (g1, g2):
        g.upload_fields(V, V)
        g.insert_particles(cell, xi)
    for _ in range(3):
        a = g1.step(download=True)
        g2.correct_particles_velocities()
        g2.move_particles()
        b = g2.distribute_particle_velocities_to_grid()
        assert np.array_equal(a, b, equal_nan=True)
        pa, pb = g1.particles(), g2.particles()
        for k in ("cell", "id", "xi", "v"):
            assert np.array_equal(pa[k], pb[k])
        g1.upload_fields(V, None)
        g2.upload_fields(V, None)


@pytest.mark.parametrize("precision", [FP64, FP32_MIXED])
def test_step_from_empty_state(precision):
    """No particles at all: the move has nothing to do, the census reseeds
    every part of every cell (ids continue the counter from 0) and the
    projection uses the new particles; equal to the oracle (FP64 bitwise,
    FP32-mixed to FP32 rounding of the reseeded lattice points)."""
    m = M.quad_2d(12, 7, (0.0, 0.0), (1.2, 0.7), jitter=0.2, seed=60)
    V = M.uniform_field(m, [0.3, -0.1]) + 0.05 * M.fourier_perturbation(m.vertices, [1.2, 0.7], seed=61)
    g = ParticleStage(m, 0.01, 4, 2, [3, 3], precision)
    o = Oracle(m, 0.01, 4, 2, [3, 3])
    for s in (g, o):
        s.upload_fields(V, V)
    assert g.n_global_particles() == 0
    Vg = g.step(download=True)
    Vo = o.step()
    assert g.n_global_particles() == o.n_global_particles() == m.n_cells * 9
    assert g.particle_counter() == o.particle_counter() == m.n_cells * 9
    pg, po = g.particles(), o.particles()
    assert np.array_equal(pg["cell"], po["cell"]) and np.array_equal(pg["id"], po["id"])
    if precision == FP64:
        assert bitwise_equal(pg["xi"], po["xi"]) and bitwise_equal(pg["v"], po["v"])
        assert bitwise_equal(Vg, Vo)
    else:
        assert np.max(np.abs(pg["xi"] - po["xi"])) <= 1e-6
        assert field_maxrel(Vg, Vo) <= 1e-5


def test_points_on_faces_and_vertices():
    """Particles exactly on cell faces and vertices (xi in {0, 1/2, 1}):
    the part index of xi = 1 is the extra index q (the phantom part of the
    reference's census), own-cell tests at the boundary of the unit square,
    relocations that start on a shared face; bitwise equal to the oracle."""
    m = M.quad_2d(8, 6, (0.0, 0.0), (1.6, 1.2), jitter=0.15, seed=62)
    V = M.uniform_field(m, [0.25, 0.15])
    vals = np.array([0.0, 0.5, 1.0])
    gx, gy = np.meshgrid(vals, vals, indexing="ij")
    pts = np.stack([gx.ravel(), gy.ravel()], axis=1)
    cell = np.repeat(np.arange(m.n_cells, dtype=np.int32), len(pts))
    xi = np.tile(pts, (m.n_cells, 1))
    g = ParticleStage(m, 0.02, 5, 2, [2, 2], FP64)
    o = Oracle(m, 0.02, 5, 2, [2, 2])
    for s in (g, o):
        s.upload_fields(V, V)
        s.insert_particles(cell, xi)
    _run_steps(g, o, lambda s: V, 3, "faces and vertices")


def test_run_twice_bitwise():
    m = M.cylinder_channel(0.1, seed=21)
    V = M.cylinder_flow(m, 1.0, eps=0.02, seed=22)
    outs = []
    for _ in range(2):
        g = ParticleStage(m, 0.004, 10, 3, [4, 4], FP64)
        g.upload_fields(V, V)
        cell, xi = M.jittered_lattice(m, [4, 4], 0.05, 23)
        g.insert_particles(cell, xi)
        for _s in range(4):
            Vp = g.step(download=True)
        outs.append((g.particles(), Vp))
        g.close()
    assert_particles_match(outs[0][0], outs[1][0], 0.0, "run twice", bitwise=True)
    assert bitwise_equal(outs[0][1], outs[1][1])


def test_walls_fixup():
    m = M.channel_2d(40, 8, 4.0, 1.0)
    V = M.poiseuille_2d(m, 1.0, 1.0, 0.05, seed=24)
    g, o = _pair(m, V, 0.01, 5, 3, [3, 3], seed=25)
    wall = np.concatenate([m.boundary_vertices["bottom"], m.boundary_vertices["top"]])
    wall = np.unique(wall)
    bc = np.zeros((len(wall), 2))
    mask = np.full(len(wall), 2, dtype=np.int32)   # normal component only
    body = np.arange(0, len(wall), 7)
    mask[body] = 3                                   # both components (body-style nodes)
    bc[body, 1] = 0.01
    for s in (g, o):
        s.set_walls(wall, wall, bc, mask)
    _run_steps(g, o, lambda s: V, 4, "walls")


@pytest.mark.parametrize("fused", [False, True])
def test_moving_mesh_resort(fused):
    """cylinderDyM order: step, then move vertices, re-sort without advection,
    reseed empty parts (cylinderDyM/cylinderDyM.cc:1147-1160); the fused
    pf2_move_mesh gives the same bits as the three calls."""
    m = M.cylinder_channel(0.08, seed=26)
    V = M.cylinder_flow(m, 1.0, eps=0.02, seed=27)
    g, o = _pair(m, V, 0.004, 5, 3, [3, 3], seed=28, keep_vext=True)
    X0 = m.vertices.copy()
    for s in range(3):
        g.step()
        o.step()
        d = np.hypot(X0[:, 0] - 0.2, X0[:, 1] - 0.2)
        w = np.clip(1.0 - (d - 0.05) / 0.15, 0.0, 1.0)        # displacement decays away from the cylinder
        X = X0.copy()
        X[:, 1] += 0.004 * np.sin(0.7 * (s + 1)) * w
        if fused:
            g.move_mesh(X, recompute_maps=(s == 1))
        else:
            g.update_vertices(X, recompute_maps=(s == 1))
            g.sort_particles_into_subdomains_and_cells()
            g.check_empty_cells()
        o.update_vertices(X, recompute_maps=(s == 1))
        o.sort_particles_into_subdomains_and_cells()
        o.check_empty_cells()
        assert_particles_match(g.particles(), o.particles(), REL, f"moving mesh {s}", bitwise=True)
        assert np.array_equal(g.cell_counts(), o.cell_counts())
        assert g.particle_counter() == o.particle_counter()
        sg, so = g.stats(), o.stats()
        assert sg["lost"] == so["lost"] and sg["spawned"] == so["spawned"] and sg["deleted"] == so["deleted"]


@pytest.mark.parametrize("case", ["q1_3x3", "q1_5x5", "q2", "box3d"])
def test_steps_between_mesh_moves(case):
    """Steps and fused mesh moves in turn, with no other call between them:
    after a mesh move the particles stay in a pending layout that the next
    step's move regroups (the survivors marked, the new particles after the
    current ones).  The particles, counts and statistics stay bitwise equal
    to the oracle's step, update_vertices, re-sort, check_empty_cells; the
    5x5 lattice with MAX 2 also takes the census's second pass and big
    cells, Q2 fields the degree-2 new particles, and a 3-D box hexahedra."""
    if case == "box3d":
        m = M.box_3d(6, 5, 4, (0, 0, 0), (1.2, 1.0, 0.8), jitter=0.15, seed=13)
        V = M.uniform_field(m, [0.5, 0.2, -0.1]) + 0.3 * M.fourier_perturbation(m.vertices, [1.2, 1.0, 0.8], seed=14)
        q, max_per_part, dt = [2, 2, 2], 2, 0.02
        X0 = m.vertices.copy()
        w = np.sin(np.pi * X0[:, 0] / 1.2) * np.sin(np.pi * X0[:, 2] / 0.8)
    else:
        m = M.cylinder_channel(0.08, seed=26)
        if case == "q2":
            m = M.q2(m)
        V = M.cylinder_flow(m, 1.0, eps=0.02, seed=27)
        q, max_per_part, dt = ([5, 5], 2, 0.004) if case == "q1_5x5" else ([3, 3], 3, 0.004)
        X0 = m.vertices.copy()
        d = np.hypot(X0[:, 0] - 0.2, X0[:, 1] - 0.2)
        w = np.clip(1.0 - (d - 0.05) / 0.15, 0.0, 1.0)
    g, o = _pair(m, V, dt, 5, max_per_part, q, seed=28, keep_vext=True)
    for s in range(5):
        g.step()
        o.step()
        X = X0.copy()
        X[:, 1] += 0.004 * np.sin(0.7 * (s + 1)) * w
        g.move_mesh(X)
        o.update_vertices(X)
        o.sort_particles_into_subdomains_and_cells()
        o.check_empty_cells()
        assert g.n_global_particles() == o.n_global_particles(), s
    Vg, Vo = g.step(download=True), o.step()
    assert bitwise_equal(Vg, Vo)
    assert_particles_match(g.particles(), o.particles(), REL, "steps between mesh moves", bitwise=True)
    assert np.array_equal(g.cell_counts(), o.cell_counts())
    # after a mesh move, other uses of the pending layout: the projection
    # alone, and correct then move as separate calls
    X = X0.copy()
    X[:, 1] -= 0.002 * w
    g.move_mesh(X)
    o.update_vertices(X)
    o.sort_particles_into_subdomains_and_cells()
    o.check_empty_cells()
    assert bitwise_equal(g.distribute_particle_velocities_to_grid(), o.distribute_particle_velocities_to_grid())
    X[:, 1] += 0.003 * w
    g.move_mesh(X)
    o.update_vertices(X)
    o.sort_particles_into_subdomains_and_cells()
    o.check_empty_cells()
    for s_ in (g, o):
        s_.correct_particles_velocities()
        s_.move_particles()
    assert bitwise_equal(g.distribute_particle_velocities_to_grid(), o.distribute_particle_velocities_to_grid())
    assert np.array_equal(g.cell_counts(), o.cell_counts())
    assert g.particle_counter() == o.particle_counter()
    sg, so = g.stats(), o.stats()
    assert sg["lost"] == so["lost"] and sg["spawned"] == so["spawned"] and sg["deleted"] == so["deleted"]


@pytest.mark.parametrize("case", ["insert_then_move", "two_updates", "reseed_then_move"])
def test_moving_mesh_resort_keeps_only_exact_xi(case):
    """The re-sort keeps the reference coordinates of particles whose cell
    kept its vertices only when they came from an inverse map in that cell
    under the current geometry.  Not so: particles inserted and moved before
    any advection, any particle after two geometry updates without a re-sort
    between them, particles reseeded after the last advection.  All three
    stay bitwise equal to the oracle, which recomputes every particle."""
    m = M.cylinder_channel(0.08, seed=34)
    V = M.cylinder_flow(m, 1.0, eps=0.02, seed=35)
    g, o = _pair(m, V, 0.004, 5, 3, [3, 3], seed=36, keep_vext=True)
    X0 = m.vertices.copy()
    d = np.hypot(X0[:, 0] - 0.2, X0[:, 1] - 0.2)
    w = np.clip(1.0 - (d - 0.05) / 0.15, 0.0, 1.0)
    X1, X2 = X0.copy(), X0.copy()
    X1[:, 1] += 0.004 * w
    X2[:, 1] -= 0.003 * w
    if case == "insert_then_move":
        g.move_mesh(X1)
        o.update_vertices(X1)
    else:
        g.step()
        o.step()
        if case == "two_updates":
            g.update_vertices(X1)
            g.update_vertices(X2)
            g.sort_particles_into_subdomains_and_cells()
            o.update_vertices(X1)
            o.update_vertices(X2)
        else:
            g.check_empty_cells()
            o.check_empty_cells()
            g.move_mesh(X1)
            o.update_vertices(X1)
    o.sort_particles_into_subdomains_and_cells()
    if case != "two_updates":
        o.check_empty_cells()
    assert_particles_match(g.particles(), o.particles(), REL, f"re-sort {case}", bitwise=True)
    assert np.array_equal(g.cell_counts(), o.cell_counts())


@pytest.mark.parametrize("q", [[4, 4], [5, 5]])
def test_crowded_cells_bitwise(q):
    """Cells holding more candidates than the per-group LDS rows (64 for up
    to 16 lattice parts, 128 above) go through the block-per-cell census;
    results stay bitwise equal."""
    m = M.quad_2d(10, 6, (0.0, 0.0), (2.0, 1.2), jitter=0.15, seed=30)
    V = M.uniform_field(m, [0.4, 0.1]) + 0.1 * M.fourier_perturbation(m.vertices, [2.0, 1.2], seed=31)
    cell, xi = M.jittered_lattice(m, q, 0.05, 32)
    rng = np.random.default_rng(33)
    extra_cells = np.concatenate([np.full(150, 5), np.full(300, 17), np.full(90, 33)]).astype(np.int32)
    extra_xi = rng.uniform(0.0, 1.0, size=(len(extra_cells), 2))
    cell = np.concatenate([cell, extra_cells])
    xi = np.concatenate([xi, extra_xi])
    order = np.argsort(cell, kind="stable")
    g = ParticleStage(m, 0.05, 5, 3, q, FP64, capacity=20000)
    o = Oracle(m, 0.05, 5, 3, q)
    for s in (g, o):
        s.upload_fields(V, V)
        s.insert_particles(cell[order], xi[order])
    _run_steps(g, o, lambda s: V, 3, f"crowded q={q}")


def test_overfull_cell_reports_capacity_error():
    """More than 2048 candidates in one cell: the step reports a capacity
    error instead of a wrong result."""
    from pfem2_amd import Pf2Error
    m = M.quad_2d(4, 4, (0.0, 0.0), (1.0, 1.0))
    V = M.uniform_field(m, [0.0, 0.0])
    rng = np.random.default_rng(34)
    n = 2100
    g = ParticleStage(m, 0.01, 2, 3, [2, 2], FP64, capacity=3 * n)
    g.upload_fields(V, V)
    g.insert_particles(np.full(n, 6, dtype=np.int32), rng.uniform(0.0, 1.0, size=(n, 2)))
    g.step()
    with pytest.raises(Pf2Error, match="2048"):
        g.n_global_particles()


@pytest.mark.parametrize("graphs", [False, True])
@pytest.mark.parametrize("precision", [FP64, FP32_MIXED])
def test_capacity_overflow_is_sticky(precision, graphs):
    """A reseed that needs more particles than the capacity: 64 particles in
    cell 0 of a 4x4 mesh with capacity 64; the census keeps 3 per part of
    cell 0 (12) and spawns 4 in each of the 15 empty cells (60), 72 > 64.
    The write skips the cells that do not fit and flags the overflow; two
    more steps queued before the host looks (eager or graph replays) do no
    work, and every synchronising call reports the capacity error."""
    from pfem2_amd import Pf2Error
    m = M.quad_2d(4, 4, (0.0, 0.0), (1.0, 1.0))
    V = M.uniform_field(m, [0.0, 0.0])
    rng = np.random.default_rng(35)
    g = ParticleStage(m, 0.01, 2, 3, [2, 2], precision, capacity=64)
    g.upload_fields(V, V)
    g.insert_particles(np.zeros(64, dtype=np.int32), rng.uniform(0.05, 0.95, size=(64, 2)))
    assert g.n_global_particles() == 64
    g.set_graphs(graphs)
    for _ in range(3):
        g.step()
    with pytest.raises(Pf2Error, match="capacity"):
        g.n_global_particles()
    with pytest.raises(Pf2Error, match="capacity"):
        g.cell_counts()
    with pytest.raises(Pf2Error, match="capacity"):
        g.step(download=True)
    g.close()
    # the library stays usable: a fresh handle with room steps normally
    h = ParticleStage(m, 0.01, 2, 3, [2, 2], precision)
    h.upload_fields(V, V)
    h.seed_particles()
    h.step()
    assert h.n_global_particles() == m.n_cells * 4
    h.close()
