"""HIP-graph replay of the step and the serving loop.

* A step replayed from a captured HIP graph gives the same bits as the same
  step launched eagerly, on every replay, with eager work and host
  synchronisation between the replays (FP64 and FP32-mixed, 2-D and 3-D).
* A simulation served by pf2_server gives the same bits as the same
  simulation run directly on a handle, whether it is served alone or among
  other simulations ("alone == in mix").
* The footprint estimate used for admission bounds what a handle takes.
"""
import numpy as np
import pytest

from pfem2_amd import FP32_MIXED, FP64, ParticleStage, Server, estimate_bytes
from pfem2_amd import mesh as M

pytestmark = pytest.mark.gpu


def _device_field(V):
    import torch
    t = torch.as_tensor(np.ascontiguousarray(V, dtype=np.float64)).cuda().contiguous()
    torch.cuda.synchronize()
    return t


def _case(kind):
    if kind == "cyl":
        m = M.cylinder_channel(0.06, seed=41)
        V = M.cylinder_flow(m, 1.0, eps=0.02, seed=42)
        return m, V, (0.003, 8, 2, [4, 4])
    m = M.box_3d(6, 6, 6, jitter=0.15, seed=43)
    V = M.uniform_field(m, [0.3, -0.2, 0.25])
    return m, V, (0.02, 6, 2, [2, 2, 2])


def _run(m, V, prm, precision, graphs, steps, checks):
    dt, S, MAX, q = prm
    st = ParticleStage(m, dt, S, MAX, q, precision)
    st.upload_fields(V, V)
    st.seed_particles()
    st.set_graphs(graphs)
    dV = _device_field(V)
    out = []
    for s in range(steps):
        st.step_with_field(dV.data_ptr())
        if s in checks:
            # eager work and host synchronisation between replays
            p = st.particles()
            out.append((st.download_field(), p["cell"], p["id"], p["xi"], p["v"], st.cell_counts()))
    st.close()
    return out


@pytest.mark.parametrize("kind", ["cyl", "box3d"])
@pytest.mark.parametrize("precision", [FP64, FP32_MIXED])
def test_graph_replay_matches_eager(kind, precision):
    m, V, prm = _case(kind)
    checks = {0, 1, 2, 4, 7}
    print(f"graph replay case={kind} precision={precision}", flush=True)
    a = _run(m, V, prm, precision, False, 8, checks)
    b = _run(m, V, prm, precision, True, 8, checks)
    assert len(a) == len(b) == len(checks)
    for ra, rb in zip(a, b):
        for xa, xb in zip(ra, rb):
            assert np.array_equal(xa, xb, equal_nan=True)


def _run_plain(m, V, prm, precision, graphs, steps, checks, uploads):
    """pf2_step (no field): the node pass refreshes the per-cell copies of V
    for the next step; an upload in between makes them stale again."""
    dt, S, MAX, q = prm
    st = ParticleStage(m, dt, S, MAX, q, precision)
    st.upload_fields(V, V)
    st.seed_particles()
    st.set_graphs(graphs)
    out = []
    for s in range(steps):
        if s in uploads:
            st.upload_fields(0.9 * V, V)
        Vp = st.step(download=True)
        if s in checks:
            p = st.particles()
            out.append((Vp, p["cell"], p["id"], p["xi"], p["v"], st.cell_counts()))
    st.close()
    return out


@pytest.mark.parametrize("kind", ["cyl", "box3d"])
@pytest.mark.parametrize("precision", [FP64, FP32_MIXED])
def test_graph_replay_of_plain_step_matches_eager(kind, precision):
    m, V, prm = _case(kind)
    checks = {0, 1, 3, 4, 6}
    uploads = {3, 5}
    print(f"plain step graph replay case={kind} precision={precision}", flush=True)
    a = _run_plain(m, V, prm, precision, False, 7, checks, uploads)
    b = _run_plain(m, V, prm, precision, True, 7, checks, uploads)
    assert len(a) == len(b) == len(checks)
    for ra, rb in zip(a, b):
        for xa, xb in zip(ra, rb):
            assert np.array_equal(xa, xb, equal_nan=True)


def _direct(m, V, prm, steps):
    dt, S, MAX, q = prm
    st = ParticleStage(m, dt, S, MAX, q, FP64)
    st.upload_fields(V, V)
    st.seed_particles()
    dV = _device_field(V)
    for _ in range(steps):
        st.step_with_field(dV.data_ptr())
    Vout = st.download_field()
    n = st.n_global_particles()
    st.close()
    return Vout, n


def _requests():
    out = []
    for k, (nx, ny) in enumerate([(60, 12), (90, 18), (40, 8), (120, 24)]):
        m = M.channel_2d(nx, ny, 5.0, 1.0)
        V = M.poiseuille_2d(m, 0.5 + 0.3 * k, 1.0, 0.02, seed=50 + k)
        out.append((m, V, (0.01, 5, 2, [3 + k % 2, 3 + k % 2]), 6 + 3 * k))
    m = M.cylinder_channel(0.05, seed=60)
    out.append((m, M.cylinder_flow(m, 1.0, eps=0.02, seed=61), (0.003, 6, 2, [4, 4]), 9))
    return out


def _server_kw(mode):
    # batched: every request in batched steps; mixed: those of at most 1000
    # cells batched, the others on their own streams
    return {"batched": mode != "single", "batch_max_cells": 1000 if mode == "mixed" else 0}


@pytest.mark.parametrize("batched", ["batched", "single", "mixed"])
def test_server_alone_equals_mix_equals_direct(batched):
    """A request served alone, in a mix of five and run directly on a handle
    gives the same bits; batched: the mix advances with one launch per
    kernel of the step for all active requests (mixed: the small ones)."""
    reqs = _requests()
    target = 4   # the cylinder request
    m, V, prm, steps = reqs[target]
    ref_V, ref_n = _direct(m, V, prm, steps)

    print("server alone", flush=True)
    srv = Server(hbm_budget=4 << 30, steps_per_chunk=3, use_graphs=True, **_server_kw(batched))
    rid = srv.submit(m, V, steps, *prm)
    srv.drain()
    V1, n1, lat1 = srv.result(rid)
    st = srv.stats()
    srv.close()
    assert st["finished"] == 1 and st["failed"] == 0
    assert n1 == ref_n
    assert np.array_equal(V1, ref_V, equal_nan=True)

    print("server mix", flush=True)
    srv = Server(hbm_budget=4 << 30, steps_per_chunk=2, chunks_ahead=2, use_graphs=True, **_server_kw(batched))
    ids = []
    for k, (mk, Vk, pk, sk) in enumerate(reqs):
        ids.append(srv.submit(mk, Vk, sk, *pk))
        srv.poll()
    srv.drain()
    res = [srv.result(i) for i in ids]
    st = srv.stats()
    srv.close()
    assert st["finished"] == len(reqs) and st["failed"] == 0
    assert st["max_active"] >= 2
    V2, n2, _ = res[target]
    assert n2 == ref_n
    assert np.array_equal(V2, ref_V, equal_nan=True)
    # every other request also equals its direct run
    for k in (0, 3):
        mk, Vk, pk, sk = reqs[k]
        Vd, nd = _direct(mk, Vk, pk, sk)
        assert res[k][1] == nd
        assert np.array_equal(res[k][0], Vd, equal_nan=True)


def test_server_admission_by_footprint():
    reqs = _requests()
    sizes = [estimate_bytes(m, *prm) for m, V, prm, s in reqs]
    # a budget that holds the largest request but not all of them at once
    budget = int(max(sizes) * 1.6) + (8 << 20)
    srv = Server(hbm_budget=budget, steps_per_chunk=2, use_graphs=False)
    ids = [srv.submit(m, V, s, *prm) for m, V, prm, s in reqs]
    srv.drain()
    st = srv.stats()
    assert st["finished"] == len(reqs) and st["failed"] == 0
    assert st["peak_bytes"] <= budget
    for i in ids:
        assert srv.result(i) is not None
    srv.close()
    # a request larger than the whole budget is refused at submission
    srv = Server(hbm_budget=8 << 20, use_graphs=False)
    m, V, prm, s = reqs[3]
    with pytest.raises(Exception):
        srv.submit(m, V, s, *prm)
    srv.close()


@pytest.mark.parametrize("precision", [FP64, FP32_MIXED])
def test_footprint_estimate_bounds_handle(precision):
    for m, V, prm, s in _requests():
        est = estimate_bytes(m, *prm, precision=precision)
        st = ParticleStage(m, *prm, precision=precision)
        got = st.memory_bytes()
        st.close()
        assert got <= est <= got + (2 << 20)


def _moving_case():
    m = M.cylinder_channel(0.06, seed=81)
    V = M.cylinder_flow(m, 1.0, eps=0.02, seed=82)
    d = np.hypot(m.vertices[:, 0] - 0.2, m.vertices[:, 1] - 0.2)
    w = np.clip(1.0 - (d - 0.05) / 0.1, 0.0, 1.0)
    W = np.zeros_like(m.vertices)
    W[:, 1] = w
    return m, V, W


@pytest.mark.parametrize("precision", [FP64, FP32_MIXED])
def test_prescribed_motion_equals_host_vertices(precision):
    """pf2_move_mesh_scaled(a) = pf2_move_mesh(X0 + a W) bit for bit."""
    m, V, W = _moving_case()
    runs = []
    for scaled in (False, True):
        st = ParticleStage(m, 0.003, 6, 2, [3, 3], precision, keep_velocity_ext=True)
        st.upload_fields(V, V)
        st.seed_particles()
        if scaled:
            st.set_motion(W)
        X0 = m.vertices.copy()
        for k in range(1, 5):
            st.step()
            a = 0.004 * np.sin(0.9 * k)
            if scaled:
                st.move_mesh_scaled(a)
            else:
                st.move_mesh(X0 + a * W)
            st.upload_fields(V, None)
        p = st.particles()
        runs.append((st.step(download=True), p["cell"], p["id"], p["xi"], p["v"]))
        st.close()
    for a, b in zip(*runs):
        assert np.array_equal(a, b, equal_nan=True)


@pytest.mark.parametrize("batched", [True, False])
def test_server_moving_request_alone_equals_mix_equals_direct(batched):
    m, V, W = _moving_case()
    dt, S, MAX, q, steps = 0.003, 6, 2, [3, 3], 7
    amp, omega, phase = 0.004, 150.0, 0.3
    st = ParticleStage(m, dt, S, MAX, q, FP32_MIXED, keep_velocity_ext=True)
    st.upload_fields(V, V)
    st.seed_particles()
    st.set_motion(W)
    dV = _device_field(V)
    for k in range(1, steps + 1):
        st.step_with_field(dV.data_ptr())
        st.move_mesh_scaled(amp * np.sin(omega * k * dt + phase))
    ref_V, ref_n = st.download_field(), st.n_global_particles()
    st.close()
    res = []
    for mix in (False, True):
        srv = Server(hbm_budget=4 << 30, steps_per_chunk=3, use_graphs=True, batched=batched)
        ids = []
        if mix:
            for mk, Vk, pk, sk in _requests()[:3]:
                ids.append(srv.submit(mk, Vk, sk, *pk))
        rid = srv.submit_moving(m, V, steps, dt, S, MAX, q, W, amp, omega, phase, precision=FP32_MIXED)
        srv.drain()
        st_ = srv.stats()
        assert st_["failed"] == 0 and st_["finished"] == len(ids) + 1
        res.append(srv.result(rid))
        srv.close()
    for Vr, nr, _ in res:
        assert nr == ref_n
        assert np.array_equal(Vr, ref_V, equal_nan=True)
