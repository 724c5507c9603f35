"""Benchmark of the PFEM2 particle stage on one MI355X.

Value (the largest configuration of the metric that fits one GPU): the
moving-mesh cylinder channel, ~400k graded quads around a cylinder that
oscillates across the channel (amplitude 0.1 D), 25 particles per cell
(5x5 lattice, +-0.05 seeded jitter, ~10.2M particles at the start), the
flow around the cylinder blended into the wall's velocity near it,
dt = 0.0005, 10 advection substeps, MAX = 2, on the FP64 path (the
reference's arithmetic, bitwise equal to the CPU oracle).  One step =
correct -> move (10 substeps of advection + relocation, then reseed/delete)
-> project, for every particle, after the FEM stand-in (the step's nodal V,
a device-resident field, copied into the stage), then the mesh move (vertex
update, re-sort of every particle, reseed).  All inputs are resident in HBM
before the timed region.

The line also carries the same workload on the FP32-mixed path, the
~100k-cell cylinder channel (16 particles per cell, dt = 0.001) in both
precisions with the FP32-mixed path's per-field error against FP64 and its
targets (met or not), the CPU oracle on a sample of the value's workload
(one core, and one simulation per host core at once), and the request
stream of the latency metric (256 independent simulations, Poisson
arrivals, 2k-200k cells, 9/16/25 particles per cell, 50-500 steps, plus 8
moving-mesh simulations of ~400k cells; admission by HBM footprint) with
its p50/p99 completion latency under "serving" (each simulation replays its
own step graphs on its own stream), and the same stream with batched steps
under "serving_batched".

Other workloads (--workload): cylinder (the 100k-cell channel alone),
channel (200x20 quads, FP64, with a host FEM stand-in between particle
steps, against the CPU oracle over the whole run), pipe3d (826k hexahedra
in a pipe with a Q2 field, 2x2x2 lattice), moving (the value's workload alone), serve (the
request stream alone).  Prints one JSON line.
"""
import argparse
import json
import math
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SEEDS = {"mesh": 2024, "field": 7, "particles": 11, "serve": 99, "pipe": 31, "moving": 57}
HBM_PEAK_GBS = 8000.0
# committed rocprofv3 --pmc summaries (profiles/hbm_traffic_<name>.txt) per workload and precision
PROFILES = {("moving", 0): "moving400k_fp64", ("cylinder", 0): "cylinder100k_fp64",
            ("cylinder", 1): "cylinder100k_fp32mixed"}
DTYPE32 = "fp32-mixed (f32 reference coords/velocities, f64 relocation and node sums)"


class Workload:
    """Mesh, field, lattice and stepping parameters of one benchmark case."""

    def __init__(self, name, m, V, q, dt, substeps, max_per_part, jitter=0.05, keep_vext=False, motion=None,
                 field=None, desc=""):
        self.name, self.m, self.V, self.q = name, m, V, q
        self.dt, self.substeps, self.max_per_part = dt, substeps, max_per_part
        self.jitter, self.keep_vext, self.motion, self.desc = jitter, keep_vext, motion, desc
        self.field = field   # field(step) -> nodal V of that step (None: V every step)


def cylinder_workload(args):
    from pfem2_amd import mesh as M
    m = M.cylinder_channel(args.scale, seed=SEEDS["mesh"])
    V = M.cylinder_flow(m, 1.0, eps=0.01, seed=SEEDS["field"])
    return Workload("cylinder", m, V, [4, 4], args.dt, args.substeps, args.max_per_part,
                    desc="2D channel with circular obstacle (2.2x0.41, r=0.05), graded quad mesh, 16 particles/cell, "
                         "potential flow + 1% Fourier perturbation, step = correct+move(10 substeps)+project")


def channel_workload():
    from pfem2_amd import mesh as M
    m = M.channel_2d(200, 20, 10.0, 1.0)
    V = M.poiseuille_2d(m, 1.0, 1.0, 0.01, seed=SEEDS["field"])
    return Workload("channel", m, V, [4, 4], 0.01, 10, 3,
                    desc="200x20 quads on [0,10]x[0,1], Poiseuille Umax=1 + 1% Fourier perturbation, 4x4 lattice "
                         "+-0.05 jitter, dt=0.01, 10 substeps, MAX=3, FP64")


def pipe_workload():
    from pfem2_amd import mesh as M
    m = M.q2(M.pipe_3d(64, 64, 256))
    V = M.pipe_flow(m, 1.0, 0.5, 0.01, seed=SEEDS["pipe"])
    return Workload("pipe3d", m, V, [2, 2, 2], 0.01, 10, 2,
                    desc="3D pipe r=0.5 L=4, 64x64x256 hexahedra clipped to the circle, Q2 Poiseuille field in z + 1% "
                         "Fourier perturbation, 2x2x2 lattice +-0.05 jitter, dt=0.01, 10 substeps, MAX=2")


def moving_workload(scale=2.0):
    """~400k quads around a cylinder oscillating across the channel
    (amplitude 0.1 D, 2 Hz); the displacement blends from rigid at the
    cylinder to zero at 3 radii."""
    from pfem2_amd import mesh as M
    m = M.cylinder_channel(scale, seed=SEEDS["moving"])
    V = M.cylinder_flow(m, 1.0, eps=0.01, seed=SEEDS["field"])
    X0 = m.vertices.copy()
    c = np.array([0.2, 0.2])
    R, Rout, A, omega = 0.05, 0.15, 0.01, 2.0 * math.pi * 2.0
    r = np.linalg.norm(X0 - c, axis=1)
    t = np.clip((r - R) / (Rout - R), 0.0, 1.0)
    w = 0.5 * (1.0 + np.cos(math.pi * t))

    def motion(step, dt):
        X = X0.copy()
        X[:, 1] += A * math.sin(omega * step * dt) * w
        return X

    def field(step, dt):
        # the flow around the cylinder blended into the wall's velocity near
        # it, so that no fluid crosses the moving wall
        Vs = V * (1.0 - w)
        Vs[1] += w * A * omega * math.cos(omega * step * dt)
        return Vs

    # dt = 0.0005: the default workload's per-substep Courant number at half its cell size
    wl = Workload("moving", m, V, [5, 5], 0.0005, 10, 2, keep_vext=True, motion=motion, field=field,
                  desc="~400k graded quads around a cylinder oscillating vertically (amplitude 0.1 D, 2 Hz), "
                       "flow blended into the wall velocity near it, 25 particles/cell, dt=0.0005, per step: "
                       "correct+move+project, then the mesh move (vertex update, re-sort, reseed)")
    wl.blend, wl.amp, wl.omega = w, A, omega     # X(t) = X0 + amp sin(omega t) (0, blend)
    return wl


def kernel_bytes(name, info, m, dim=2, vext=False):
    """Algorithmic HBM bytes of one launch of a step kernel, from the work
    counts of the step (compulsory traffic: every array element touched once).
    vext: the handle keeps the last advection velocity (moving mesh), which
    the move writes and the regroup carries along."""
    nc, nd = m.n_cells, m.n_dofs
    nv = 1 << dim
    n0, n1 = info["n_before"], info["n_after"]
    q = info["queued"]
    arr = max(info["arrivals"], 0)
    if name in ("k_move", "k_move32"):
        wx, wv = (8, 8) if name == "k_move" else (4, 4)
        reads = n0 * (4 + dim * wx * (3 if name == "k_move" else 1) + dim * wv)
        writes = n0 * dim * wv + (n0 - q[0]) * (dim * wx * (2 if name == "k_move" else 1) + 8) \
            + q[0] * (dim * wx * (2 if name == "k_move" else 1) + 4 + 4 + 4 + 8 * dim)
        mesh = nc * nv * 4 + nd * dim * wv * 2 + (m.n_vertices * dim * 8 if name == "k_move" else nc * 24)
        return reads + writes + mesh + (n0 * dim * 8 if vext else 0)
    if name == "k_resort":
        # own-cell test of every particle after a mesh move: cell, id, xi in,
        # destination and part key out; the q[2] it parks for k_resort_search
        # also read x and vext and store the parking record (x, xi,
        # velocity_ext, substep, queue entry); per cell: vertices, moved flag
        return n0 * (4 + 4 + dim * 8 + 8) + q[2] * (2 * dim * 8 + 3 * dim * 8 + 4 + 8) + nc * (nv * dim * 8 + 1)
    if name == "k_resort32":
        # the FP32 state: cell, id, FP32 reference point in; destination and
        # part key out; the q[2] parked: FP64 position in, substep flag and
        # queue entry out
        return n0 * (4 + 4 + dim * 4 + 8) + q[2] * (dim * 8 + 4 + 8) + nc * (nv * dim * 8 + 1)
    if name == "k_resort_search32":
        # per queued particle: entry, FP64 position and velocity_ext, FP32
        # reference point, vertices in; reference point, destination, key out
        return int(q[2] * (8 + 4 + 2 * dim * 8 + dim * 4 + nv * dim * 8 + dim * 4 + 8))
    if name == "k_resort_search":
        # the relocations of the re-sort (q[2]): entry, record and vertices
        # in; x, xi, destination and part key out
        return int(q[2] * (8 + 4 + 3 * dim * 8 + nv * dim * 8 + 2 * dim * 8 + 8))
    if name.startswith("k_search_resume"):
        # one launch per relocation round; round r handles q[r] particles
        mm = re.match(r"(k_search_resume(?:32)?)(?:_r(\d))?$", name)
        base, r = mm.group(1), int(mm.group(2) or 0)
        wx = 8 if base == "k_search_resume" else 4
        per = 8 + 4 + 4 + dim * wx * (2 if base == "k_search_resume" else 1) + 8 * dim + 8 + dim * wx
        return int(q[r] * per) if r < len(q) else None
    if name == "k_cell_count":
        return (n0 + arr) * 12 + arr * 4 + n0 * 4 + nc * (8 + 4 + 4 + 20)
    if name in ("k_cell_write", "k_cell_write32"):
        wx = 8 if name == "k_cell_write" else 4
        per_src = 4 + 4 + 4 + dim * wx * (3 if name == "k_cell_write" else 2) + (dim * 8 if vext else 0)
        per_dst = 4 + 4 + dim * wx * (3 if name == "k_cell_write" else 2) + (dim * 8 if vext else 0)
        return n1 * (per_src + per_dst) + nc * 32
    if name == "k_cell_pend":
        # the pending layout after a re-sort (the census marked the
        # survivors): per cell the spawn count in and the resets out; new
        # particles written whole (few)
        return nc * (8 + 8)
    if name == "k_cell_map":
        # per-cell census results in, the output map and per-cell resets out
        return nc * (36 + 12) + n1 * (4 + 4)
    if name in ("k_cell_gather", "k_cell_gather32"):
        # map and destination of each output position, the record copied in order
        rec = dim * 8 * 3 if name == "k_cell_gather" else (16 if dim == 2 else 24)
        per_src = 4 + 4 + 4 + rec + (dim * 8 if vext else 0)
        per_dst = 4 + 4 + rec + (dim * 8 if vext else 0)
        return n1 * (per_src + per_dst)
    if name in ("k_cell_partials", "k_cell_partials32"):
        wx = 8 if name == "k_cell_partials" else 4
        return n1 * dim * wx * 2 + nc * (4 + nv * (dim + 1) * 8)
    if name == "k_node":
        return nc * nv * (dim + 1) * 8 + nd * (dim * 8 * 2 + 4) + nc * nv * 4
    return None


def make_stage(W, precision):
    from pfem2_amd import ParticleStage
    from pfem2_amd import mesh as M
    st = ParticleStage(W.m, W.dt, W.substeps, W.max_per_part, W.q, precision, keep_velocity_ext=W.keep_vext)
    st.upload_fields(W.V, W.V)
    cell, xi = M.jittered_lattice(W.m, W.q, W.jitter, SEEDS["particles"])
    st.insert_particles(cell, xi)
    return st


def run_path(args, W, precision, torch):
    """Timed steps (graph replay) of one precision path, then synchronised
    per-step latencies, an eager (no graph) timing and per-kernel HIP-event
    times with the work counts of the same steps."""
    st = make_stage(W, precision)
    n0 = st.n_global_particles()
    dV = torch.from_numpy(np.ascontiguousarray(W.V)).to("cuda")   # FEM stand-in output, resident in HBM
    torch.cuda.synchronize()
    ptr = dV.data_ptr()
    lib = st._lib
    k = [0]
    Fs = None
    total = args.warmup + 2 * args.steps + max(args.steps, 20) + 12
    if W.motion is not None:
        # the prescribed motion X0 + a(t) W resident in HBM: each step forms its
        # vertices on the device (X0 + a W, bit for bit the host's
        # X0 + (amp sin(omega t)) blend), so no vertex upload is timed
        Wd = np.zeros_like(W.m.vertices)
        Wd[:, 1] = W.blend
        st.set_motion(Wd)
    # fields of every step, prepared and resident in HBM outside the timed region
    if W.field is not None:
        Fs = [torch.from_numpy(np.ascontiguousarray(W.field(i, W.dt))).to("cuda") for i in range(total)]
        torch.cuda.synchronize()

    def step():
        p = Fs[k[0] % len(Fs)].data_ptr() if Fs is not None else ptr
        rc = lib.pf2_step_with_field(st._h, p)
        if rc:
            st._check(rc)
        if W.motion is not None:
            # moving mesh, in the order of the reference's moving-mesh driver:
            # vertices, re-sort of the particles, reseed (one fused call)
            st.move_mesh_scaled(W.amp * math.sin(W.omega * ((k[0] % total) + 1) * W.dt))
        k[0] += 1

    # a field that changes every step would be re-captured every step
    st.set_graphs(Fs is None)
    for _ in range(args.warmup):
        step()
    n_start = st.n_global_particles()     # synchronises: the count entering the timed window
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    st.synchronize()
    t1 = time.perf_counter()
    n1 = st.n_global_particles()
    ms = (t1 - t0) * 1e3 / args.steps
    # particles stepped per timed step: the counts entering and leaving the
    # timed window, averaged (the count changes by reseeding and losses only)
    updates = 0.5 * (n_start + n1)
    lat = []
    for _ in range(max(args.steps, 20)):
        a = time.perf_counter()
        step()
        st.synchronize()
        lat.append((time.perf_counter() - a) * 1e3)
    st.set_graphs(False)
    te0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    st.synchronize()
    ms_eager = (time.perf_counter() - te0) * 1e3 / args.steps
    ms_host_vertices = None
    if W.motion is not None:
        # the same steps with the vertices uploaded from host memory each step
        # (the PCIe-inclusive rate of a caller whose mesh motion lives on the host)
        Xh = [np.ascontiguousarray(W.motion(((k[0] + i) % total) + 1, W.dt)) for i in range(args.steps)]
        th0 = time.perf_counter()
        for i in range(args.steps):
            p = Fs[k[0] % len(Fs)].data_ptr() if Fs is not None else ptr
            st._check(lib.pf2_step_with_field(st._h, p))
            st.move_mesh(Xh[i])
            k[0] += 1
        st.synchronize()
        ms_host_vertices = (time.perf_counter() - th0) * 1e3 / args.steps
    st.set_profiling(True)
    infos = []
    for _ in range(min(args.steps, 10)):
        step()
        infos.append(st.step_info())
    kt = st.kernel_times()
    nprof = len(infos)
    st.set_profiling(False)
    info = {key: (np.mean([i[key] for i in infos], axis=0).tolist() if key == "queued" else
                  float(np.mean([i[key] for i in infos]))) for key in infos[0]}
    res = {
        "updates_per_s": updates / (ms * 1e-3),
        "ms_per_step": ms,
        "ms_per_step_eager": ms_eager,
        "ms_per_step_host_vertices": ms_host_vertices,
        "particles_inserted": int(n0),
        "particles_start": int(n_start),
        "particles": int(n1),
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "kernels_ms_avg": {key: v[0] / max(v[1], 1) for key, v in kt.items()},
        "kernels_ms_per_step": {key: v[0] / nprof for key, v in kt.items()},
        "kernel_launches_per_step": {key: v[1] / nprof for key, v in kt.items()},
        "step_info": info,
        "keep_vext": W.keep_vext,
        "stats": st.stats(),
        "memory_bytes": st.memory_bytes(),
    }
    st.close()
    return res


def accuracy(W, steps):
    """Per-field error of the FP32-mixed path against the FP64 path (which is
    bitwise equal to the CPU oracle) after `steps` steps."""
    from pfem2_amd import FP32_MIXED, FP64
    a = make_stage(W, FP64)
    b = make_stage(W, FP32_MIXED)
    n0 = a.n_global_particles()
    for _ in range(steps):
        Va = a.step(download=True)
        Vb = b.step(download=True)
        a.upload_fields(W.V, None)
        b.upload_fields(W.V, None)
    pa, pb = a.particles(), b.particles()
    ia = {int(k): n for n, k in enumerate(pa["id"]) if k <= n0}
    pairs = [(ia[int(k)], n) for n, k in enumerate(pb["id"]) if k <= n0 and int(k) in ia]
    ai = np.array([p[0] for p in pairs])
    bi = np.array([p[1] for p in pairs])
    same = pa["cell"][ai] == pb["cell"][bi]
    m = W.m
    L = float(np.max(m.vertices.max(axis=0) - m.vertices.min(axis=0)))
    out = {
        "steps": steps,
        "V_rel_l2": float(np.linalg.norm(Vb - Va) / np.linalg.norm(Va)),
        "V_max_rel": float(np.max(np.abs(Vb - Va)) / np.max(np.abs(Va))),
        "x_max_rel": float(np.max(np.abs(pb["x"][bi] - pa["x"][ai])) / L),
        "v_max_rel": float(np.max(np.abs(pb["v"][bi][same] - pa["v"][ai][same])) / np.max(np.abs(pa["v"]))),
        "owner_mismatch": float(1.0 - same.mean()),
        "count_rel_diff": float(abs(len(pb["id"]) - len(pa["id"])) / len(pa["id"])),
    }
    a.close()
    b.close()
    return out


def _oracle(W):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_binding import Oracle
    from pfem2_amd import mesh as M
    o = Oracle(W.m, W.dt, W.substeps, W.max_per_part, W.q)
    o.upload_fields(W.V, W.V)
    cell, xi = M.jittered_lattice(W.m, W.q, W.jitter, SEEDS["particles"])
    o.insert_particles(cell, xi)
    return o


def cpu_steps(W, o, seconds, max_steps=50):
    """Full steps of workload W on one oracle until about `seconds` of CPU
    time (at least one step).  A step is what the GPU step is: correct +
    move + project (timed inside the oracle) and, for a moving mesh, the
    vertex update, re-sort and reseed (timed here).  Returns (particle-updates,
    seconds, steps)."""
    done, t, upd = 0, 0.0, 0
    while (done == 0 or t < seconds) and done < max_steps:
        if W.field is not None:
            o.upload_fields(W.field(done, W.dt), None)
        n = o.n_global_particles()
        t += o.time_steps(1)
        if W.motion is not None:
            X = W.motion(done + 1, W.dt)
            a = time.perf_counter()
            o.update_vertices(X)
            o.sort_particles_into_subdomains_and_cells()
            o.check_empty_cells()
            t += time.perf_counter() - a
        elif W.field is None:
            o.upload_fields(W.V, None)
        upd += n
        done += 1
    return upd, t, done


def host_threads():
    """Host cores this process may use, at most 16 (a one-GPU box's share)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline(W, seconds=10.0, threads=None):
    """The oracle (the CPU restatement of the reference's particle stage,
    FP64, built -O3) on a bounded sample of workload W: one simulation on one
    core, then `threads` independent simulations at once (the reference runs
    one MPI rank per core; the oracle releases the interpreter lock, so the
    simulations run in parallel threads), each stepped for about `seconds`."""
    import threading
    o = _oracle(W)
    upd1, t1, k1 = cpu_steps(W, o, seconds)
    o.close()
    threads = threads or host_threads()
    res = [None] * threads
    ready = threading.Barrier(threads + 1)

    def run(j):
        oj = _oracle(W)
        ready.wait()
        res[j] = cpu_steps(W, oj, seconds, max_steps=k1)
        oj.close()

    ts = [threading.Thread(target=run, args=(j,)) for j in range(threads)]
    for t in ts:
        t.start()
    ready.wait()
    a = time.perf_counter()
    for t in ts:
        t.join()
    wall = time.perf_counter() - a
    updn = sum(r[0] for r in res)
    desc = f"{W.name} workload ({W.m.n_cells} cells, ~{int(upd1 / k1)} particles)"
    return {"value": updn / wall, "unit": "particle-updates/s", "cores": threads, "kind": "port",
            "sample": f"{threads} independent simulations of the {desc} at once, {k1} full step(s) each "
                      f"(step = correct+move+project{', mesh move' if W.motion else ''}), {wall:.1f} s wall",
            "single_core": {"value": upd1 / t1, "cores": 1,
                            "sample": f"one simulation, {k1} full step(s), {t1:.1f} s"}}


def b_alg(W, n_particles, precision):
    """Compulsory HBM bytes per particle-update: particle state read and
    written once, nodal V, V_old and the projection (f64), vertex
    coordinates (twice on a moving mesh: read and written), the cell DoF
    map and the particle offsets."""
    from pfem2_amd import FP64
    m, d = W.m, W.m.dim
    wx, wv = (8, 8) if precision == FP64 else (4, 4)
    ndc = getattr(m, "dofs_per_cell", 1 << d)
    b = 2 * (8 + d * (wx + wv))
    b += m.n_dofs / n_particles * 3 * d * 8
    b += m.n_vertices / n_particles * d * 8 * (2 if W.motion is not None else 1)
    b += m.n_cells / n_particles * (ndc * 4 + 8)
    return b


def stage_roofline(W, r, precision):
    """The whole particle stage against HBM: B_alg x particle-updates/s."""
    b = b_alg(W, 0.5 * (r["particles_start"] + r["particles"]), precision)
    ach = b * r["updates_per_s"] / 1e9
    return {"bound": "hbm", "bytes_per_update": b, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": ach / HBM_PEAK_GBS}


def measured_traffic(kernel, profile):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 --pmc
    summary of the workload (profiles/hbm_traffic_<profile>.txt: FETCH_SIZE
    and WRITE_SIZE scaled by a plain copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_{profile}.txt")
    if not profile or not os.path.exists(path):
        return None
    for line in open(path):
        mm = re.match(r"\s*kernel=(\S+)\s+hbm_bytes_per_launch=(\d+)", line)
        if mm and mm.group(1) == kernel:
            return int(mm.group(2))
    return None


def measured_valu(kernel, profile):
    """VALU busy percentage of `kernel` from the committed rocprofv3 --pmc
    summary of the workload (profiles/valu_<profile>.txt), or None."""
    path = os.path.join(ROOT, "profiles", f"valu_{profile}.txt")
    if not profile or not os.path.exists(path):
        return None
    for line in open(path):
        mm = re.match(r"\s*kernel=(\S+)\s+valu_busy_pct=([0-9.]+)", line)
        if mm and mm.group(1) == kernel:
            return float(mm.group(2))
    return None


def roofline(res, m, profile=None):
    """Roofline of the kernel with the largest share of the step.  The FP64
    relocation is one launch (one round with refilling lanes); the FP32
    rounds are two kernels (the last round is an instantiation of its own)
    and stay apart.  A kernel launched more than once per step has its bytes
    and duration per launch averaged, as the profiler's per-kernel figures
    are."""
    per_step = res["kernels_ms_per_step"]
    fam = {}
    for key, v in per_step.items():
        f = re.sub(r"_r\d$", "", key) if key.startswith("k_search_resume_r") else key
        fam.setdefault(f, []).append(key)
    dom = max(fam, key=lambda f: sum(per_step[k] for k in fam[f]))
    members = fam[dom]
    vext = res.get("keep_vext", False)
    bs = [kernel_bytes(k, res["step_info"], m, m.dim, vext) for k in members]
    b = sum(x for x in bs if x) / len(members) if all(bs) else None
    dur = sum(res["kernels_ms_avg"][k] for k in members) / len(members) * 1e-3
    achieved = b / dur / 1e9 if b else None
    others = {}
    for key in per_step:
        kb = kernel_bytes(key, res["step_info"], m, m.dim, vext)
        if kb:
            others[key] = {"ms_avg": res["kernels_ms_avg"][key], "bytes_per_launch": kb,
                           "GBs": kb / (res["kernels_ms_avg"][key] * 1e-3) / 1e9,
                           "share_of_step": per_step[key] / sum(per_step.values())}
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None,
            "traffic": measured_traffic(dom, profile), "valu_busy_pct": measured_valu(dom, profile),
            "kernel": dom, "launches_per_step": len(members), "bytes_per_launch": b, "avg_launch_ms": dur * 1e3,
            "all_kernels": others}


def serve(args, rate=None, batched=None, batch_max_cells=None):
    """The request stream: independent simulations with Poisson arrivals,
    served by pf2_server (FP32-mixed, admission by footprint; batched: every
    step advances all active simulations with one launch per kernel, or,
    with batch_max_cells, those of at most that many cells).
    Moving-mesh requests (~400k quads around a cylinder oscillating with a
    seeded phase, 25 particles per cell) arrive within the first half of the
    stream, so that the memory in use climbs toward the budget."""
    from pfem2_amd import FP32_MIXED, Server
    from pfem2_amd import mesh as M
    rng = np.random.default_rng(SEEDS["serve"])
    reqs = []
    for i in range(args.serve_requests):
        nc = float(np.exp(rng.uniform(np.log(2e3), np.log(2e5))))
        ny = max(4, int(round(math.sqrt(nc / 5.0))))
        nx = max(4, int(round(nc / ny)))
        q = int(rng.choice([3, 4, 5]))
        umax = float(rng.uniform(0.2, 2.0))
        steps = int(rng.integers(50, 501))
        m = M.channel_2d(nx, ny, 5.0, 1.0)
        V = M.poiseuille_2d(m, umax, 1.0, 0.01, seed=1000 + i)
        dt = 0.5 * min(5.0 / nx, 1.0 / ny) / umax
        reqs.append({"m": m, "V": V, "steps": steps, "dt": dt, "q": q, "moving": None})
    rate = args.serve_rate if rate is None else rate
    batched = args.serve_batched if batched is None else batched
    batch_max_cells = args.serve_batch_max_cells if batch_max_cells is None else batch_max_cells
    arrivals = list(np.cumsum(rng.exponential(1.0 / rate, len(reqs))))
    if args.serve_moving > 0:
        W0 = moving_workload()
        Wd = np.zeros_like(W0.m.vertices)
        Wd[:, 1] = W0.blend
        span = arrivals[-1]
        for j in range(args.serve_moving):
            reqs.append({"m": W0.m, "V": W0.V, "steps": int(rng.integers(40, 101)), "dt": W0.dt, "q": 5,
                         "moving": (Wd, W0.amp, W0.omega, float(rng.uniform(0.0, 2.0 * math.pi)))})
            arrivals.append(float(rng.uniform(0.0, 0.5 * span)))
    order = np.argsort(arrivals, kind="stable")
    reqs = [reqs[k] for k in order]
    arrivals = [arrivals[k] for k in order]
    print(f"# serving {len(reqs)} requests ({args.serve_moving} moving) over ~{arrivals[-1]:.1f} s", file=sys.stderr,
          flush=True)
    srv = Server(hbm_budget=int(args.serve_budget_gib * (1 << 30)), use_graphs=True, batched=batched,
                 batch_max_cells=batch_max_cells)
    ids = []
    t0 = time.perf_counter()
    i = 0
    while True:
        now = time.perf_counter() - t0
        while i < len(reqs) and arrivals[i] <= now:
            r = reqs[i]
            if r["moving"] is None:
                ids.append(srv.submit(r["m"], r["V"], r["steps"], r["dt"], 10, 2, [r["q"], r["q"]], FP32_MIXED))
            else:
                Wd, amp, omega, phase = r["moving"]
                ids.append(srv.submit_moving(r["m"], r["V"], r["steps"], r["dt"], 10, 2, [r["q"], r["q"]], Wd, amp,
                                             omega, phase, FP32_MIXED))
            i += 1
        npend, nact = srv.poll()
        if i == len(reqs) and npend == 0 and nact == 0:
            break
        if nact == 0 and npend == 0:
            time.sleep(max(0.0, min(arrivals[i] - now, 0.001)))
    wall = time.perf_counter() - t0
    st = srv.stats()
    lat = np.array([srv.result(r)[2] for r in ids])
    srv.close()
    mov = np.array([r["moving"] is not None for r in reqs])
    cells = [r["m"].n_cells for r in reqs]
    out = {"requests": len(reqs), "moving_requests": int(mov.sum()), "finished": int(st["finished"]),
           "failed": int(st["failed"]), "p50_ms": st["p50_ms"], "p99_ms": st["p99_ms"], "mean_ms": st["mean_ms"],
           "max_ms": float(lat.max()), "updates_per_s": st["updates_per_s"],
           "particle_updates": st["particle_updates"], "wall_s": wall,
           "peak_hbm_gib": st["peak_bytes"] / 2 ** 30, "max_active": int(st["max_active"]),
           "arrival_rate_per_s": rate, "budget_gib": args.serve_budget_gib, "batched": bool(batched),
           "batch_max_cells": int(batch_max_cells) if batched else None,
           "cells_mean": float(np.mean(cells)), "precision": "fp32-mixed",
           "stream": f"{args.serve_requests} static requests, Poisson arrivals ({rate}/s), cells "
                     "log-uniform 2k-200k (channel quads), lattice 3/4/5, Umax U[0.2,2], 50-500 steps; "
                     f"{int(mov.sum())} moving-mesh requests (407k quads, 5x5 lattice, 40-100 steps, oscillating "
                     "cylinder) arriving in the first half; 10 substeps, MAX=2"}
    if mov.any():
        out["static_p50_ms"] = float(np.percentile(lat[~mov], 50))
        out["static_p99_ms"] = float(np.percentile(lat[~mov], 99))
        out["moving_p50_ms"] = float(np.percentile(lat[mov], 50))
        out["moving_p99_ms"] = float(np.percentile(lat[mov], 99))
    return out


def copy_calibration(torch, mib=512):
    """A plain device copy of known size (for scaling rocprofv3 FETCH/WRITE counters)."""
    n = mib * 1024 * 1024 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each_way": n * 4}))


def base_line(args, W, r, dtype):
    return {
        "metric": "particle-updates/s (advect+project+correct) and p50/p99 sim latency",
        "value": r["updates_per_s"],
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": dtype,
        "data": "synthetic",
        "config": {"workload": f"{W.name}: {W.desc}", "cells": W.m.n_cells, "particles": r["particles"], "dt": W.dt,
                   "substeps": W.substeps, "max_per_part": W.max_per_part, "seeds": SEEDS},
        "p50_ms": r["p50_ms"],
        "p99_ms": r["p99_ms"],
        # the mesh vertices uploaded from host memory every step instead of
        # formed in HBM (moving mesh only): PCIe-inclusive, never the value
        "ms_per_step_host_vertices": r["ms_per_step_host_vertices"],
    }


def only_main(args, W, torch):
    """One precision path alone (for profiler runs)."""
    from pfem2_amd import FP32_MIXED, FP64
    r = run_path(args, W, FP64 if args.only == "fp64" else FP32_MIXED, torch)
    print(f"# {W.name} {args.only} {r['updates_per_s']:.3e} updates/s, {r['ms_per_step']:.3f} ms/step",
          file=sys.stderr, flush=True)
    return {"path": args.only, "workload": W.name, **r, "roofline": roofline(r, W.m)}


def path_summary(W, r, precision):
    return {"updates_per_s": r["updates_per_s"], "ms_per_step": r["ms_per_step"],
            "ms_per_step_eager": r["ms_per_step_eager"], "ms_per_step_host_vertices": r["ms_per_step_host_vertices"],
            "p50_ms": r["p50_ms"], "p99_ms": r["p99_ms"],
            "particles_start": r["particles_start"], "particles_end": r["particles"],
            "kernels_ms_per_step": r["kernels_ms_per_step"], "roofline": roofline(r, W.m),
            "stage_roofline": stage_roofline(W, r, precision), "stats": r["stats"], "step_info": r["step_info"],
            "memory_gib": r["memory_bytes"] / 2 ** 30}


def run_both(args, W, torch):
    from pfem2_amd import FP32_MIXED, FP64
    print(f"# {W.name}: mesh {W.m.n_cells} cells {W.m.n_vertices} vertices", file=sys.stderr, flush=True)
    r64 = run_path(args, W, FP64, torch)
    print(f"# {W.name} fp64 {r64['updates_per_s']:.3e} updates/s, {r64['ms_per_step']:.3f} ms/step "
          f"(eager {r64['ms_per_step_eager']:.3f})", file=sys.stderr, flush=True)
    r32 = run_path(args, W, FP32_MIXED, torch)
    print(f"# {W.name} fp32-mixed {r32['updates_per_s']:.3e} updates/s, {r32['ms_per_step']:.3f} ms/step "
          f"(eager {r32['ms_per_step_eager']:.3f})", file=sys.stderr, flush=True)
    return r64, r32


FP32_TARGETS = {"error_per_field": 1e-5, "speedup_over_fp64": 3.0, "hbm_frac": 0.6}


def fp32_targets(acc, speedup, frac):
    """The FP32-mixed path's contract, met or not, per item."""
    out = {"speedup_over_fp64": {"value": speedup, "target": FP32_TARGETS["speedup_over_fp64"],
                                 "met": speedup >= FP32_TARGETS["speedup_over_fp64"]},
           "stage_hbm_frac": {"value": frac, "target": FP32_TARGETS["hbm_frac"],
                              "met": frac >= FP32_TARGETS["hbm_frac"]}}
    if acc:
        for k in ("V_rel_l2", "V_max_rel", "x_max_rel", "v_max_rel"):
            out[k] = {"value": acc[k], "target": FP32_TARGETS["error_per_field"],
                      "met": acc[k] <= FP32_TARGETS["error_per_field"]}
    return out


def workload_main(args, W, torch, with_cpu=True, with_accuracy=True):
    """Both precision paths of one workload; the FP64 path (the reference's
    arithmetic, bitwise equal to the oracle) is the line's value."""
    from pfem2_amd import FP32_MIXED, FP64
    r64, r32 = run_both(args, W, torch)
    acc = accuracy(W, args.accuracy_steps) if with_accuracy and args.accuracy_steps > 0 else None
    if acc:
        print(f"# fp32-mixed vs fp64: {acc}", file=sys.stderr, flush=True)
    cpu = cpu_baseline(W, args.cpu_seconds) if with_cpu and not args.no_cpu else None
    out = base_line(args, W, r64, "f64")
    out["roofline"] = roofline(r64, W.m, profile=PROFILES.get((W.name, FP64)))
    out["stage_roofline"] = stage_roofline(W, r64, FP64)
    out["cpu_baseline"] = cpu
    s32 = path_summary(W, r32, FP32_MIXED)
    s32["error_vs_fp64"] = acc
    speed = r32["updates_per_s"] / r64["updates_per_s"]
    s32["targets"] = fp32_targets(acc, speed, s32["stage_roofline"]["frac"])
    out["paths"] = {"fp64": path_summary(W, r64, FP64), "fp32_mixed": s32, "speedup_fp32_over_fp64": speed}
    return out


def channel_main(args, torch):
    """200x20 channel, FP64, end to end: every step runs the particle stage
    on the GPU, then the host FEM stand-in (fractional step with a pressure
    solve) turns the projected field into the next V.  The CPU oracle runs
    the same sequence; the final fields are compared bit for bit.  Particle
    stage and host FEM are timed separately."""
    from pfem2_amd import FP64
    from pfem2_amd.fem import ChannelFEM
    W = channel_workload()
    steps = args.channel_steps
    fem_g = ChannelFEM(W.m, W.dt)
    st = make_stage(W, FP64)
    n0 = st.n_global_particles()
    gpu_s = fem_s = 0.0
    upd = 0
    for _ in range(steps):
        a = time.perf_counter()
        Vp = st.step(download=True)
        b = time.perf_counter()
        V = fem_g.step(Vp)
        c = time.perf_counter()
        st.upload_fields(V, None)
        gpu_s += (b - a) + (time.perf_counter() - c)
        fem_s += c - b
    Vg = Vp
    n1 = st.n_global_particles()
    st.close()
    fem_o = ChannelFEM(W.m, W.dt)
    o = _oracle(W)
    cpu_s = 0.0
    Vo = None
    for _ in range(steps):
        upd += o.n_global_particles()
        cpu_s += o.time_steps(1)
        Vo = o.field()          # the projection of this step
        o.upload_fields(fem_o.step(Vo), None)
    gpu_rate = 0.5 * (n0 + n1) * steps / gpu_s
    out = base_line(args, W, {"updates_per_s": gpu_rate, "ms_per_step": gpu_s * 1e3 / steps, "particles": n1,
                              "p50_ms": None, "p99_ms": None}, "f64")
    out["config"]["workload"] += "; host FEM stand-in (fractional step, pressure solve) between particle steps"
    out["steps"] = steps
    out["warmup"] = 0
    out["roofline"] = None
    out["host_fem_ms_per_step"] = fem_s * 1e3 / steps
    out["cpu_baseline"] = {"value": upd / cpu_s, "unit": "particle-updates/s", "cores": 1, "kind": "port",
                           "sample": f"the whole run: {steps} steps of {W.m.n_cells} cells, {cpu_s:.1f} s"}
    out["parity"] = {"final_field_bitwise_equal": bool(np.array_equal(Vg, Vo, equal_nan=True)),
                     "particles_gpu": int(n1), "particles_oracle": int(o.n_global_particles())}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--scale", type=float, default=1.0)
    ap.add_argument("--dt", type=float, default=0.001)
    ap.add_argument("--substeps", type=int, default=10)
    ap.add_argument("--max-per-part", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--accuracy-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--workload", choices=["default", "cylinder", "channel", "pipe3d", "moving", "serve",
                                           "serve-sweep"], default="default")
    ap.add_argument("--serve-requests", type=int, default=256)
    ap.add_argument("--serve-rate", type=float, default=10.0, help="Poisson arrival rate, requests/s")
    ap.add_argument("--serve-budget-gib", type=float, default=200.0)
    ap.add_argument("--serve-moving", type=int, default=8, help="moving-mesh requests mixed into the stream")
    ap.add_argument("--serve-batch-max-cells", type=int, default=0,
                    help="batched serving: batch only simulations of at most this many cells (0: all)")
    ap.add_argument("--serve-batched", action="store_true",
                    help="serve with batched steps (pf2_batch_step) instead of one step graph per simulation")
    ap.add_argument("--serve-rates", default="5,10,20,40,80,160",
                    help="arrival rates of --workload serve-sweep (requests/s)")
    ap.add_argument("--channel-steps", type=int, default=100)
    ap.add_argument("--copy-calibration", action="store_true")
    ap.add_argument("--only", choices=["fp64", "fp32"], default=None,
                    help="run one precision path of the workload and print its measurements only")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    if args.copy_calibration:
        copy_calibration(torch)
        return
    if args.workload == "serve-sweep":
        # p50/p99 latency and throughput against the offered load, batched
        # and per-simulation serving
        rows = []
        for rate in [float(x) for x in args.serve_rates.split(",")]:
            for bt, bmax in ((False, 0), (True, 0), (True, 16384)):
                r = serve(args, rate=rate, batched=bt, batch_max_cells=bmax)
                print(f"# rate {rate}/s batched={bt} (max cells {bmax}): p50 {r['p50_ms']:.1f} ms "
                      f"p99 {r['p99_ms']:.1f} ms "
                      f"{r['updates_per_s']:.3e} updates/s, max active {r['max_active']}", file=sys.stderr, flush=True)
                rows.append({k: r[k] for k in ("arrival_rate_per_s", "batched", "batch_max_cells", "p50_ms", "p99_ms",
                                               "updates_per_s",
                                               "max_active", "peak_hbm_gib", "finished", "failed", "wall_s")})
        out = {"metric": "p50/p99 sim latency against offered load", "value": None, "unit": "ms",
               "n_gpus": 1, "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": DTYPE32,
               "data": "synthetic", "config": {"workload": rows and r["stream"]}, "sweep": rows}
        print(json.dumps(out))
        return
    if args.workload == "serve":
        s = serve(args)
        out = {"metric": "p50/p99 sim latency and particle-updates/s of a request stream", "value": s["p50_ms"],
               "unit": "ms (p50 latency)", "n_gpus": 1, "higher_is_better": False, "scaling": "weak",
               "vs_baseline": None, "dtype": DTYPE32, "data": "synthetic", "config": {"workload": s["stream"]},
               "serving": s}
        print(json.dumps(out))
        return
    if args.workload == "channel":
        print(json.dumps(channel_main(args, torch)))
        return
    if args.only:
        W = {"cylinder": lambda: cylinder_workload(args), "pipe3d": pipe_workload, "moving": moving_workload}[
            args.workload]()
        print(json.dumps(only_main(args, W, torch)))
        return
    if args.workload == "pipe3d":
        print(json.dumps(workload_main(args, pipe_workload(), torch, with_accuracy=False)))
        return
    if args.workload == "cylinder":
        print(json.dumps(workload_main(args, cylinder_workload(args), torch)))
        return
    # default: the largest single-GPU configuration (the moving-mesh cylinder,
    # ~400k quads, ~10M particles) with the FP64 path as the value, then the
    # 100k-cell cylinder channel in both precisions with the FP32-mixed
    # path's error against FP64, then the request stream
    out = workload_main(args, moving_workload(), torch, with_accuracy=False)
    if args.workload == "moving":
        print(json.dumps(out))
        return
    cyl = workload_main(args, cylinder_workload(args), torch, with_cpu=False)
    out["cylinder100k"] = {"workload": cyl["config"]["workload"], "cells": cyl["config"]["cells"], **cyl["paths"]}
    if not args.no_serve:
        s = serve(args)
        print(f"# serving: p50 {s['p50_ms']:.1f} ms p99 {s['p99_ms']:.1f} ms, {s['updates_per_s']:.3e} updates/s, "
              f"max active {s['max_active']}", file=sys.stderr, flush=True)
        out["serving"] = s
        out["sim_latency_p50_ms"] = s["p50_ms"]
        out["sim_latency_p99_ms"] = s["p99_ms"]
        sb = serve(args, batched=True)
        print(f"# serving, batched: p50 {sb['p50_ms']:.1f} ms p99 {sb['p99_ms']:.1f} ms, "
              f"{sb['updates_per_s']:.3e} updates/s", file=sys.stderr, flush=True)
        out["serving_batched"] = sb
    print(json.dumps(out))


if __name__ == "__main__":
    main()
