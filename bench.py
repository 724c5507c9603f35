"""Benchmark of the PFEM2 particle stage on one MI355X.

Workload (the metric's configuration that fits one GPU): 2-D channel with a
circular obstacle (Turek-Schaefer geometry 2.2 x 0.41, cylinder r = 0.05),
graded quad mesh of ~100k cells with seeded vertex jitter, 16 particles per
cell (4x4 lattice, +-0.05 seeded jitter, ~1.6M particles at the start),
velocity = potential flow around the cylinder + 1% seeded Fourier
perturbation, dt = 0.001, 10 advection substeps, MAX = 2 particles per cell
part.

One step = correct -> move (10 substeps of advection + relocation, then
reseed/delete) -> project, for every particle, plus the FEM stand-in (the new
nodal V, a device-resident analytic field, copied into the stage).  All
inputs are resident in HBM before the timed region.

Prints one JSON line.  `value` is the FP32-mixed path's whole-job
throughput; "paths" holds the FP64 path (bitwise equal to the CPU oracle)
and the per-field error of the FP32-mixed path against it.
"""
import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SEEDS = {"mesh": 2024, "field": 7, "particles": 11}
HBM_PEAK_GBS = 8000.0
Q = [4, 4]


def kernel_bytes(name, info, m, dim=2):
    """Algorithmic HBM bytes of one launch of a step kernel, from the work
    counts of the step (compulsory traffic: every array element touched once)."""
    nc, nd = m.n_cells, m.n_dofs
    nv = 1 << dim
    n0, n1 = info["n_before"], info["n_after"]
    q = info["queued"]
    arr = max(info["arrivals"], 0)
    f64 = not name.endswith("32") and name not in ("k_group", "k_cell_count")
    w_xi = 8 if (f64 or name == "k_cell_count") else 4
    if name in ("k_move", "k_move32"):
        wx, wv = (8, 8) if name == "k_move" else (4, 4)
        reads = n0 * (4 + dim * wx * (3 if name == "k_move" else 1) + dim * wv)
        writes = n0 * dim * wv + (n0 - q[0]) * (dim * wx * (2 if name == "k_move" else 1) + 8) \
            + q[0] * (dim * wx * (2 if name == "k_move" else 1) + 4 + 4 + 4 + 8 * dim)
        mesh = nc * nv * 4 + nd * dim * wv * 2 + (m.n_vertices * dim * 8 if name == "k_move" else nc * 24)
        return reads + writes + mesh
    if name in ("k_search_resume", "k_search_resume32"):
        wx = 8 if name == "k_search_resume" else 4
        per = 4 + 4 + 4 + 4 + dim * wx * (2 if name == "k_search_resume" else 1) + 8 * dim + 8 + dim * wx
        return int(sum(q) * per / max(len(q), 1))
    if name == "k_group":
        return n0 * 8 + arr * 8 + nc * 4
    if name == "k_cell_count":
        return (n0 + arr) * 12 + arr * 4 + n0 * 4 + nc * (8 + 4 + 4 + 20)
    if name in ("k_cell_write", "k_cell_write32"):
        wx = 8 if name == "k_cell_write" else 4
        per_src = 4 + 4 + 4 + dim * wx * (3 if name == "k_cell_write" else 2)
        per_dst = 4 + 4 + dim * wx * (3 if name == "k_cell_write" else 2)
        return n1 * (per_src + per_dst) + nc * 32
    if name in ("k_cell_partials", "k_cell_partials32"):
        wx = 8 if name == "k_cell_partials" else 4
        return n1 * dim * wx * 2 + nc * (4 + nv * (dim + 1) * 8)
    if name == "k_node":
        return nc * nv * (dim + 1) * 8 + nd * (dim * 8 * 2 + 4) + nc * nv * 4
    return None


def make_stage(m, V, precision, args):
    from pfem2_amd import ParticleStage
    from pfem2_amd import mesh as M
    st = ParticleStage(m, args.dt, args.substeps, args.max_per_part, Q, precision)
    st.upload_fields(V, V)
    cell, xi = M.jittered_lattice(m, Q, 0.05, SEEDS["particles"])
    st.insert_particles(cell, xi)
    return st


def run_path(args, m, V, precision, torch):
    st = make_stage(m, V, precision, args)
    n0 = st.n_global_particles()
    dV = torch.from_numpy(np.ascontiguousarray(V)).to("cuda")   # FEM stand-in output, resident in HBM
    torch.cuda.synchronize()
    ptr = dV.data_ptr()
    lib = st._lib

    def step():
        rc = lib.pf2_upload_fields_device(st._h, ptr, None)
        if rc:
            st._check(rc)
        rc = lib.pf2_step(st._h, None)
        if rc:
            st._check(rc)

    for _ in range(args.warmup):
        step()
    st.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    st.synchronize()
    t1 = time.perf_counter()
    n1 = st.n_global_particles()
    ms = (t1 - t0) * 1e3 / args.steps
    updates = 0.5 * (n0 + n1)
    # per-step latency of synchronised steps
    lat = []
    for _ in range(max(args.steps, 20)):
        a = time.perf_counter()
        step()
        st.synchronize()
        lat.append((time.perf_counter() - a) * 1e3)
    # per-kernel HIP-event times on the stage's stream
    st.set_profiling(True)
    infos = []
    for _ in range(args.steps):
        step()
        infos.append(st.step_info())
    kt = st.kernel_times()
    st.set_profiling(False)
    info = {k: (np.mean([i[k] for i in infos], axis=0).tolist() if k == "queued" else
                float(np.mean([i[k] for i in infos]))) for k in infos[0]}
    res = {
        "updates_per_s": updates / (ms * 1e-3),
        "ms_per_step": ms,
        "particles": int(n1),
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "kernels_ms_avg": {k: v[0] / max(v[1], 1) for k, v in kt.items()},
        "kernels_ms_per_step": {k: v[0] / args.steps for k, v in kt.items()},
        "kernel_launches_per_step": {k: v[1] / args.steps for k, v in kt.items()},
        "step_info": info,
        "stats": st.stats(),
    }
    st.close()
    return res


def accuracy(args, m, V, steps):
    """Per-field error of the FP32-mixed path against the FP64 path (which is
    bitwise equal to the CPU oracle) after `steps` steps."""
    from pfem2_amd import FP32_MIXED, FP64
    a = make_stage(m, V, FP64, args)
    b = make_stage(m, V, FP32_MIXED, args)
    n0 = a.n_global_particles()
    for _ in range(steps):
        Va = a.step(download=True)
        Vb = b.step(download=True)
        a.upload_fields(V, None)
        b.upload_fields(V, None)
    pa, pb = a.particles(), b.particles()
    ia = {int(k): n for n, k in enumerate(pa["id"]) if k <= n0}
    pairs = [(ia[int(k)], n) for n, k in enumerate(pb["id"]) if k <= n0 and int(k) in ia]
    ai = np.array([p[0] for p in pairs])
    bi = np.array([p[1] for p in pairs])
    same = pa["cell"][ai] == pb["cell"][bi]
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


def cpu_baseline(m, V, args, seconds=10.0):
    """Oracle (CPU restatement, 1 thread) on a bounded sample of the workload:
    the same mesh and particles, stepped until about `seconds` of CPU time."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_binding import Oracle
    from pfem2_amd import mesh as M
    o = Oracle(m, args.dt, args.substeps, args.max_per_part, Q)
    o.upload_fields(V, V)
    cell, xi = M.jittered_lattice(m, Q, 0.05, SEEDS["particles"])
    o.insert_particles(cell, xi)
    done, t, upd = 0, 0.0, 0.0
    while t < seconds and done < 50:
        n = o.n_global_particles()
        dt = o.time_steps(1)
        o.upload_fields(V, None)
        t += dt
        upd += n
        done += 1
    return {"value": upd / t, "unit": "particle-updates/s", "cores": 1, "kind": "port",
            "sample": f"{done} full steps of the same workload ({m.n_cells} cells, ~{int(upd / done)} particles), "
                      f"{t:.1f} s single-threaded"}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` for this workload from the committed
    rocprofv3 --pmc summary (FETCH_SIZE and WRITE_SIZE scaled by a plain copy
    kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_cylinder100k.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        mm = re.match(r"\s*kernel=(\S+)\s+hbm_bytes_per_launch=(\d+)", line)
        if mm and mm.group(1) == kernel:
            return int(mm.group(2))
    return None


def roofline(res, m):
    per_step = res["kernels_ms_per_step"]
    fam = {}
    for k, v in per_step.items():
        fam[k] = v
    dom = max(fam, key=lambda k: fam[k])
    b = kernel_bytes(dom, res["step_info"], m)
    dur = res["kernels_ms_avg"][dom] * 1e-3
    achieved = b / dur / 1e9 if b else None
    others = {}
    for k in per_step:
        kb = kernel_bytes(k, res["step_info"], m)
        if kb:
            others[k] = {"ms_avg": res["kernels_ms_avg"][k], "bytes_per_launch": kb,
                         "GBs": kb / (res["kernels_ms_avg"][k] * 1e-3) / 1e9,
                         "share_of_step": per_step[k] / sum(per_step.values())}
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS if achieved else None, "traffic": measured_traffic(dom),
            "kernel": dom, "bytes_per_launch": b, "avg_launch_ms": res["kernels_ms_avg"][dom],
            "all_kernels": others}


def copy_calibration(torch, mib=512):
    """A plain device copy of known size (for scaling rocprofv3 FETCH/WRITE counters)."""
    n = mib * 1024 * 1024 // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_each_way": n * 4}))


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
    ap.add_argument("--only", choices=["fp64", "fp32"], default=None)
    ap.add_argument("--copy-calibration", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    if args.copy_calibration:
        copy_calibration(torch)
        return
    from pfem2_amd import FP32_MIXED, FP64
    from pfem2_amd import mesh as M
    m = M.cylinder_channel(args.scale, seed=SEEDS["mesh"])
    V = M.cylinder_flow(m, 1.0, eps=0.01, seed=SEEDS["field"])
    print(f"# mesh {m.n_cells} cells {m.n_vertices} vertices", file=sys.stderr, flush=True)
    r64 = r32 = None
    if args.only in (None, "fp64"):
        r64 = run_path(args, m, V, FP64, torch)
        print(f"# fp64 {r64['updates_per_s']:.3e} updates/s, {r64['ms_per_step']:.3f} ms/step", file=sys.stderr,
              flush=True)
    if args.only in (None, "fp32"):
        r32 = run_path(args, m, V, FP32_MIXED, torch)
        print(f"# fp32-mixed {r32['updates_per_s']:.3e} updates/s, {r32['ms_per_step']:.3f} ms/step",
              file=sys.stderr, flush=True)
    if args.only:
        r = r64 or r32
        print(json.dumps({"path": args.only, **r, "roofline": roofline(r, m)}))
        return
    acc = accuracy(args, m, V, args.accuracy_steps) if args.accuracy_steps > 0 else None
    print(f"# fp32-mixed vs fp64: {acc}", file=sys.stderr, flush=True)
    cpu = None if args.no_cpu else cpu_baseline(m, V, args, args.cpu_seconds)
    out = {
        "metric": "particle-updates/s (advect+project+correct) and p50/p99 sim latency",
        "value": r32["updates_per_s"],
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r32["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32-mixed (f32 reference coords/velocities, f64 relocation and node sums)",
        "data": "synthetic",
        "config": {"workload": "2D channel with circular obstacle (2.2x0.41, r=0.05), graded quad mesh, "
                               "16 particles/cell, potential flow + 1% Fourier perturbation, "
                               "step = correct+move(10 substeps)+project",
                   "cells": m.n_cells, "particles": r32["particles"], "dt": args.dt, "substeps": args.substeps,
                   "max_per_part": args.max_per_part, "seeds": SEEDS},
        "p50_ms": r32["p50_ms"],
        "p99_ms": r32["p99_ms"],
        "roofline": roofline(r32, m),
        "cpu_baseline": cpu,
        "paths": {
            "fp64": {"updates_per_s": r64["updates_per_s"], "ms_per_step": r64["ms_per_step"],
                     "p50_ms": r64["p50_ms"], "p99_ms": r64["p99_ms"],
                     "kernels_ms_per_step": r64["kernels_ms_per_step"], "roofline": roofline(r64, m),
                     "stats": r64["stats"], "step_info": r64["step_info"]},
            "fp32_mixed": {"updates_per_s": r32["updates_per_s"], "ms_per_step": r32["ms_per_step"],
                           "kernels_ms_per_step": r32["kernels_ms_per_step"], "stats": r32["stats"],
                           "step_info": r32["step_info"], "error_vs_fp64": acc},
            "speedup_fp32_over_fp64": r32["updates_per_s"] / r64["updates_per_s"],
        },
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
