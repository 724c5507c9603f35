"""Benchmark of the PFEM2 particle stage on one MI355X.

Workload (the metric's configuration that fits one GPU): 2-D channel with a
circular obstacle (Turek-Schaefer geometry 2.2 x 0.41, cylinder r = 0.05),
graded quad mesh of ~100k cells with seeded vertex jitter, 16 particles per
cell (4x4 lattice, +-0.05 seeded jitter, ~1.6M particles), velocity =
potential flow around the cylinder + 1% seeded Fourier perturbation, dt =
0.001, 10 advection substeps, MAX = 2 particles per cell part.

One step = correct -> move (10 substeps of advection + relocation, then
reseed/delete) -> project, for every particle, plus the FEM stand-in (the new
nodal V, a device-resident analytic field, copied into the stage).  All
inputs are resident in HBM before the timed region.

Prints one JSON line (see the "paths" entry for the FP64 path and the
FP32-mixed path; `value` is the FP32-mixed path's whole-job throughput).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

SEEDS = {"mesh": 2024, "field": 7, "particles": 11}
HBM_PEAK_GBS = 8000.0


def kernel_bytes(kind, dim, n_p, m):
    """Algorithmic HBM bytes of one launch of the dominant kernel."""
    nc, nd = m.n_cells, m.n_dofs
    nv = 1 << dim
    if kind == "k_move":
        # read cell + x + xi + v, write x + xi + v + dest + key (FP64 state)
        per_p = 4 + 3 * 8 * dim + 3 * 8 * dim + 8
        mesh_b = nd * 2 * 8 * dim + m.n_vertices * 8 * dim + nc * nv * 8
    else:
        # read cell + xi + v, write xi + v + dest + key (FP32 state)
        per_p = 4 + 2 * 4 * dim + 2 * 4 * dim + 8
        mesh_b = nd * 2 * 4 * dim + nc * (nv * 4 + (6 if dim == 2 else 21) * 4)
    return per_p * n_p + mesh_b, per_p


def run_path(args, m, V, precision, torch):
    from pfem2_amd import ParticleStage
    from pfem2_amd import mesh as M
    st = ParticleStage(m, args.dt, args.substeps, args.max_per_part, [4, 4], precision)
    st.upload_fields(V, V)
    cell, xi = M.jittered_lattice(m, [4, 4], 0.05, SEEDS["particles"])
    st.insert_particles(cell, xi)
    n0 = st.n_global_particles()
    dV = torch.from_numpy(np.ascontiguousarray(V)).to("cuda")   # FEM stand-in output, resident
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
    # per-step latency (synchronised steps)
    lat = []
    for _ in range(min(args.steps, 20)):
        a = time.perf_counter()
        step()
        st.synchronize()
        lat.append((time.perf_counter() - a) * 1e3)
    # per-kernel HIP-event times on the stage's stream
    st.set_profiling(True)
    for _ in range(args.steps):
        step()
    kt = st.kernel_times()
    st.set_profiling(False)
    n2 = st.n_global_particles()
    res = {
        "updates_per_s": updates / (ms * 1e-3),
        "ms_per_step": ms,
        "particles": int(n1),
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "kernels_ms_avg": {k: v[0] / max(v[1], 1) for k, v in kt.items()},
        "kernel_launches": {k: v[1] for k, v in kt.items()},
        "particles_profiled": int(n2),
        "stats": st.stats(),
    }
    st.close()
    return res


def cpu_baseline(m, V, args, seconds=10.0):
    """Oracle (CPU restatement, 1 thread) on a bounded sample of the workload:
    the same mesh and particles, stepped until about `seconds` of CPU time."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_binding import Oracle
    from pfem2_amd import mesh as M
    o = Oracle(m, args.dt, args.substeps, args.max_per_part, [4, 4])
    o.upload_fields(V, V)
    cell, xi = M.jittered_lattice(m, [4, 4], 0.05, SEEDS["particles"])
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
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (every box has one GPU); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    from pfem2_amd import FP32_MIXED, FP64
    from pfem2_amd import mesh as M
    m = M.cylinder_channel(args.scale, seed=SEEDS["mesh"])
    V = M.cylinder_flow(m, 1.0, eps=0.01, seed=SEEDS["field"])
    print(f"# mesh {m.n_cells} cells {m.n_vertices} vertices", file=sys.stderr, flush=True)
    r64 = run_path(args, m, V, FP64, torch)
    print(f"# fp64 {r64['updates_per_s']:.3e} updates/s, {r64['ms_per_step']:.3f} ms/step", file=sys.stderr,
          flush=True)
    r32 = run_path(args, m, V, FP32_MIXED, torch)
    print(f"# fp32-mixed {r32['updates_per_s']:.3e} updates/s, {r32['ms_per_step']:.3f} ms/step", file=sys.stderr,
          flush=True)
    # roofline of the dominant kernel of the reported path
    km = max(r32["kernels_ms_avg"], key=lambda k: r32["kernels_ms_avg"][k] * r32["kernel_launches"][k])
    kname = "k_move32" if "k_move32" in r32["kernels_ms_avg"] else km
    bytes_launch, per_p = kernel_bytes("k_move32", 2, r32["particles_profiled"], m)
    dur = r32["kernels_ms_avg"][kname] * 1e-3
    achieved = bytes_launch / dur / 1e9
    roof = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": kname,
            "bytes_per_launch": bytes_launch, "bytes_per_particle": per_p}
    b64, pp64 = kernel_bytes("k_move", 2, r64["particles_profiled"], m)
    d64 = r64["kernels_ms_avg"]["k_move"] * 1e-3
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
        "dtype": "fp32-mixed (f32 reference coords/velocities, f64 relocation)",
        "data": "synthetic",
        "config": {"workload": "2D channel with circular obstacle (2.2x0.41, r=0.05), graded quad mesh, "
                               "16 particles/cell, potential flow + 1% Fourier perturbation",
                   "cells": m.n_cells, "particles": r32["particles"], "dt": args.dt, "substeps": args.substeps,
                   "max_per_part": args.max_per_part, "seeds": SEEDS},
        "p50_ms": r32["p50_ms"],
        "p99_ms": r32["p99_ms"],
        "roofline": roof,
        "cpu_baseline": cpu,
        "paths": {
            "fp64": {"updates_per_s": r64["updates_per_s"], "ms_per_step": r64["ms_per_step"],
                     "p50_ms": r64["p50_ms"], "p99_ms": r64["p99_ms"], "kernels_ms_avg": r64["kernels_ms_avg"],
                     "k_move_GBs": b64 / d64 / 1e9, "k_move_bytes_per_particle": pp64, "stats": r64["stats"]},
            "fp32_mixed": {"updates_per_s": r32["updates_per_s"], "ms_per_step": r32["ms_per_step"],
                           "kernels_ms_avg": r32["kernels_ms_avg"], "stats": r32["stats"]},
            "speedup_fp32_over_fp64": r32["updates_per_s"] / r64["updates_per_s"],
        },
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
