#!/usr/bin/env python3
"""Benchmark of the MOG force step on one MI355X.

Workload: two colliding exponential disks, n = 1e6 (5e5 each, scale radius
8, sech^2 height 0.5, centres 40 apart, relative speed 5, disks inclined 30
degrees), k = 240 grids of g = 16, W = 100, G = 100, dt = 1e-5, damping
0.999, float32 state, seeded synthetic data.  A step is one full force step
(clear, deposit, spectral solve, gather + update) on particles resident in
HBM; K steps are replayed from a hipGraph between two synchronisations.

Printed (one JSON line): steps/s (value), particle-grid interactions/s,
the relative RMS force error of the GPU against the FP64 oracle on a grid
sample, per-phase HIP-event times, the roofline of the dominant kernel and
the CPU baseline (the NumPy FP64 oracle timed on this host on a sample).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "steps/s & particle-grid ops/s at n=1e6,k=240,g=16 at fixed force error"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--k", type=int, default=240)
    ap.add_argument("--g", type=int, default=16)
    ap.add_argument("--compat", default="reference", choices=["reference", "physical"])
    ap.add_argument("--workload", default="two_disks", choices=["two_disks", "plummer", "uniform"],
                    help="two_disks: n=1e6 colliding disks (k=240, g=16); plummer: 64 Plummer "
                         "spheres, a=1.5 (use with --n 4000000 --g 32); uniform: U[0,W)^3")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "bf16x3", "fp16"])
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--cpu-grids", type=int, default=None,
                    help="grids of the oracle's timed CPU sample (all n particles)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--profile-reps", type=int, default=20)
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default tuning instead of autotuning first")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    from mog_solver_amd.config import SolverConfig
    from mog_solver_amd.solver import MogSolver
    from mog_solver_amd import workloads
    from mog_solver_amd.transforms import compute_transforms

    W, G, dt, damping = 100.0, 100.0, 1e-5, 0.999
    n, k, g = args.n, args.k, args.g
    cfg_d = dict(n_grids=k, grid_size=g, world_size=W, G=G, dt=dt, damping=damping,
                 solver="fft", compat=args.compat, precision=args.precision)
    cfg = SolverConfig.from_dict(cfg_d)
    off, rot = compute_transforms(k, g, W, seed=0)
    if args.workload == "two_disks":
        pos_h, vel_h = workloads.two_disks(n, W, G, g, seed=2)
        wl = "two colliding exponential disks"
    elif args.workload == "plummer":
        pos_h, vel_h = workloads.plummer_spheres(n, n_spheres=64, scale=1.5, world_size=W, seed=3)
        wl = "64 Plummer spheres (a=1.5)"
    else:
        pos_h, vel_h = workloads.uniform_cube(n, W, seed=1)
        wl = "uniform cube"
    print(f"[bench] workload ready n={n} k={k} g={g}", file=sys.stderr, flush=True)

    solver = MogSolver(cfg, rot=rot, off=off)
    pos = torch.from_numpy(pos_h).cuda()
    vel = torch.from_numpy(vel_h).cuda()
    torch.cuda.synchronize()

    # force error against the FP64 oracle on a grid sample (first cpu_grids
    # grids, every particle), from the initial state
    if args.cpu_grids is None:  # about 10 s of single-core oracle work on this host
        args.cpu_grids = max(1, int(48 * 1_000_000 / n))
    kk = max(1, min(args.cpu_grids, k))
    sample_cfg = SolverConfig.from_dict(dict(cfg_d, n_grids=kk))
    sample = MogSolver(sample_cfg, rot=rot[:kk], off=off[:kk])
    gpu_avg = sample.forces(pos).cpu().numpy()[:, :3].astype(np.float64)
    sample.close()

    tuned = None
    if not args.no_autotune:  # outside the timed region, on copies of the state
        from mog_solver_amd.autotune import autotune

        tuned = autotune(solver, pos, vel, cache_path=None)
        print(f"[bench] autotuned ({tuned['evaluated']} candidates): {tuned['tuning']}",
              file=sys.stderr, flush=True)

    use_graph = not args.no_graph
    solver.run(pos, vel, args.warmup, use_graph=use_graph)
    torch.cuda.synchronize()
    print(f"[bench] warmup done ({args.warmup} steps)", file=sys.stderr, flush=True)
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    start.record()
    solver.run(pos, vel, args.steps, use_graph=use_graph)
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms = start.elapsed_time(stop)
    steps_per_s = args.steps / (ms / 1e3)
    print(f"[bench] timed {args.steps} steps: {ms:.2f} ms (wall {wall*1e3:.2f} ms)", file=sys.stderr, flush=True)

    # per-phase HIP-event times on the launch stream
    reps = max(1, args.profile_reps)
    acc = {}
    for _ in range(reps):
        ph = solver.step_profiled(pos, vel)
        for key, v in ph.items():
            acc[key] = acc.get(key, 0.0) + v
    phases = {key: v / reps for key, v in acc.items()}
    cells = k * g**3
    alg_bytes = {
        "deposit": 16.0 * n + 8.0 * cells,
        "gather_update": 64.0 * n + 16.0 * cells,
        "solve": 24.0 * cells,
        "clear": 8.0 * cells,
    }
    dom = max(("deposit", "gather_update", "solve"), key=lambda key: phases[key])
    achieved = alg_bytes[dom] / (phases[dom] / 1e3) / 1e9
    roofline = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": None,
        "alg_bytes_per_launch": alg_bytes[dom],
        "avg_launch_ms": round(phases[dom], 4),
    }

    cpu = None
    force_err = None
    if not args.no_cpu_baseline:
        from oracle import mog_oracle as O

        t1 = time.perf_counter()
        want, _ = O.forces(pos_h, rot[:kk], off[:kk], kk, g, W, G, compat=args.compat)
        cpu_s = time.perf_counter() - t1
        force_err = O.rel_rms(gpu_avg, want)
        est_step = cpu_s * k / kk
        cpu = {
            "value": round(1.0 / est_step, 6),
            "unit": "steps/s",
            "cores": 1,
            "kind": "port",
            "sample": f"NumPy FP64 oracle force step of all {n} particles on {kk} of {k} grids "
                      f"({cpu_s:.2f} s), scaled by k/{kk} to one {k}-grid step",
            "interactions_per_s": round(n * kk / cpu_s, 1),
            "host_cpu_count": os.cpu_count(),
        }
        print(f"[bench] cpu baseline {cpu_s:.2f} s for {kk} grids; force rel rms {force_err:.3e}",
              file=sys.stderr, flush=True)

    out = {
        "metric": METRIC,
        "value": round(steps_per_s, 3),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms / args.steps, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32" if args.precision == "fp32" else f"fp32 state, {args.precision} DFT",
        "data": "synthetic",
        "config": {
            "workload": f"{wl}, n={n}, k={k} grids of g={g}, W={W}, "
                        f"G={G}, dt={dt}, damping={damping}, compat={args.compat}, "
                        f"DFT precision {args.precision}",
            "n": n, "k": k, "g": g, "parallelism": "single-gpu",
            "graph": use_graph,
        },
        "interactions_per_s": round(n * k * steps_per_s, 1),
        "force_rel_rms_vs_fp64": None if force_err is None else float(f"{force_err:.3e}"),
        "force_sample_grids": kk,
        "tuning": solver.get_tuning(),
        "autotuned": tuned is not None,
        "phases_ms": {key: round(v, 4) for key, v in phases.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
