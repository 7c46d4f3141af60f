#!/usr/bin/env python3
"""Benchmark of the MOG force step on one MI355X.

Workload: two colliding exponential disks, n = 1e6 (5e5 each, scale radius
8, sech^2 height 0.5, centres 40 apart, relative speed 5, disks inclined 30
degrees), k = 240 grids of g = 16, W = 100, G = 100, dt = 1e-5, damping
0.999, float32 state, seeded synthetic data.  A step is one full force step
(clear, deposit, spectral solve, gather + update) on particles resident in
HBM; K steps are replayed from a hipGraph between two synchronisations.

Printed (one JSON line): steps/s (value), particle-grid interactions/s,
the relative RMS force error of the GPU against the FP64 oracle (all k
grids, initial state), per-phase HIP-event times, the roofline of the
dominant kernel and the CPU baseline: the NumPy FP64 oracle timed on this
host, on one core (a sample of grids) and as a process pool over all grids.
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
# datasheet-derived: 256 CUs x 128 B/clk of LDS (or L1) x 2.4 GHz; dense FP32 vector
LDS_PEAK_TBS = 256 * 128 * 2.4e9 / 1e12
FP32_VECTOR_PEAK_TFS = 157.3

# HBM-side bytes per launch, from two rocprofv3 --pmc runs of this script
# (FETCH_SIZE, then WRITE_SIZE; KiB per dispatch, mean over the dispatches)
# with --copy-calibration 256 --no-autotune on the workload: the 256 MiB
# copy reads as FETCH 128 MiB and WRITE 256 MiB, so FETCH_SIZE is scaled by
# 2.0 and WRITE_SIZE by 1.0.  Figures and commands:
# profiles/pmc_traffic_two_disks_n1e6_k240_g16.md and
# profiles/pmc_traffic_plummer_n4e6_k240_g32.md.  Used only when the run
# matches (workload, n, k, g, precision); the deposit's figure also needs the
# same (grids per workgroup, workgroup target), since its re-reads of the
# positions and its flushes depend on its tiling.  The deposit phase is its
# prepass plus the LDS kernel.
PMC_TRAFFIC_KIB = {
    ("two_disks", 1_000_000, 240, 16, "fp32"): {
        ("deposit", 2, 2048): (7_818.1 + 10_730.2, 15_625.0 + 34_463.9),
        "gather_update": (44_247.9, 72_998.5),
        "solve": (3_734.8, 14_242.9),
    },
    ("plummer", 4_000_000, 240, 32, "fp32"): {
        ("deposit", 1, 4096): (31_256.8 + 67_334.1, 62_500.0 + 681_771.4),
        "gather_update": (544_925.3, 313_054.0),
        "solve": (69_800.9, 190_846.7),
    },
}
# Vector instructions per wave-wide particle-grid pair (SQ_INSTS_VALU / 3.75e6
# for the kernels of the same tiling, profiles/sq_counters_two_disks_n1e6_k240_g16.md):
# the deposit and the gather are bound by vector-instruction issue.
SQ_VALU_PER_WAVE_PAIR = {
    ("two_disks", 1_000_000, 240, 16, "fp32"): {("deposit", 2, 2048): 28.53, "gather_update": 27.88},
}
SIMDS, CLOCK_HZ = 1024, 2.4e9
FETCH_FACTOR, WRITE_FACTOR = 2.0, 1.0


def _oracle_grid_sum(job):
    """Sum over a range of grids of the FP64 oracle's per-grid world force
    (a worker of the CPU baseline's process pool)."""
    pos_h, rot, off, g, W, G, compat = job
    from oracle import mog_oracle as O

    avg, _ = O.forces(pos_h, rot, off, rot.shape[0], g, W, G, compat=compat)
    return avg * rot.shape[0]


def _ready(_):
    return os.getpid()


def start_cpu_pool(workers):
    """The CPU baseline's worker processes, started before this process
    touches the GPU (workers import NumPy and the oracle only)."""
    import concurrent.futures as cf
    import multiprocessing as mp

    # one thread per worker process: the pool is the parallelism (the
    # children inherit the environment at start-up)
    keys = ("OMP_NUM_THREADS", "OPENBLAS_NUM_THREADS", "MKL_NUM_THREADS")
    saved = {key: os.environ.get(key) for key in keys}
    for key in keys:
        os.environ[key] = "1"
    try:
        ex = cf.ProcessPoolExecutor(max_workers=workers, mp_context=mp.get_context("spawn"))
        list(ex.map(_ready, range(workers)))
    finally:
        for key, v in saved.items():
            if v is None:
                os.environ.pop(key, None)
            else:
                os.environ[key] = v
    return ex


def cpu_pool_forces(ex, workers, pos_h, rot, off, g, W, G, compat):
    """FP64 oracle avgForce over all grids, grids sharded across the pool's
    processes; returns (forces, wall seconds)."""
    k = rot.shape[0]
    bounds = np.linspace(0, k, workers + 1).astype(int)
    jobs = [(pos_h, rot[a:b], off[a:b], g, W, G, compat) for a, b in zip(bounds[:-1], bounds[1:]) if b > a]
    t0 = time.perf_counter()
    total = sum(ex.map(_oracle_grid_sum, jobs))
    return total / k, time.perf_counter() - t0


def parse_tune(items) -> dict:
    """``["key=INT", ...]`` of ``--tune`` as a dict (SystemExit on a malformed
    item; unknown keys are rejected by ``MogSolver.set_tuning``)."""
    fields = {}
    for item in items:
        key, sep, val = item.partition("=")
        key = key.strip()
        try:
            if not sep or not key:
                raise ValueError(item)
            fields[key] = int(val)
        except ValueError:
            raise SystemExit(f"--tune expects KEY=INT, got {item!r}") from None
    return fields


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
    ap.add_argument("--drift-steps", type=int, default=100,
                    help="steps of the FP64 momentum/energy drift report (0 = off)")
    ap.add_argument("--profile-reps", type=int, default=20)
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default tuning instead of autotuning first")
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=INT",
                    help="set one tuning field after autotuning (or instead of it with "
                         "--no-autotune), e.g. deposit_wg_target=2048; repeatable")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy a MIB-MiB device buffer three times (the reference "
                         "dispatch for scaling rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    return ap.parse_args()


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    pool = workers = None
    if not args.no_cpu_baseline:
        # at most 16 processes: the host CPU share of one GPU
        workers = max(1, min(16, os.cpu_count() or 1))
        pool = start_cpu_pool(workers)
    import torch

    from mog_solver_amd.config import SolverConfig
    from mog_solver_amd.solver import MogSolver
    from mog_solver_amd import workloads
    from mog_solver_amd.transforms import compute_transforms

    if args.copy_calibration > 0:
        src = torch.ones(args.copy_calibration * 262144, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

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

    tuned = None
    if not args.no_autotune:  # outside the timed region, on copies of the state
        from mog_solver_amd.autotune import autotune

        tuned = autotune(solver, pos, vel, cache_path=None)
        print(f"[bench] autotuned ({tuned['evaluated']} candidates): {tuned['tuning']}",
              file=sys.stderr, flush=True)

    if args.tune:
        solver.set_tuning(**parse_tune(args.tune))
        print(f"[bench] tuning set: {solver.get_tuning()}", file=sys.stderr, flush=True)

    # force error against the FP64 oracle, from the initial state, with the
    # tuning that is timed below (the DFT path is the one choice that changes
    # rounding): all k grids, and a sample of the first cpu_grids grids
    # (every particle) for the one-core baseline
    if args.cpu_grids is None:  # about 10 s of single-core oracle work on this host
        args.cpu_grids = max(1, int(48 * 1_000_000 / n))
    kk = max(1, min(args.cpu_grids, k))
    timed_tuning = solver.get_tuning()
    sample_cfg = SolverConfig.from_dict(dict(cfg_d, n_grids=kk))
    sample = MogSolver(sample_cfg, rot=rot[:kk], off=off[:kk])
    sample.set_tuning(spectral_mfma=timed_tuning["spectral_mfma"])
    gpu_avg = sample.forces(pos).cpu().numpy()[:, :3].astype(np.float64)
    sample.close()
    gpu_full = solver.forces(pos).cpu().numpy()[:, :3].astype(np.float64)

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
    # the gather also clears the spare density grids the next step deposits
    # into (8 B per cell); the profiled "clear" phase is empty
    alg_bytes = {
        "deposit": 16.0 * n + 8.0 * cells,
        "gather_update": 64.0 * n + 24.0 * cells,
        "solve": 24.0 * cells,
        "clear": 0.0,
    }
    dom = max(("deposit", "gather_update", "solve"), key=lambda key: phases[key])
    achieved = alg_bytes[dom] / (phases[dom] / 1e3) / 1e9
    tuning = solver.get_tuning()
    key = (args.workload, n, k, g, args.precision)
    entry = dom
    if dom == "deposit":
        entry = ("deposit", tuning["deposit_grids_per_wg"], tuning["deposit_wg_target"])
    measured = PMC_TRAFFIC_KIB.get(key, {})
    traffic = None
    if entry in measured:
        fetch_kib, write_kib = measured[entry]
        traffic = round((fetch_kib * FETCH_FACTOR + write_kib * WRITE_FACTOR) * 1024.0)
    # the limit the SQ counters show: vector-instruction issue.  Issue rate
    # of the timed phase = profiled instructions per pair x pairs / (SIMDs x
    # cycles); peak = one wave64 instruction per 2 SIMD-cycles
    issue = None
    per_pair = SQ_VALU_PER_WAVE_PAIR.get(key, {}).get(entry)
    if per_pair is not None:
        instr = per_pair * float(n) * k / 64.0
        simd_cycles = SIMDS * CLOCK_HZ * phases[dom] / 1e3
        issue = {
            "limit": "vector-instruction issue",
            "valu_instr_per_wave_pair": per_pair,
            "simd_cycles_per_instr": round(simd_cycles / instr, 3),
            "frac": round(2.0 * instr / simd_cycles, 4),
            "peak": "1 wave64 instruction per 2 SIMD-cycles; about 2.7 for FP32 add/FMA and bit operations "
                    "with vector operands, 4.3 for shifts, rint, max3, conversions and scalar operands",
        }
    roofline = {
        "bound": "hbm",
        "kernel": dom,
        "achieved": round(achieved, 2),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 5),
        "traffic": traffic,
        "alg_bytes_per_launch": alg_bytes[dom],
        "avg_launch_ms": round(phases[dom], 4),
        "issue": issue,
    }

    # achieved fractions per phase: HBM (algorithmic bytes), and the level
    # that bounds the scatter/gather (8 B per LDS atomic read-modify-write,
    # 16 B per cell fetch) and the transforms (flop count of the path used)
    inter = float(n) * k
    # the FP32 vector path runs FFTs from g = 16 up (counted as full complex
    # 3D FFTs, 5 N log2 N flops each: one forward and three inverse); the
    # matrix-core paths and smaller grids run dense DFTs
    fft = g >= 16 and args.precision == "fp32" and not tuning.get("spectral_mfma", 0)
    dft_flops = (4.0 * 5.0 * g**3 * math.log2(g**3) * k) if fft else 96.0 * k * g**4
    phase_fractions = {}
    for key in ("clear", "deposit", "solve", "gather_update"):
        sec = phases[key] / 1e3
        hbm = alg_bytes[key] / sec / 1e9
        entry = {"ms": round(phases[key], 4), "hbm_GBps": round(hbm, 1), "hbm_frac": round(hbm / HBM_PEAK_GBS, 5)}
        if key == "deposit":
            tbs = 8.0 * inter / sec / 1e12
            entry.update(lds_atomic_TBps=round(tbs, 3), lds_frac=round(tbs / LDS_PEAK_TBS, 4))
        elif key == "gather_update":
            tbs = 16.0 * inter / sec / 1e12
            entry.update(cell_fetch_TBps=round(tbs, 3), cache_frac=round(tbs / LDS_PEAK_TBS, 4))
        elif key == "solve":
            tfs = dft_flops / sec / 1e12
            entry.update(dft="fft" if fft else "dense", dft_flops_per_launch=dft_flops,
                         dft_TFLOPs=round(tfs, 3), fp32_frac=round(tfs / FP32_VECTOR_PEAK_TFS, 4))
        phase_fractions[key] = entry
    phase_fractions["peaks"] = {"hbm_GBps": HBM_PEAK_GBS, "lds_TBps": round(LDS_PEAK_TBS, 2),
                                "fp32_vector_TFLOPs": FP32_VECTOR_PEAK_TFS}

    # FP64 diagnostics over --drift-steps steps from the initial state (on a
    # copy, outside the timed region): momentum and energy drift
    drift = None
    if args.drift_steps > 0:
        p_d = torch.from_numpy(pos_h).cuda()
        v_d = torch.from_numpy(vel_h).cuda()
        d0 = solver.diagnostics(p_d, v_d)
        solver.run(p_d, v_d, args.drift_steps, use_graph=use_graph)
        d1 = solver.diagnostics(p_d, v_d)
        mv = float(np.sum(pos_h[:, 3].astype(np.float64) * np.linalg.norm(vel_h[:, :3].astype(np.float64), axis=1)))
        e0, e1 = d0["kinetic"] + d0["potential"], d1["kinetic"] + d1["potential"]
        drift = {
            "steps": args.drift_steps,
            "kinetic": [d0["kinetic"], d1["kinetic"]],
            "potential": [d0["potential"], d1["potential"]],
            "energy_change_rel_potential": float(f"{(e1 - e0) / abs(d0['potential']):.4e}"),
            "momentum_change_rel_sum_m_v": float(f"{np.linalg.norm(d1['momentum'] - d0['momentum']) / mv:.4e}"),
            "note": "FP64 (mog_diagnostics); damping removes kinetic energy every step, so E is not conserved",
        }
        del p_d, v_d
        print(f"[bench] drift over {args.drift_steps} steps: {drift}", file=sys.stderr, flush=True)

    cpu = None
    force_err = force_err_sample = None
    if not args.no_cpu_baseline:
        from oracle import mog_oracle as O

        # (i) one core: the oracle on the first kk grids, scaled to k grids
        t1 = time.perf_counter()
        want, _ = O.forces(pos_h, rot[:kk], off[:kk], kk, g, W, G, compat=args.compat)
        cpu_s = time.perf_counter() - t1
        force_err_sample = O.rel_rms(gpu_avg, want)
        # (ii) the process pool sharding all k grids over the host's cores;
        # also the force error over all k grids
        want_full, pool_s = cpu_pool_forces(pool, workers, pos_h, rot, off, g, W, G, args.compat)
        pool.shutdown()
        force_err = O.rel_rms(gpu_full, want_full)
        cpu = {
            "value": round(1.0 / pool_s, 6),
            "unit": "steps/s",
            "cores": workers,
            "kind": "port",
            "sample": f"NumPy FP64 oracle force step (deposit, FFT solve, gather) of all {n} particles "
                      f"on all {k} grids, grids sharded over {workers} processes: {pool_s:.2f} s wall",
            "interactions_per_s": round(n * k / pool_s, 1),
            "single_core": {
                "value": round(kk / (cpu_s * k), 6),
                "unit": "steps/s",
                "cores": 1,
                "sample": f"{kk} of {k} grids in {cpu_s:.2f} s, scaled by k/{kk}",
            },
            "host_cpu_count": os.cpu_count(),
        }
        print(f"[bench] cpu baseline: 1 core {cpu_s:.2f} s for {kk} grids; {workers} processes "
              f"{pool_s:.2f} s for {k} grids; force rel rms {force_err:.3e} (all grids)",
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
        "force_rel_rms_vs_fp64_first_grids": None if force_err_sample is None else float(f"{force_err_sample:.3e}"),
        "force_sample_grids": kk,
        "force_error_dft_path": "matrix-core fp32 DFT" if (args.precision == "fp32" and timed_tuning["spectral_mfma"]
                                                             and g >= 16) else
                                ("vector FFT" if args.precision == "fp32" else f"matrix-core {args.precision} DFT"),
        "tuning": tuning,
        "diagnostics": drift,
        "autotuned": tuned is not None,
        "phases_ms": {key: round(v, 4) for key, v in phases.items()},
        "roofline": roofline,
        "phase_fractions": phase_fractions,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
