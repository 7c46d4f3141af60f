"""Throughput of the batched ensemble on the uneven Plummer ensemble.

``python -m mog_solver_amd.ensemble_report [--sims 64] [--steps 50]``

Builds the 64-simulation Plummer ensemble (n log-uniform in [1e4, 5e5],
n_grids in {32, 64, 128, 240}, grid_size in {16, 32}, scale radius 5,
seed 4), autotunes the batch's deposit task shape, and times
(a) the batch: one persistent launch per phase, the step replayed from one
    hipGraph;
(b) the same simulations stepped one after another, each with its own
    context and captured graph (the way a user of the single-simulation
    interface would run them).
Inputs are resident on the device; each timing is bracketed by device
synchronisations.  Prints one JSON object.
"""
from __future__ import annotations

import argparse
import json
import sys
import time


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--sims", type=int, default=64)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--seed", type=int, default=4)
    ap.add_argument("--no-autotune", action="store_true")
    args = ap.parse_args(argv)

    import torch

    from . import workloads
    from .batch import MogBatch, autotune
    from .config import SolverConfig
    from .solver import MogSolver

    if not torch.cuda.is_available():
        sys.exit("ensemble_report needs a HIP device")
    ens = workloads.plummer_ensemble(n_sims=args.sims, seed=args.seed)
    cfgs = [SolverConfig.from_dict(c) for c, _, _ in ens]
    states = [(p, v) for _, p, v in ens]
    counts = [p.shape[0] for p, _ in states]
    batch = MogBatch(cfgs, counts)
    pos, vel = batch.pack(states)
    tuning = batch.get_tuning() if args.no_autotune else autotune(batch, pos, vel)

    def clock(fn, steps):
        fn(2)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn(steps)
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    t_batch = clock(lambda s: batch.run(pos, vel, s, use_graph=True), args.steps)

    solvers = [MogSolver(c, rot=r, off=o) for c, (r, o) in zip(cfgs, batch.transforms)]
    ps = [t.clone() for t in batch.views(pos)]
    vs = [t.clone() for t in batch.views(vel)]

    def sequential(steps):
        for _ in range(steps):
            for m, p, v in zip(solvers, ps, vs):
                m.run(p, v, 1, use_graph=True)

    t_seq = clock(sequential, max(5, args.steps // 5))
    interactions = sum(c.n_grids * n for c, n in zip(cfgs, counts))
    cells = sum(c.n_grids * c.grid_size ** 3 for c in cfgs)
    out = {
        "workload": f"{args.sims} Plummer spheres (a=5), n log-uniform [1e4, 5e5], "
                    f"n_grids in {{32,64,128,240}}, grid_size in {{16,32}}, seed {args.seed}",
        "particles": int(sum(counts)),
        "interactions_per_step": int(interactions),
        "cells": int(cells),
        "tuning": tuning,
        "batch_ms_per_step": round(t_batch * 1e3, 4),
        "batch_steps_per_s": round(1.0 / t_batch, 3),
        "batch_interactions_per_s": round(interactions / t_batch, 1),
        "sequential_ms_per_step": round(t_seq * 1e3, 4),
        "speedup_vs_sequential": round(t_seq / t_batch, 3),
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
