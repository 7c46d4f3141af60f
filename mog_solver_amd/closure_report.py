"""Train and evaluate the learned closure on the two-colliding-disks workload.

``python -m mog_solver_amd.closure_report [--n 1000000] [--snapshots 200] ...``

1. Advance the k_full-grid simulation (n particles, two exponential disks)
   and record ``--snapshots`` (features of the k_small ensemble, k_full
   force) pairs, one every ``--every`` steps.
2. Train the closure MLP on the GPU (fused BF16 training kernel + Adam).
3. From the final state run ``--eval-steps`` closure steps.  At every step
   the closure force is compared with the k_full force at the same positions
   (relative RMS); every ``--direct-every`` steps the k_full, k_small and
   closure forces are compared with FP64 direct summation (softening one
   cell) at a ``--subsample`` of particles, and with the estimator floor:
   a second, independent k_full ensemble (other rotations and offsets) and
   a ``--k-ref``-grid ensemble standing in for the k -> infinity limit.  A
   k_full ensemble is a random draw, so no predictor can be closer to one
   draw than the draw's own scatter about the ensemble mean.
4. Autotune both solvers' tilings (unless ``--no-autotune``), then time a
   plain k_full step (hipGraph replay), a closure step (features, MLP,
   update) and the ``ClosureSolver`` as shipped (default audit every 10
   steps against k_full and FP64 direct summation, default threshold and
   fallback), reporting the share of its steps that used the closure; all
   with inputs resident on the device.

Prints one JSON object.
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np


def _rel(a, b):
    d = a - b
    return float(np.linalg.norm(d) / max(np.linalg.norm(b), 1e-300))


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n", type=int, default=1_000_000)
    ap.add_argument("--k-full", type=int, default=240)
    ap.add_argument("--k-small", type=int, default=24)
    ap.add_argument("--g", type=int, default=16)
    ap.add_argument("--snapshots", type=int, default=200)
    ap.add_argument("--every", type=int, default=5)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-final", type=float, default=1e-4)
    ap.add_argument("--eval-steps", type=int, default=100)
    ap.add_argument("--direct-every", type=int, default=10)
    ap.add_argument("--subsample", type=int, default=4096)
    ap.add_argument("--time-steps", type=int, default=100)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--k-ref", type=int, default=2400,
                    help="grids of the reference ensemble of the floor measurement (0 = off)")
    ap.add_argument("--target-k", type=int, default=0,
                    help="train on the forces of a k-grid ensemble (independent draw) instead of k_full's")
    ap.add_argument("--compat", default="reference", choices=["reference", "physical"])
    ap.add_argument("--no-autotune", action="store_true",
                    help="time both solvers with the library's default tuning")
    args = ap.parse_args(argv)

    import torch

    from . import closure as CL
    from . import workloads
    from .config import SolverConfig
    from .solver import MogSolver
    from .transforms import compute_transforms

    if not torch.cuda.is_available():
        sys.exit("closure_report needs a HIP device")
    W, G, dt, damping = 100.0, 100.0, 1e-5, 0.999

    def make(k, seed):
        cfg = SolverConfig.from_dict(dict(n_grids=k, grid_size=args.g, world_size=W, G=G, dt=dt, damping=damping,
                                          solver="fft", compat=args.compat))
        off, rot = compute_transforms(k, args.g, W, seed=seed)
        return MogSolver(cfg, rot=rot, off=off)

    def log(msg):
        print(f"[closure_report] {msg}", file=sys.stderr, flush=True)

    n = args.n
    full = make(args.k_full, args.seed + 1)
    small = make(args.k_small, args.seed + 2)
    pos, vel = workloads.two_disks(n=n, grid_size=args.g)
    p = torch.from_numpy(pos).cuda()
    v = torch.from_numpy(vel).cuda()
    rho0 = CL.default_rho0(n, 1.0, args.g)
    out = dict(workload=f"two_disks n={n}", k_full=args.k_full, k_small=args.k_small, g=args.g,
               snapshots=args.snapshots, every=args.every, epochs=args.epochs, compat=args.compat,
               target_k=args.target_k or args.k_full)

    t0 = time.perf_counter()
    target = make(args.target_k, args.seed + 5) if args.target_k else None
    snaps = CL.make_snapshots(full, small, p, v, args.snapshots, every=args.every, rho0=rho0, target=target)
    if target is not None:
        target.close()
    torch.cuda.synchronize()
    out["snapshot_s"] = time.perf_counter() - t0
    log(f"{len(snaps)} snapshots in {out['snapshot_s']:.1f}s")

    mlp = CL.ClosureMLP(seed=args.seed, bf16=True)
    t0 = time.perf_counter()
    hist = CL.train(mlp, snaps, epochs=args.epochs, lr=args.lr, lr_final=args.lr_final)
    torch.cuda.synchronize()
    out["train_s"] = time.perf_counter() - t0
    out["train_steps"] = args.epochs * len(snaps)
    out["loss_first_epoch"], out["loss_last_epoch"] = hist[0], hist[-1]
    out["train_rel_rms"] = CL.evaluate_all(mlp, snaps)
    log(f"trained in {out['train_s']:.1f}s: loss {hist[0]:.3e} -> {hist[-1]:.3e}")
    del snaps
    torch.cuda.empty_cache()

    # evaluation on the steps that follow the training window
    cs = CL.ClosureSolver(full, small, mlp, rho0, audit_every=0)
    G_eff = workloads.effective_G(G, W, args.g)
    rng = np.random.default_rng(args.seed)
    sub = np.sort(rng.choice(n, size=min(args.subsample, n), replace=False)).astype(np.int32)
    sub_t = torch.from_numpy(sub.astype(np.int64)).cuda()
    vs_full = {"closure": [], "k_small_ngp": [], "k_small_cic": []}
    direct = {"k_full": [], "k_small_ngp": [], "k_small_cic": [], "closure": []}
    full_b = make(args.k_full, args.seed + 3)
    kref = make(args.k_ref, args.seed + 4) if args.k_ref else None
    floor = {"k_full_vs_independent_k_full": []}
    if kref is not None:
        floor.update({f"{name}_vs_k_ref": [] for name in ("k_full", "closure", "k_small_cic", "k_small_ngp")})
    for step in range(args.eval_steps):
        feat, mean, pred = cs._buffers(p)
        CL.closure_features(small, p, rho0, feat=feat, mean=mean)
        mlp.forward(feat, mean, out=pred)
        ref = full.forces(p)
        ngp = CL.ngp_mean(feat)
        vs_full["closure"].append(CL.rel_rms(pred, ref))
        vs_full["k_small_ngp"].append(CL.rel_rms(ngp, ref))
        vs_full["k_small_cic"].append(CL.rel_rms(mean, ref))
        if args.direct_every and step % args.direct_every == 0:
            nd = CL.direct_sum(p, sub, W, G_eff, W / args.g).cpu().numpy()
            for name, t in (("k_full", ref), ("k_small_ngp", ngp), ("k_small_cic", mean), ("closure", pred)):
                direct[name].append(_rel(t[sub_t, :3].double().cpu().numpy(), nd))
            floor["k_full_vs_independent_k_full"].append(CL.rel_rms(ref, full_b.forces(p)))
            if kref is not None:
                fr = kref.forces(p)
                for name, t in (("k_full", ref), ("closure", pred), ("k_small_cic", mean), ("k_small_ngp", ngp)):
                    floor[f"{name}_vs_k_ref"].append(CL.rel_rms(t, fr))
        small.update(p, v, pred)
    out["eval_steps"] = args.eval_steps
    out["vs_k_full_rel_rms"] = {k2: dict(mean=float(np.mean(v2)), max=float(np.max(v2))) for k2, v2 in vs_full.items()}
    out["vs_fp64_direct_sum_rel_rms"] = {k2: float(np.mean(v2)) for k2, v2 in direct.items()}
    out["direct_sum"] = dict(subsample=int(sub.size), softening=W / args.g, G_eff=G_eff)
    out["floor_rel_rms"] = {k2: float(np.mean(v2)) for k2, v2 in floor.items() if v2}
    out["floor"] = dict(k_ref=args.k_ref, note="k_full vs an independent k_full draw; every ensemble and the "
                        "closure vs the k_ref-grid ensemble (the k -> infinity stand-in)")
    full_b.close()
    if kref is not None:
        kref.close()
    log(f"floor: {out['floor_rel_rms']}")
    log(f"eval vs k_full: {out['vs_k_full_rel_rms']}")

    # timing, inputs resident on the device
    def clock(fn, steps):
        fn(3)
        torch.cuda.synchronize()
        t = time.perf_counter()
        fn(steps)
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps

    if not args.no_autotune:  # both solvers, on copies of the state
        from .autotune import autotune

        # the k_small solver is tuned on its plain step, which can do without
        # the Morton visiting order; the closure's feature pass cannot, so
        # that order is kept
        resort = small.get_tuning()["resort_every"]
        out["tuning"] = {name: autotune(sv, p, v, cache_path=None)["tuning"]
                         for name, sv in (("k_full", full), ("k_small", small))}
        small.set_tuning(resort_every=resort)
        out["tuning"]["k_small"] = small.get_tuning()
        log(f"autotuned: {out['tuning']}")
    K = args.time_steps
    t_full = clock(lambda s: full.run(p, v, s, use_graph=True), K)
    runner = CL.ClosureSolver(full, small, mlp, rho0, audit_every=0)

    def closure_steps(s):
        for _ in range(s):
            runner.step(p, v)

    t_closure = clock(closure_steps, K)
    audited = CL.ClosureSolver(full, small, mlp, rho0, audit_every=10, threshold=math.inf, direct_subsample=0)

    def audited_steps(s):
        for _ in range(s):
            audited.step(p, v)

    t_audit = clock(audited_steps, K)
    shipped = CL.ClosureSolver(full, small, mlp, rho0)  # defaults: audits, threshold, fallback, direct sum

    def shipped_steps(s):
        for _ in range(s):
            shipped.step(p, v)

    t_shipped = clock(shipped_steps, K)
    out["ms_per_step"] = dict(k_full=t_full * 1e3, closure=t_closure * 1e3, closure_audit_every_10=t_audit * 1e3,
                              closure_solver_defaults=t_shipped * 1e3)
    out["speedup_vs_k_full"] = dict(closure=t_full / t_closure, closure_audit_every_10=t_full / t_audit,
                                    closure_solver_defaults=t_full / t_shipped)
    out["closure_solver_defaults"] = dict(
        threshold=shipped.threshold, audit_every=shipped.audit_every, fallback_steps=shipped.fallback_steps,
        direct_subsample=shipped.direct_subsample, direct_ratio=shipped.direct_ratio,
        closure_step_share=shipped.closure_steps / max(shipped.steps, 1),
        audits=[dict(step=a.step, vs_full=round(a.error_vs_full, 4),
                     vs_direct=None if a.error_vs_direct is None else round(a.error_vs_direct, 4),
                     full_vs_direct=None if a.full_vs_direct is None else round(a.full_vs_direct, 4),
                     fallback=a.fallback, reason=a.reason) for a in shipped.audits])
    print(json.dumps(out))


if __name__ == "__main__":
    main()
