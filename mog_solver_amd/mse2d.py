"""Force-MSE harness of the 2D MOG path (grid size G vs ensemble size N).

``python -m mog_solver_amd.mse2d [--n-side 512] [--sizes 8 16 32 64]
[--ensembles 1 32 128] [--trials 3] [--time-steps 50]``

The reference's paper reports the 2D force MSE of MOG against an analytic
solution for G in {8, 16, 32, 64} and N in {1, 32, 128} (paper.tex:43-54,
"Table: 2D Force MSE Convergence"); the code that produced it is not part
of the reference's sources, so this harness is designed from that text:

* particles: a quiet stratified sample (one jittered particle per lattice
  cell, mass-weighted) of an isolated Gaussian disk, sigma = 5, cut at
  r = 20, at the centre of the W = 100 box -- isolated, as the reference's
  rotate-about-the-centre-and-wrap grids require (a rotated periodic field
  is not periodic in the rotated grid);
* analytic force by Gauss's law in 2D with the MOG coupling
  (del^2 phi = 2 pi G h^2 rho, h = W / G_cells):
  F(r) = G h^2 (M(<r)/r - pi M r / W^2) r_hat, the second term the
  neutralising background of the periodic solve;
* the MOG force at every particle from N randomly rotated and offset
  G x G grids (``Mog2D``), trials with independent transforms;
* MSE = mean over particles |F_MOG - F|^2, and relative to mean |F|^2;
* with ``--time-steps`` (default 50), the 2D step's throughput on the same
  particles with the largest ensemble, per grid size.

The MOG coupling scales with h^2, so only the relative MSE compares across
G.  The absolute values depend on units and on the (unpublished) setup, so
the table reproduces the paper's claim, not its numbers: the error falls
with the ensemble size N at every G.  Prints one JSON object.
"""
from __future__ import annotations

import argparse
import json
import sys

import numpy as np


def sample_blob(n_side=512, W=100.0, sigma=5.0, radius=20.0, peak_mass=100.0, seed=0):
    """A quiet, stratified sample of an isolated Gaussian disk at the box
    centre: one particle uniform in each cell of an n_side x n_side lattice
    over the square of half-width ``radius`` around the centre, kept inside
    the circle r < radius, with mass peak_mass exp(-r^2 / (2 sigma^2)).
    (n,4) float32 (x, y, 0, m)."""
    rng = np.random.default_rng(seed)
    d = 2.0 * radius / n_side
    j = np.arange(n_side)
    X, Y = np.meshgrid(j, j, indexing="xy")
    x = (X.ravel() + rng.uniform(0.0, 1.0, X.size)) * d - radius
    y = (Y.ravel() + rng.uniform(0.0, 1.0, Y.size)) * d - radius
    r2 = x * x + y * y
    keep = r2 < radius * radius
    pos = np.zeros((int(keep.sum()), 4), dtype=np.float32)
    pos[:, 0] = x[keep] + W / 2
    pos[:, 1] = y[keep] + W / 2
    pos[:, 3] = peak_mass * np.exp(-r2[keep] / (2.0 * sigma * sigma))
    return pos


def deposited_mass(pos):
    """The masses the deposit sees: uint(m * 1000) / 1000 (float32 product)."""
    return np.floor(np.asarray(pos[:, 3], dtype=np.float32) * np.float32(1000.0)).astype(np.float64) / 1000.0


def analytic_blob_force(pos, W, G, h):
    """Force (the MOG sign convention, grad phi) of an axisymmetric isolated
    mass distribution about the box centre in the periodic box, with the 2D
    MOG coupling del^2 phi = 2 pi G h^2 rho: by Gauss's law
    F(r) = G h^2 (M(<r) / r - pi M r / W^2) r_hat, the second term being
    the neutralising background of the periodic solve (the square lattice
    of images adds no quadratic term).  M(<r) is the particles' own
    enclosed deposited mass."""
    xy = pos[:, :2].astype(np.float64) - W / 2
    r = np.sqrt(np.sum(xy * xy, 1))
    m = deposited_mass(pos)
    order = np.argsort(r)
    enc = np.empty_like(r)
    enc[order] = np.cumsum(m[order]) - 0.5 * m[order]
    M = float(m.sum())
    mag = G * h * h * (enc / np.maximum(r, 1e-12) - np.pi * M * r / (W * W))
    return xy / np.maximum(r, 1e-12)[:, None] * mag[:, None]


def mse_table(n_side=512, sizes=(8, 16, 32, 64), ensembles=(1, 32, 128), trials=3, W=100.0, G=100.0,
              sigma=5.0, radius=20.0, seed=0):
    import torch

    from .mog2d import Mog2D

    pos_h = sample_blob(n_side, W, sigma, radius, seed=seed)
    pos = torch.from_numpy(pos_h).cuda()
    rows = []
    for g in sizes:
        h = W / g
        ref = analytic_blob_force(pos_h, W, G, h)
        ref2 = float(np.mean(np.sum(ref * ref, 1)))
        for N in ensembles:
            errs = []
            for t in range(trials):
                m = Mog2D(N, g, W, G, seed=1000 * t + N)
                f = m.forces(pos).cpu().numpy()[:, :2].astype(np.float64)
                m.close()
                errs.append(float(np.mean(np.sum((f - ref) ** 2, 1))))
            mse = float(np.mean(errs))
            rows.append({"G": g, "N": N, "mse": mse, "rel_mse": mse / ref2, "trials": trials})
    return {"n": int(pos_h.shape[0]), "W": W, "G_const": G, "sigma": sigma, "radius": radius, "rows": rows}


def step_throughput(pos_h, sizes=(8, 16, 32, 64), n_grids=128, steps=50, W=100.0, G=100.0, dt=1e-4,
                    damping=0.999):
    """Milliseconds per 2D step (deposit, solve, gather + update of
    ``mog2d_step``; launches, no graph) and particle-grid interactions per
    second for the harness's particles, per grid size, with ``n_grids``
    grids; the state is a copy, advanced ``3 + steps`` steps."""
    import torch

    from .mog2d import Mog2D

    rows = []
    for g in sizes:
        m = Mog2D(n_grids, g, W, G, seed=7)
        p = torch.from_numpy(pos_h).cuda()
        v = torch.zeros_like(p)
        for _ in range(3):
            m.step(p, v, dt, damping)
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record()
        for _ in range(steps):
            m.step(p, v, dt, damping)
        stop.record()
        torch.cuda.synchronize()
        ms = start.elapsed_time(stop) / steps
        m.close()
        rows.append({"G": g, "N": n_grids, "ms_per_step": round(ms, 4),
                     "interactions_per_s": round(pos_h.shape[0] * n_grids / (ms / 1e3), 1)})
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--n-side", type=int, default=512, help="lattice side of the stratified sample")
    ap.add_argument("--sizes", type=int, nargs="+", default=[8, 16, 32, 64])
    ap.add_argument("--ensembles", type=int, nargs="+", default=[1, 32, 128])
    ap.add_argument("--trials", type=int, default=3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--time-steps", type=int, default=50,
                    help="timed 2D steps per grid size with the largest ensemble (0: skip)")
    args = ap.parse_args(argv)
    import torch

    if not torch.cuda.is_available():
        sys.exit("mse2d needs a HIP device")
    out = mse_table(args.n_side, tuple(args.sizes), tuple(args.ensembles), args.trials, seed=args.seed)
    for r in out["rows"]:
        print(f"G={r['G']:3d} N={r['N']:4d}  MSE {r['mse']:.4e}  rel {r['rel_mse']:.4e}", file=sys.stderr)
    if args.time_steps > 0:
        pos_h = sample_blob(args.n_side, seed=args.seed)
        out["throughput"] = step_throughput(pos_h, tuple(args.sizes), max(args.ensembles), args.time_steps)
        for r in out["throughput"]:
            print(f"G={r['G']:3d} N={r['N']:4d}  {r['ms_per_step']:.4f} ms/step  "
                  f"{r['interactions_per_s']:.3e} interactions/s", file=sys.stderr)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
