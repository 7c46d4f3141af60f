"""GPU parity on the benchmark workloads themselves, against the FP64 oracle.

* Two colliding disks (n = 1e6, k = 240, g = 16; the headline benchmark):
  density grids bit-exact on a 48-grid sample (every 5th of the 240 grids;
  a grid's density depends only on its own transform), forces within 1e-5
  relative RMS over all 240 grids on both FP32 transform paths, and the
  graph-replayed step equal to its phases bit for bit.
* 64 Plummer spheres (a = 1.5) at g = 32, k = 240, n = 4e6, with the DFT in
  FP32, BF16x3 and scaled FP16: density bit-exact on a 12-grid sample;
  force grids of that sample against the FP64 solve; particle forces of a
  24-grid solver (the first 24 of the 240 transforms) against the oracle on
  the same 24 grids.
* The 64-simulation Plummer ensemble (n log-uniform in [1e4, 5e5],
  k in {32, 64, 128, 240}, g in {16, 32}): one batched step bit-exact
  against each simulation stepped alone, and the forces of three of its
  simulations (recovered from the velocities of that step, which start at
  zero) against the oracle.
"""
import numpy as np
import pytest

from oracle import mog_oracle as O
from mog_solver_amd.config import SolverConfig
from mog_solver_amd import workloads

from test_gpu_parity import dev, make_solver

pytestmark = pytest.mark.gpu

W, G = 100.0, 100.0


def test_two_disks_n1e6_k240_g16(torch_cuda):
    torch = torch_cuda
    n, k, g = 1_000_000, 240, 16
    pos, vel = workloads.two_disks(n, W, G, g, seed=2)
    s, rot, off = make_solver(k, g, W, G)
    p = dev(torch, pos)
    grids = s.deposit(p).cpu().numpy()
    assert np.all(grids.sum(axis=(1, 2, 3)) == n * 1000)
    sample = np.arange(0, k, 5)
    want = O.deposit(pos, rot[sample], off[sample], O.Params.make(sample.size, g, W))
    assert np.array_equal(grids[sample], want), f"{np.count_nonzero(grids[sample] != want)} cells differ"
    want_f, _ = O.forces(pos, rot, off, k, g, W, G)
    errs = {}
    for mfma in (0, 1):
        s.set_tuning(spectral_mfma=mfma)
        got = s.forces(p).cpu().numpy()
        errs[mfma] = O.rel_rms(got[:, :3], want_f)
    print(f"two disks: force rel rms vs FP64 (all 240 grids): vector FFT {errs[0]:.3e}, matrix-core {errs[1]:.3e}")
    assert max(errs.values()) < 1e-5, errs
    # the graph-replayed step (the benchmark's timed path) = its phases
    for mfma in (0, 1):
        s.set_tuning(spectral_mfma=mfma)
        p1, v1 = dev(torch, pos), dev(torch, vel)
        s.run(p1, v1, 1, use_graph=True)
        p2, v2 = dev(torch, pos), dev(torch, vel)
        s.update(p2, v2, s.forces(p2))
        torch.cuda.synchronize()
        assert torch.equal(p1, p2) and torch.equal(v1, v2), mfma


_C3 = {}


def _c3_state():
    if not _C3:
        n, k, g = 4_000_000, 240, 32
        pos, _ = workloads.plummer_spheres(n, n_spheres=64, scale=1.5, world_size=W, seed=3)
        rot, off = O.se3_transforms(k, g, W, 0)
        sample = np.arange(0, k, 20)
        P = O.Params.make(sample.size, g, W)
        grids = O.deposit(pos, rot[sample], off[sample], P)
        ks = 24
        want_f, _ = O.forces(pos, rot[:ks], off[:ks], ks, g, W, G)
        _C3.update(pos=pos, rot=rot, off=off, sample=sample, grids=grids,
                   solved=O.solve(grids, g, W, G), ks=ks, want_f=want_f)
    return _C3


# (precision, per-grid force-field bar, particle-force bar)
@pytest.mark.parametrize("precision,grid_tol,force_tol", [
    ("fp32", 2e-6, 1e-5), ("bf16x3", 5e-5, 1e-5), ("fp16", 5e-3, 5e-4)])
def test_plummer_n4e6_k240_g32_precisions(torch_cuda, precision, grid_tol, force_tol):
    torch = torch_cuda
    k, g = 240, 32
    c = _c3_state()
    s, rot, off = make_solver(k, g, W, G, precision=precision)
    assert np.array_equal(rot, c["rot"]) and np.array_equal(off, c["off"])
    p = dev(torch, c["pos"])
    grids = s.deposit(p)
    assert np.array_equal(grids.cpu().numpy()[c["sample"]], c["grids"])
    fg = s.solve(grids).cpu().numpy()[c["sample"]].astype(np.float64)
    gerr = [O.rel_rms(fg[i, ..., :3].reshape(-1, 3), c["solved"][i, ..., :3].reshape(-1, 3))
            for i in range(c["sample"].size)]
    ks = c["ks"]
    from mog_solver_amd.solver import MogSolver

    sub = MogSolver(SolverConfig.from_dict(dict(n_grids=ks, grid_size=g, world_size=W, G=G, dt=1e-5, damping=0.999,
                                                solver="fft", precision=precision)),
                    rot=c["rot"][:ks], off=c["off"][:ks])
    ferr = O.rel_rms(sub.forces(p).cpu().numpy()[:, :3], c["want_f"])
    print(f"64 Plummer g=32 {precision}: grid rel rms max {max(gerr):.2e}, particle force ({ks} grids) {ferr:.2e}")
    assert max(gerr) < grid_tol, gerr
    assert ferr < force_tol


def test_plummer_ensemble_64_sims(torch_cuda):
    from mog_solver_amd.batch import MogBatch
    from mog_solver_amd.solver import MogSolver

    torch = torch_cuda
    ens = workloads.plummer_ensemble()
    assert len(ens) == 64
    cfgs = [SolverConfig.from_dict(cd) for cd, _, _ in ens]
    sts = [(p, v) for _, p, v in ens]
    assert all(np.all(v == 0) for _, v in sts)
    batch = MogBatch(cfgs, [p.shape[0] for p, _ in sts])
    p, v = batch.pack(sts)
    batch.step(p, v)
    gp = [t.cpu().numpy() for t in batch.views(p)]
    gv = [t.cpu().numpy() for t in batch.views(v)]
    for s, (cfg, (pos, vel)) in enumerate(zip(cfgs, sts)):
        rot, off = batch.transforms[s]
        one = MogSolver(cfg, rot=rot, off=off)
        one.set_tuning(spectral_mfma=0)
        pp, vv = dev(torch, pos), dev(torch, vel)
        one.step(pp, vv)
        assert np.array_equal(pp.cpu().numpy(), gp[s]), f"sim {s} positions"
        assert np.array_equal(vv.cpu().numpy(), gv[s]), f"sim {s} velocities"
        one.close()
    # three simulations against the oracle: the smallest of each grid size
    # and the smallest with k = 240
    by_n = sorted(range(len(cfgs)), key=lambda i: sts[i][0].shape[0])
    idx = [next(i for i in by_n if cfgs[i].grid_size == 16), next(i for i in by_n if cfgs[i].grid_size == 32)]
    idx.append(next(i for i in by_n if cfgs[i].n_grids == 240 and i not in idx))
    for s in idx:
        cfg = cfgs[s]
        rot, off = batch.transforms[s]
        want, _ = O.forces(sts[s][0], rot, off, cfg.n_grids, cfg.grid_size, W, G)
        dtp = np.float64(np.float32(cfg.dt * cfg.timescale))
        acc = -gv[s][:, :3].astype(np.float64) / (dtp * np.float64(np.float32(cfg.damping)))
        err = O.rel_rms(acc, want)
        print(f"ensemble sim {s} (n={sts[s][0].shape[0]}, k={cfg.n_grids}, g={cfg.grid_size}): force rel rms {err:.2e}")
        assert err < 1e-5
