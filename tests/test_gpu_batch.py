"""Batched ensemble: one persistent launch per phase over uneven simulations.

Bar: bit-exact.  A batch step must leave every simulation's positions and
velocities identical to stepping that simulation alone (MogSolver with the
FP32 vector spectral solve), whose parity with the FP64 oracle is checked in
test_gpu_parity.py; graph replay and every deposit task shape give the same
bits.
"""
import numpy as np
import pytest

from mog_solver_amd.config import SolverConfig
from mog_solver_amd import workloads

pytestmark = pytest.mark.gpu

SIMS = [  # (n_grids, grid_size, n, compat, solver)
    (3, 16, 40000, "reference", "fft"),
    (8, 32, 5000, "physical", "fft"),
    (17, 8, 12000, "reference", "fft"),
    (5, 4, 100, "physical", "fft"),
    (2, 16, 0, "reference", "fft"),
    (6, 16, 35000, "physical", "convolution"),
    (4, 32, 33000, "reference", "fft"),
]


def configs():
    out = []
    for k, g, n, compat, solver in SIMS:
        out.append(SolverConfig.from_dict(dict(n_grids=k, grid_size=g, world_size=100.0, G=100.0, dt=1e-4,
                                               damping=0.999, solver=solver, compat=compat, num_particles=n)))
    return out


def states():
    rng = np.random.default_rng(7)
    out = []
    for i, (_, _, n, _, _) in enumerate(SIMS):
        pos, vel = workloads.plummer_sphere(n, scale=8.0, rng=rng)
        vel[:, :3] = rng.normal(0, 50, (n, 3)).astype(np.float32)
        out.append((pos, vel))
    return out


def individual_steps(torch, cfgs, batch, sts, nsteps):
    from mog_solver_amd.solver import MogSolver

    res = []
    for s, (cfg, (pos, vel)) in enumerate(zip(cfgs, sts)):
        rot, off = batch.transforms[s]
        solver = MogSolver(cfg, rot=rot, off=off)
        solver.set_tuning(spectral_mfma=0)
        p, v = torch.from_numpy(pos.copy()).cuda(), torch.from_numpy(vel.copy()).cuda()
        if p.shape[0]:
            for _ in range(nsteps):
                solver.step(p, v)
        res.append((p.cpu().numpy(), v.cpu().numpy()))
    return res


def test_batch_step_bit_exact_vs_individual(torch_cuda):
    from mog_solver_amd.batch import MogBatch

    torch = torch_cuda
    cfgs, sts = configs(), states()
    batch = MogBatch(cfgs, [p.shape[0] for p, _ in sts], seeds=list(range(10, 10 + len(cfgs))))
    p, v = batch.pack(sts)
    for _ in range(3):
        batch.step(p, v)
    want = individual_steps(torch, cfgs, batch, sts, 3)
    for s, (gp, gv, (wp, wv)) in enumerate(zip(batch.views(p), batch.views(v), want)):
        assert np.array_equal(gp.cpu().numpy(), wp), f"sim {s} positions"
        assert np.array_equal(gv.cpu().numpy(), wv), f"sim {s} velocities"


def test_batch_graph_and_tuning_variants_agree(torch_cuda):
    from mog_solver_amd.batch import MogBatch

    torch = torch_cuda
    cfgs, sts = configs(), states()
    batch = MogBatch(cfgs, [p.shape[0] for p, _ in sts], seeds=list(range(len(cfgs))))
    p0, v0 = batch.pack(sts)
    ref_p, ref_v = p0.clone(), v0.clone()
    for _ in range(4):
        batch.step(ref_p, ref_v)
    for chunk, threads, graph in [(16384, 1024, True), (1000, 256, False), (4096, 512, True)]:
        batch.set_tuning(deposit_chunk=chunk, deposit_threads=threads)
        assert batch.get_tuning()["n_tasks"] > 0
        p, v = p0.clone(), v0.clone()
        batch.run(p, v, 2, use_graph=graph)
        batch.run(p, v, 2, use_graph=graph)  # graph replayed, not rebuilt
        assert torch.equal(p, ref_p) and torch.equal(v, ref_v), (chunk, threads, graph)


@pytest.mark.parametrize("graph", [True, False])
def test_batch_run_state_across_reorders(torch_cuda, graph):
    """A batch run keeps the ensemble's state in visiting order across a
    rebuild of the order (every 16 steps) and writes it back at the end:
    the same bits as single batch steps through the caller's arrays."""
    from mog_solver_amd.batch import MogBatch

    torch = torch_cuda
    cfgs, sts = configs(), states()
    batch = MogBatch(cfgs, [p.shape[0] for p, _ in sts], seeds=list(range(20, 20 + len(cfgs))))
    p0, v0 = batch.pack(sts)
    ref_p, ref_v = p0.clone(), v0.clone()
    for _ in range(18):
        batch.step(ref_p, ref_v)
    p, v = p0.clone(), v0.clone()
    batch.run(p, v, 18, use_graph=graph)
    assert torch.equal(p, ref_p) and torch.equal(v, ref_v)


def test_batch_rejects_unsupported(torch_cuda):
    from mog_solver_amd._abi import MogError
    from mog_solver_amd.batch import MogBatch

    base = dict(n_grids=2, world_size=100.0, G=100.0, dt=1e-5, damping=1.0, solver="fft")
    for extra in (dict(grid_size=64), dict(grid_size=16, precision="bf16x3")):
        with pytest.raises(MogError):
            MogBatch([SolverConfig.from_dict({**base, **extra})], [10])
    b = MogBatch([SolverConfig.from_dict({**base, "grid_size": 8})], [10])
    with pytest.raises(MogError):
        b.set_tuning(deposit_chunk=100)
    with pytest.raises(MogError):
        b.set_tuning(deposit_threads=300)


def test_batch_ensemble_workload_runs(torch_cuda):
    """A slice of the uneven Plummer ensemble: finite state, particles in
    the box, and the same result as the individual simulations."""
    from mog_solver_amd.batch import MogBatch

    torch = torch_cuda
    ens = workloads.plummer_ensemble(n_sims=6, n_min=1e4, n_max=6e4, grids=(32, 64), seed=4)
    cfgs = [SolverConfig.from_dict(c) for c, _, _ in ens]
    sts = [(p, v) for _, p, v in ens]
    batch = MogBatch(cfgs, [p.shape[0] for p, _ in sts])
    p, v = batch.pack(sts)
    batch.run(p, v, 3)
    assert bool(torch.isfinite(p).all()) and bool(torch.isfinite(v).all())
    assert float(p[:, :3].min()) >= 0.0 and float(p[:, :3].max()) < 100.0
    want = individual_steps(torch, cfgs, batch, sts, 3)
    for s, (gp, (wp, _)) in enumerate(zip(batch.views(p), want)):
        assert np.array_equal(gp.cpu().numpy(), wp), f"sim {s}"


def test_batch_masses_match_individual(torch_cuda):
    """Fractional, zero, negative and very large masses (the 64-bit global
    path of the batched deposit), in one-grid tasks (g = 32, two particles
    per lane) and multi-grid tasks (g = 16), with chunk sizes that leave a
    partial last pair: the same state as the individual simulations."""
    from mog_solver_amd.batch import MogBatch

    torch = torch_cuda
    sims = [(3, 32, 9001), (9, 16, 7003)]
    cfgs = [SolverConfig.from_dict(dict(n_grids=k, grid_size=g, world_size=100.0, G=100.0, dt=1e-6,
                                        damping=0.999, solver="fft", num_particles=n))
            for k, g, n in sims]
    rng = np.random.default_rng(21)
    sts = []
    for _, _, n in sims:
        pos = np.zeros((n, 4), np.float32)
        pos[:, :3] = rng.uniform(0, 100, (n, 3))
        pos[:, 3] = rng.uniform(0, 3, n).astype(np.float32)
        pos[:50, 3] = 0.0
        pos[50:100, 3] = -1.0
        pos[100:140, 3] = np.float32(5e9)  # milli-mass 5e12: beyond 32 bits
        sts.append((pos, np.zeros((n, 4), np.float32)))
    batch = MogBatch(cfgs, [p.shape[0] for p, _ in sts], seeds=[3, 4])
    want = individual_steps(torch, cfgs, batch, sts, 1)
    for chunk, threads in [(65536, 1024), (1000, 256), (3333, 512)]:
        batch.set_tuning(deposit_chunk=chunk, deposit_threads=threads)
        p, v = batch.pack(sts)
        batch.step(p, v)
        for s, (gp, gv, (wp, wv)) in enumerate(zip(batch.views(p), batch.views(v), want)):
            assert np.array_equal(gp.cpu().numpy(), wp), (chunk, threads, s)
            assert np.array_equal(gv.cpu().numpy(), wv), (chunk, threads, s)
