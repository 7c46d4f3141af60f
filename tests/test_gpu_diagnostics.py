"""FP64 diagnostics on the GPU (mog_diagnostics) and a 100-step trajectory
of the CPU-runnable configuration against the oracle.

Bars: momentum, kinetic energy and mass equal the oracle's FP64 sums to
1e-12 relative; the grid potential energy (FP64 transforms on both sides,
summed in different orders) to 1e-10.  The 100-step trajectory is chaotic
at cell faces, so it is compared statistically (fraction of particles
within 1e-4 of the oracle, drift of momentum and energy between the two,
L1 distance of the final density grids).
"""
import numpy as np
import pytest

from oracle import mog_oracle as O
from mog_solver_amd import workloads

from test_gpu_parity import dev, make_solver

pytestmark = pytest.mark.gpu


def _check_diag(got, want):
    scale_p = want["mass"] * np.sqrt(2 * max(want["kinetic"], 1e-300) / max(want["mass"], 1e-300))
    assert np.all(np.abs(got["momentum"] - want["momentum"]) <= 1e-12 * max(scale_p, 1e-300) + 1e-300)
    assert got["kinetic"] == pytest.approx(want["kinetic"], rel=1e-12, abs=1e-300)
    assert got["mass"] == pytest.approx(want["mass"], rel=1e-12, abs=1e-300)
    assert got["potential"] == pytest.approx(want["potential"], rel=1e-10, abs=1e-300)


@pytest.mark.parametrize("k,g,n", [(16, 16, 65536), (5, 8, 20000), (3, 32, 40000), (2, 64, 30000), (4, 2, 1000)])
def test_diagnostics_match_oracle(torch_cuda, k, g, n):
    torch = torch_cuda
    W, G = 100.0, 100.0
    rng = np.random.default_rng(k * 7 + g)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = np.mod(rng.normal(50, 15, (n, 3)), W)
    pos[:, 3] = rng.uniform(0.5, 2.0, n)
    vel = np.zeros_like(pos)
    vel[:, :3] = rng.normal(0, 30, (n, 3))
    s, rot, off = make_solver(k, g, W, G)
    got = s.diagnostics(dev(torch, pos), dev(torch, vel))
    want = O.diagnostics(pos, vel, rot, off, k, g, W, G)
    _check_diag(got, want)
    # the spare density grids are left all zero: the next step is unchanged
    p1, v1 = dev(torch, pos), dev(torch, vel)
    s.step(p1, v1)
    s2, _, _ = make_solver(k, g, W, G)
    p2, v2 = dev(torch, pos), dev(torch, vel)
    s2.step(p2, v2)
    torch.cuda.synchronize()
    assert torch.equal(p1, p2) and torch.equal(v1, v2)


def test_diagnostics_empty(torch_cuda):
    torch = torch_cuda
    s, _, _ = make_solver(4, 8)
    empty = torch.zeros((0, 4), dtype=torch.float32, device="cuda")
    d = s.diagnostics(empty, empty.clone())
    assert d["mass"] == 0.0 and d["kinetic"] == 0.0 and d["potential"] == 0.0
    assert np.all(d["momentum"] == 0.0)


def test_trajectory_c1_100_steps(torch_cuda):
    """C1 (n = 65,536 in U[0,100)^3, k = 16, g = 16, G = 100, dt = 1e-5,
    damping 0.999), 100 steps on the GPU and in the oracle (FP64 forces,
    float32 state).  Stated tolerances:

    * >= 99 % of particles within 1e-4 of the oracle's position;
    * momentum difference <= 1e-3 of sum m |v| (both sides);
    * |E_gpu - E_oracle| <= 1e-4 |U_oracle| at steps 25, 50, 75, 100 (the
      potential energy dominates: the cube starts cold);
    * the final density grids differ by <= 1e-3 in L1 (relative to the
      total deposited milli-mass).
    """
    torch = torch_cuda
    k, g, W, G, dt, damp = 16, 16, 100.0, 100.0, 1e-5, 0.999
    pos, vel = workloads.uniform_cube(65536, W, seed=1)
    s, rot, off = make_solver(k, g, W, G, dt, damp)
    p, v = dev(torch, pos), dev(torch, vel)
    cfg = dict(n_grids=k, grid_size=g, world_size=W, G=G, dt=dt, damping=damp)
    op, ov = pos.copy(), vel.copy()
    report = []
    for step in range(1, 101):
        s.step(p, v)
        op, ov, _ = O.step(op, ov, rot, off, cfg)
        if step % 25 == 0:
            got = s.diagnostics(p, v)
            want = O.diagnostics(op, ov, rot, off, k, g, W, G)
            e_got = got["kinetic"] + got["potential"]
            e_want = want["kinetic"] + want["potential"]
            mv = np.sum(op[:, 3] * np.linalg.norm(ov[:, :3].astype(np.float64), axis=1))
            dp = np.linalg.norm(got["momentum"] - want["momentum"]) / mv
            de = abs(e_got - e_want) / abs(want["potential"])
            report.append((step, dp, de))
            assert dp <= 1e-3, report
            assert de <= 1e-4, report
    gp = p.cpu().numpy()
    d = np.abs(gp[:, :3] - op[:, :3])
    d = np.minimum(d, W - d)
    within = np.mean(d.max(1) < 1e-4)
    grids_gpu = s.deposit(p).cpu().numpy()
    grids_cpu = O.deposit(op, rot, off, O.Params.make(k, g, W))
    l1 = np.abs(grids_gpu - grids_cpu).sum() / grids_cpu.sum()
    print(f"C1 100 steps: within 1e-4 {within:.5f}, grid L1 {l1:.2e}, (step, dp, dE) {report}")
    assert within >= 0.99
    assert l1 <= 1e-3
