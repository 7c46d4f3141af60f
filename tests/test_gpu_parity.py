"""GPU parity of the HIP force step against the CPU oracle.

Bars: density grids and the float32 state update are bit-exact; forces are
within 1e-5 relative RMS of the FP64 oracle (the pipeline's stated force
tolerance); force grids within 2e-6 relative RMS per grid.
"""
import numpy as np
import pytest

from oracle import mog_oracle as O
from mog_solver_amd.config import SolverConfig
from mog_solver_amd import workloads

pytestmark = pytest.mark.gpu

FORCE_TOL = 1e-5
GRID_TOL = 2e-6


def make_solver(k, g, W=100.0, G=100.0, dt=1e-5, damping=0.999, compat="reference", seed=0, precision="fp32",
                solver="fft"):
    from mog_solver_amd.solver import MogSolver

    cfg = SolverConfig.from_dict(
        dict(n_grids=k, grid_size=g, world_size=W, G=G, dt=dt, damping=damping, solver=solver, compat=compat,
             precision=precision)
    )
    rot, off = O.se3_transforms(k, g, W, seed)
    return MogSolver(cfg, rot=rot, off=off), rot, off


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def edge_positions(W, g, n_rand, seed):
    """Random positions plus values on and around cell faces and box edges."""
    rng = np.random.default_rng(seed)
    h = np.float32(W / g)
    special = [0.0, W, W / 2, np.nextafter(np.float32(W), np.float32(0)), 1e-7, -1e-7, -W / 3, 1.5 * W]
    special += list(np.arange(g + 1, dtype=np.float32) * h)
    special = np.array(special, dtype=np.float32)
    m = len(special)
    grid = np.stack(np.meshgrid(special, special[: m // 2], special[::3], indexing="ij"), -1).reshape(-1, 3)
    xyz = np.concatenate([rng.uniform(0, W, (n_rand, 3)).astype(np.float32), grid.astype(np.float32)])
    pos = np.zeros((xyz.shape[0], 4), np.float32)
    pos[:, :3] = xyz
    pos[:, 3] = 1.0
    return pos


@pytest.mark.parametrize("prepass", [-1, 1])
@pytest.mark.parametrize("k,g", [(16, 16), (5, 4), (7, 8), (3, 32), (2, 64), (2, 2)])
def test_deposit_bit_exact(torch_cuda, k, g, prepass):
    """Both forms of the LDS deposit: rows read from the positions, and
    rows prepared by the prepass (centred coordinates, milli-mass)."""
    torch = torch_cuda
    W = 100.0
    pos = edge_positions(W, g, 20000, seed=k * 100 + g)
    s, rot, off = make_solver(k, g, W)
    s.set_tuning(deposit_prepass=prepass)
    got = s.deposit(dev(torch, pos)).cpu().numpy()
    want = O.deposit(pos, rot, off, O.Params.make(k, g, W))
    assert got.dtype == np.int64 and got.shape == want.shape
    assert np.array_equal(got, want), f"{np.count_nonzero(got != want)} cells differ"


def test_deposit_c1_and_permutation(torch_cuda):
    torch = torch_cuda
    pos, _ = workloads.uniform_cube(65536, 100.0, seed=1)
    s, rot, off = make_solver(16, 16)
    want = O.deposit(pos, rot, off, O.Params.make(16, 16, 100.0))
    got = s.deposit(dev(torch, pos)).cpu().numpy()
    assert np.array_equal(got, want)
    perm = np.random.default_rng(7).permutation(pos.shape[0])
    got2 = s.deposit(dev(torch, pos[perm])).cpu().numpy()
    assert np.array_equal(got2, want)
    assert np.all(want.sum(axis=(1, 2, 3)) == 65536 * 1000)


@pytest.mark.parametrize("prepass", [-1, 1])
@pytest.mark.parametrize("k,g", [(12, 16), (3, 32)])
def test_deposit_masses_and_large_values(torch_cuda, k, g, prepass):
    """Fractional, zero, negative and very large masses (64-bit path), in
    the multi-grid (g = 16) and one-grid (g = 32) deposit tilings."""
    torch = torch_cuda
    rng = np.random.default_rng(11)
    n = 50000
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0, 100, (n, 3))
    pos[:, 3] = rng.uniform(0, 3, n).astype(np.float32)
    pos[:100, 3] = 0.0
    pos[100:200, 3] = -1.0
    pos[200:300, 3] = np.float32(5e9)  # milli-mass 5e12: beyond 32 bits
    pos[300:320, 3] = np.float32(np.nan)
    s, rot, off = make_solver(k, g)
    s.set_tuning(deposit_prepass=prepass)
    want = O.deposit(pos, rot, off, O.Params.make(k, g, 100.0))
    got = s.deposit(dev(torch, pos)).cpu().numpy()
    assert np.array_equal(got, want)


# FP32 solve against FP64 on both transform paths: the vector FFTs
# (spectral_mfma 0, the default) and the matrix-core DFTs (spectral_mfma 1,
# used from g = 16 up when selected, e.g. by the autotuner), with the
# Green's-function multiply, in both compat modes.
@pytest.mark.parametrize("g,compat,mfma", [(8, "reference", 0), (4, "reference", 0)] + [
    (g, compat, mfma) for g in (16, 32, 64) for compat in ("reference", "physical") for mfma in (0, 1)])
def test_solve_matches_fp64(torch_cuda, g, compat, mfma):
    torch = torch_cuda
    k = 3 if g < 64 else 1
    W, G = 100.0, 100.0
    rng = np.random.default_rng(g)
    n = 40000
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = np.mod(rng.normal(50, 12, (n, 3)), W)
    pos[:, 3] = 1.0
    s, rot, off = make_solver(k, g, W, G, compat=compat)
    s.set_tuning(spectral_mfma=mfma)
    assert s.get_tuning()["spectral_mfma"] == mfma
    P = O.Params.make(k, g, W)
    grids = O.deposit(pos, rot, off, P)
    got = s.solve(dev(torch, grids)).cpu().numpy().astype(np.float64)
    want = O.solve(grids, g, W, G, compat=compat)
    for i in range(k):
        a, b = got[i, ..., :3].reshape(-1, 3), want[i, ..., :3].reshape(-1, 3)
        err = O.rel_rms(a, b)
        assert err < GRID_TOL, f"grid {i}: rel rms {err}"
    mag = np.abs(got[..., 3] - want[..., 3]).max() / np.abs(want[..., 3]).max()
    assert mag < 1e-5
    if compat == "reference" and O.aliases_x_with_z(g):
        assert np.array_equal(got[..., 0], got[..., 2])


# g = 32 with more grids than CUs: the grid solve's persistent workgroups
# take several grids each and reuse their scratch slot (one per workgroup).
def test_solve_g32_more_grids_than_cus(torch_cuda):
    torch = torch_cuda
    k, g, W, G = 300, 32, 100.0, 100.0
    assert k > torch.cuda.get_device_properties(0).multi_processor_count
    rng = np.random.default_rng(32)
    n = 20000
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = np.mod(rng.normal(50, 12, (n, 3)), W)
    pos[:, 3] = 1.0
    s, rot, off = make_solver(k, g, W, G)
    grids = O.deposit(pos, rot, off, O.Params.make(k, g, W))
    got = s.solve(dev(torch, grids)).cpu().numpy().astype(np.float64)
    sample = np.arange(0, k, 7)  # among them grids a workgroup takes second
    want = O.solve(grids[sample], g, W, G)
    for j, i in enumerate(sample):
        err = O.rel_rms(got[i, ..., :3].reshape(-1, 3), want[j, ..., :3].reshape(-1, 3))
        assert err < GRID_TOL, f"grid {i}: rel rms {err}"


def clustered_grids(k, g, W, seed, n=40000):
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = np.mod(rng.normal(50, 12, (n, 3)), W)
    pos[:, 3] = 1.0
    return pos


# solver: convolution.  compat physical: the real-space Green's-gradient
# convolution (main.py:250-278 kernel) against the FP64 oracle, on the VALU
# (g < 16) and matrix-core (g >= 16) transforms; compat reference: the key
# runs the FFT solver, bit-identically.
@pytest.mark.parametrize("g,k,mfma", [(8, 4, False), (16, 3, True), (16, 3, False), (32, 2, True), (64, 1, True),
                                      (32, 2, False), (64, 1, False)])
def test_convolution_solve_matches_fp64(torch_cuda, g, k, mfma):
    torch = torch_cuda
    W, G = 100.0, 100.0
    pos = clustered_grids(k, g, W, seed=g + k)
    s, rot, off = make_solver(k, g, W, G, compat="physical", solver="convolution")
    s.set_tuning(spectral_mfma=int(mfma))
    grids = O.deposit(pos, rot, off, O.Params.make(k, g, W))
    got = s.solve(dev(torch, grids)).cpu().numpy().astype(np.float64)
    want = O.solve(grids, g, W, G, compat="physical", solver="convolution")
    for i in range(k):
        err = O.rel_rms(got[i, ..., :3].reshape(-1, 3), want[i, ..., :3].reshape(-1, 3))
        assert err < GRID_TOL, f"grid {i}: rel rms {err}"
    fft = O.solve(grids, g, W, G, compat="physical")
    assert O.rel_rms(want[0, ..., :3].reshape(-1, 3), fft[0, ..., :3].reshape(-1, 3)) > 10 * GRID_TOL


def test_convolution_key_under_reference_compat_is_fft(torch_cuda):
    torch = torch_cuda
    k, g, W = 4, 16, 100.0
    pos = dev(torch, clustered_grids(k, g, W, seed=3, n=50000))
    a, _, _ = make_solver(k, g, W, solver="convolution")
    b, _, _ = make_solver(k, g, W, solver="fft")
    fa, fb = a.forces(pos).cpu().numpy(), b.forces(pos).cpu().numpy()
    assert np.array_equal(fa, fb)


def test_convolution_forces_match_oracle(torch_cuda):
    torch = torch_cuda
    k, g, W, G = 6, 16, 100.0, 100.0
    pos = clustered_grids(k, g, W, seed=8, n=30000)
    s, rot, off = make_solver(k, g, W, G, compat="physical", solver="convolution")
    got = s.forces(dev(torch, pos)).cpu().numpy()
    want, _ = O.forces(pos, rot, off, k, g, W, G, compat="physical", solver="convolution")
    assert O.rel_rms(got, want) < FORCE_TOL


# Split-precision matrix-core DFT: per-grid force error against FP64 and the
# particle-force error after averaging over k grids.  Bounds: BF16X3 keeps
# ~16 mantissa bits per operand (hi + lo), so a grid's force field carries a
# few 1e-5 relative RMS and the k-average stays below 1e-5; scaled FP16 keeps
# 11 bits, a few 1e-4 per grid.
@pytest.mark.parametrize("g,precision,grid_tol,force_tol", [
    (16, "bf16x3", 1e-4, 1e-5), (32, "bf16x3", 1e-4, 1e-5),
    (16, "fp16", 5e-3, 2e-3), (32, "fp16", 5e-3, 2e-3)])
def test_split_precision_dft(torch_cuda, g, precision, grid_tol, force_tol):
    torch = torch_cuda
    k, W, G = 24, 100.0, 100.0
    pos, _ = workloads.plummer_spheres(60000, n_spheres=8, scale=5.0, world_size=W, seed=g)
    s, rot, off = make_solver(k, g, W, G, precision=precision)
    P = O.Params.make(k, g, W)
    grids = O.deposit(pos, rot, off, P)
    got = s.solve(dev(torch, grids)).cpu().numpy().astype(np.float64)
    want = O.solve(grids, g, W, G)
    errs = [O.rel_rms(got[i, ..., :3].reshape(-1, 3), want[i, ..., :3].reshape(-1, 3)) for i in range(k)]
    assert max(errs) < grid_tol, errs
    f = s.forces(dev(torch, pos)).cpu().numpy()
    fw, _ = O.forces(pos, rot, off, k, g, W, G)
    err = O.rel_rms(f[:, :3], fw)
    print(f"{precision} g={g}: grid rel rms max {max(errs):.2e}, particle force rel rms {err:.2e}")
    assert err < force_tol


def test_gather_matches_fp64(torch_cuda):
    torch = torch_cuda
    k, g, W = 24, 16, 100.0
    pos, _ = workloads.uniform_cube(30000, W, seed=5)
    s, rot, off = make_solver(k, g, W)
    rng = np.random.default_rng(3)
    force = rng.normal(size=(k, g, g, g, 4)).astype(np.float32)
    got = s.gather(dev(torch, pos), dev(torch, force)).cpu().numpy()
    want = O.gather(pos, rot, off, force.astype(np.float64), O.Params.make(k, g, W))
    assert np.all(got[:, 3] == 0)
    assert O.rel_rms(got[:, :3], want) < 1e-6


@pytest.mark.parametrize("prepass", [-1, 1])
@pytest.mark.parametrize("k,g", [(5, 16), (3, 32)])
def test_deposit_far_positions(torch_cuda, k, g, prepass):
    """Positions far outside the box (|w| beyond the range of the fast
    cell's float fields) next to ordinary ones: the exact sequence takes them,
    and the fast cell a pair was first added to is taken back where it was
    added."""
    torch = torch_cuda
    W = 100.0
    rng = np.random.default_rng(21)
    pos = edge_positions(W, g, 20000, seed=g)
    far = np.array([3e9, -2e8, 1e7, -5e5, 2.5e4, -333.25], np.float32)
    idx = rng.integers(0, pos.shape[0], size=(3000, 3))
    pos[idx[:, 0], 0] = far[rng.integers(0, far.size, 3000)]
    pos[idx[:, 1], 1] = far[rng.integers(0, far.size, 3000)]
    pos[idx[:, 2], 2] = far[rng.integers(0, far.size, 3000)]
    pos[::7, 3] = np.float32(3e6)  # and some masses above the 32-bit partials
    s, rot, off = make_solver(k, g, W)
    s.set_tuning(deposit_prepass=prepass)
    got = s.deposit(dev(torch, pos)).cpu().numpy()
    want = O.deposit(pos, rot, off, O.Params.make(k, g, W))
    assert np.array_equal(got, want), f"{np.count_nonzero(got != want)} cells differ"


@pytest.mark.parametrize("k,g,shift", [(3, 32, 40.0), (5, 16, 300.0)])
def test_deposit_and_gather_with_far_offsets(torch_cuda, k, g, shift):
    """Offsets moved by a multiple of W (same cells up to rounding, which the
    exact sequence reproduces): the fast-path bound |w| then exceeds the
    range of the shifted cell fields (2^(22 - s0 - 2 lg) with s0 = 2 for the
    LDS deposit, 4 for the gather), so both kernels take their index form."""
    torch = torch_cuda
    W = 100.0
    s, rot, off = make_solver(k, g, W)
    off = off.copy()
    off[:, :3] += np.float32(shift * W)
    s.set_transforms(rot, off)
    lg = int(np.log2(g))
    wbound = (4.5 * W + 0.5 * W + float(np.abs(off[:, :3]).max())) / (W / g)
    assert wbound + 1.5 > 2.0 ** (22 - 4 - 2 * lg)  # gather index form
    if shift > 100:
        assert wbound + 1.5 > 2.0 ** (22 - 2 - 2 * lg)  # deposit index form
    pos = edge_positions(W, g, 20000, seed=k + g)
    got = s.deposit(dev(torch, pos)).cpu().numpy()
    want = O.deposit(pos, rot, off, O.Params.make(k, g, W))
    assert np.array_equal(got, want), f"{np.count_nonzero(got != want)} cells differ"
    rng = np.random.default_rng(4)
    force = rng.normal(size=(k, g, g, g, 4)).astype(np.float32)
    gf = s.gather(dev(torch, pos), dev(torch, force)).cpu().numpy()
    wf = O.gather(pos, rot, off, force.astype(np.float64), O.Params.make(k, g, W))
    assert O.rel_rms(gf[:, :3], wf) < 1e-6


def test_update_bit_exact(torch_cuda):
    torch = torch_cuda
    rng = np.random.default_rng(9)
    n = 100000
    W = 100.0
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(-5, 105, (n, 3))
    pos[:, 3] = rng.uniform(0.5, 2, n)
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(0, 3e3, (n, 3))
    avg = np.zeros((n, 4), np.float32)
    avg[:, :3] = rng.normal(0, 1e6, (n, 3))
    s, _, _ = make_solver(4, 16, W, dt=1e-5, damping=0.999)
    p, v = dev(torch, pos), dev(torch, vel)
    s.update(p, v, dev(torch, avg))
    wp, wv = O.update32(pos, vel, avg, 1e-5, 0.999, W)
    assert np.array_equal(p.cpu().numpy(), wp)
    assert np.array_equal(v.cpu().numpy(), wv)


@pytest.mark.parametrize("compat", ["reference", "physical"])
def test_forces_c1_within_tolerance(torch_cuda, compat):
    torch = torch_cuda
    k, g, W, G = 16, 16, 100.0, 100.0
    pos, _ = workloads.uniform_cube(65536, W, seed=1)
    s, rot, off = make_solver(k, g, W, G, compat=compat)
    got = s.forces(dev(torch, pos)).cpu().numpy()
    want, _ = O.forces(pos, rot, off, k, g, W, G, compat=compat)
    err = O.rel_rms(got[:, :3], want)
    assert err < FORCE_TOL, err


@pytest.mark.parametrize("k,g", [(16, 16), (4, 32), (2, 64), (24, 8)])
def test_step_matches_phases_and_graph(torch_cuda, k, g):
    """mog_step == deposit+solve+gather+update, and graph replay == eager."""
    torch = torch_cuda
    W = 100.0
    pos, vel = workloads.uniform_cube(65536, W, seed=1)
    vel[:, :3] = np.random.default_rng(4).normal(0, 50, (pos.shape[0], 3))
    s, rot, off = make_solver(k, g, W)
    p1, v1 = dev(torch, pos), dev(torch, vel)
    s.step(p1, v1)
    p2, v2 = dev(torch, pos), dev(torch, vel)
    avg = s.forces(p2)
    s.update(p2, v2, avg)
    assert torch.equal(p1, p2) and torch.equal(v1, v2)
    # three eager steps vs three graph replays, checked after each
    p3, v3 = dev(torch, pos), dev(torch, vel)
    p4, v4 = dev(torch, pos), dev(torch, vel)
    for _ in range(3):
        s.step(p3, v3)
        s.run(p4, v4, 1, use_graph=True)
        torch.cuda.synchronize()
        assert torch.equal(p3, p4) and torch.equal(v3, v4)


@pytest.mark.parametrize("resort,prepass", [(1, 0), (2, 0), (5, 0), (2, 1), (5, 1)])
def test_run_state_across_reorders(torch_cuda, resort, prepass):
    """mog_run keeps the particle state in visiting order (imported at the
    start and after each rebuild of the order, velocities written back at
    each rebuild and at the end): bit-identical to single steps through the
    caller's arrays, with the order rebuilt inside a run, eager and graph,
    across successive calls.  With the deposit's prepass (forced on), the
    graph replays' gather writes the next step's prepared rows and the
    positions go back with the velocities."""
    torch = torch_cuda
    k, g, W = 12, 16, 100.0
    pos, vel = workloads.two_disks(80000, W, 100.0, g, seed=8)
    s, rot, off = make_solver(k, g, W)
    s.set_tuning(**{**s.get_tuning(), "resort_every": resort, "sorted_deposit": 0, "cell_cache": 0,
                    "deposit_prepass": prepass})
    p1, v1 = dev(torch, pos), dev(torch, vel)
    for _ in range(7):
        s.step(p1, v1)
    p2, v2 = dev(torch, pos), dev(torch, vel)
    s.run(p2, v2, 7, use_graph=True)
    p3, v3 = dev(torch, pos), dev(torch, vel)
    s.run(p3, v3, 3, use_graph=False)
    s.run(p3, v3, 4, use_graph=True)
    torch.cuda.synchronize()
    assert torch.equal(p1, p2) and torch.equal(v1, v2)
    assert torch.equal(p1, p3) and torch.equal(v1, v3)


def test_particle_count_changes_between_steps(torch_cuda):
    """The per-step cell cache grows with n; captured graphs are rebuilt."""
    torch = torch_cuda
    k, g, W = 8, 16, 100.0
    s, rot, off = make_solver(k, g, W)
    for n in (1000, 50000, 3000, 70000):
        pos, vel = workloads.uniform_cube(n, W, seed=n)
        pa, va = dev(torch, pos), dev(torch, vel)
        pb, vb = dev(torch, pos), dev(torch, vel)
        s.run(pa, va, 2, use_graph=True)
        avg = s.forces(pb)
        s.update(pb, vb, avg)
        avg = s.forces(pb)
        s.update(pb, vb, avg)
        torch.cuda.synchronize()
        assert torch.equal(pa, pb) and torch.equal(va, vb), n


@pytest.mark.parametrize("g", [16, 32])
def test_tuning_variants_agree(torch_cuda, g):
    """Every tuning choice gives the same step: deposit/gather/visiting-order
    variants bit-exactly, the vector DFT within FP32 rounding."""
    torch = torch_cuda
    k, W = 12, 100.0
    pos, vel = workloads.two_disks(80000, W, 100.0, g, seed=6)
    s, rot, off = make_solver(k, g, W)
    base = s.get_tuning()
    variants = [
        dict(),
        dict(resort_every=0),
        dict(resort_every=3, sorted_deposit=1),
        dict(resort_every=2, sorted_deposit=1, cell_cache=1),
        dict(sorted_deposit=1, cell_cache=1, deposit_grids_per_wg=1, deposit_threads=256),
        dict(deposit_grids_per_wg=4, deposit_threads=1024, deposit_wg_target=300),
        dict(deposit_grids_per_wg=8, deposit_threads=64, deposit_wg_target=50),
    ]
    results = []
    for v in variants:
        s.set_tuning(**{**base, **v})
        p, vv = dev(torch, pos), dev(torch, vel)
        s.run(p, vv, 2, use_graph=True)
        s.step(p, vv)
        torch.cuda.synchronize()
        results.append((p.cpu().numpy(), vv.cpu().numpy()))
    for (p, vv), v in zip(results[1:], variants[1:]):
        assert np.array_equal(p, results[0][0]) and np.array_equal(vv, results[0][1]), v
    # the other DFT path (matrix cores when the default is the vector FFT):
    # FP32 rounding differences only
    flipped = 1 - base["spectral_mfma"]
    s.set_tuning(**{**base, "spectral_mfma": flipped})
    assert s.get_tuning()["spectral_mfma"] == flipped
    p, vv = dev(torch, pos), dev(torch, vel)
    s.run(p, vv, 2, use_graph=True)
    s.step(p, vv)
    got_v = vv.cpu().numpy()[:, :3]
    assert not np.array_equal(got_v, results[0][1][:, :3])  # a different path ran
    assert O.rel_rms(got_v, results[0][1][:, :3]) < 1e-5


def test_autotune_keeps_the_step(torch_cuda):
    """The autotuner runs on copies of the state and picks among tunings that
    give the same step: after tuning, a step is bit-identical to the default
    tuning's (with the DFT path the tuner chose, the one choice that changes
    rounding)."""
    from mog_solver_amd.autotune import autotune

    torch = torch_cuda
    k, g, W = 8, 16, 100.0
    pos, vel = workloads.two_disks(40000, W, 100.0, g, seed=3)
    s, _, _ = make_solver(k, g, W)
    base = s.get_tuning()
    p0, v0 = dev(torch, pos), dev(torch, vel)
    res = autotune(s, p0, v0, steps=2, cache_path=None)
    assert res["evaluated"] > 10 and not res["cached"] and res["ms_per_step"] > 0
    tuned = s.get_tuning()
    assert tuned == res["tuning"]
    assert np.array_equal(p0.cpu().numpy(), pos) and np.array_equal(v0.cpu().numpy(), vel)
    out = []
    for t in (tuned, {**base, "spectral_mfma": tuned["spectral_mfma"]}):
        s.set_tuning(**t)
        p, v = dev(torch, pos), dev(torch, vel)
        s.run(p, v, 3, use_graph=True)
        out.append((p.cpu().numpy(), v.cpu().numpy()))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_tuning_rejects_invalid(torch_cuda):
    from mog_solver_amd._abi import MogError

    s, _, _ = make_solver(4, 16)
    for bad in (dict(deposit_threads=100), dict(deposit_grids_per_wg=3), dict(deposit_wg_target=0),
                dict(resort_every=-1), dict(cell_cache=1, sorted_deposit=0), dict(spectral_mfma=2)):
        with pytest.raises(MogError):
            s.set_tuning(**bad)


def test_trajectory_c1_10_steps(torch_cuda):
    """10 steps of the CPU-runnable config vs the oracle (FP64 forces,
    float32 state).  Per-particle agreement is statistical: a one-ulp force
    difference can move a particle across a cell face."""
    torch = torch_cuda
    k, g, W, G, dt, damp = 16, 16, 100.0, 100.0, 1e-5, 0.999
    pos, vel = workloads.uniform_cube(65536, W, seed=1)
    s, rot, off = make_solver(k, g, W, G, dt, damp)
    p, v = dev(torch, pos), dev(torch, vel)
    cfg = dict(n_grids=k, grid_size=g, world_size=W, G=G, dt=dt, damping=damp)
    op, ov = pos.copy(), vel.copy()
    for _ in range(10):
        s.step(p, v)
        op, ov, _ = O.step(op, ov, rot, off, cfg)
    gp, gv = p.cpu().numpy(), v.cpu().numpy()
    d = np.abs(gp[:, :3] - op[:, :3])
    d = np.minimum(d, W - d)
    assert np.mean(d.max(1) < 1e-4) > 0.999
    assert O.rel_rms(gv[:, :3], ov[:, :3]) < 1e-4
    assert np.array_equal(gp[:, 3], op[:, 3]) and np.all(gv[:, 3] == 0)


def test_empty_and_single(torch_cuda):
    torch = torch_cuda
    s, rot, off = make_solver(4, 8)
    empty = torch.zeros((0, 4), dtype=torch.float32, device="cuda")
    s.step(empty, empty.clone())
    grids = s.deposit(empty)
    assert int(grids.abs().sum()) == 0
    one = np.array([[10.0, 20.0, 30.0, 2.5]], np.float32)
    got = s.deposit(dev(torch, one)).cpu().numpy()
    want = O.deposit(one, rot, off, O.Params.make(4, 8, 100.0))
    assert np.array_equal(got, want) and got.sum() == 4 * 2500
    # a single particle feels no force from itself (odd kernel, DC removed)
    f = s.forces(dev(torch, one)).cpu().numpy()
    wf, _ = O.forces(one, rot, off, 4, 8, 100.0, 100.0)
    assert np.allclose(f[:, :3], wf, atol=1e-3 * max(1.0, np.abs(wf).max()))
