"""Initial-condition generators: the CPU twin of the Perlin-fBm sampler and
the seeded two-galaxy state (no GPU)."""
import numpy as np
import pytest

from oracle import init_oracle as I
from mog_solver_amd import workloads


def py_hash(x):
    """Scalar restatement of the 32-bit hash with Python integers."""
    m = 0xFFFFFFFF
    x &= m
    x ^= x >> 16
    x = (x * 0x85EBCA6B) & m
    x ^= x >> 13
    x = (x * 0xC2B2AE35) & m
    x ^= x >> 16
    return x


def test_hash_matches_scalar_integers():
    xs = np.random.default_rng(0).integers(0, 2**32, 2000, dtype=np.uint64).astype(np.uint32)
    got = I.ihash(xs)
    assert all(int(g) == py_hash(int(x)) for g, x in zip(got, xs))
    assert int(I.ihash(np.uint32(0))) == 0
    # avalanche: one input bit flips about half of the output bits
    a = I.ihash(xs)
    b = I.ihash(xs ^ np.uint32(1))
    flips = np.unpackbits((a ^ b).view(np.uint8)).mean()
    assert 0.4 < flips < 0.6


def test_rand_range_and_uniformity():
    r = I.urand(np.arange(200000, dtype=np.uint32))
    assert r.min() >= 0 and r.max() <= 1
    hist, _ = np.histogram(r, bins=10, range=(0, 1))
    assert np.all(np.abs(hist / len(r) - 0.1) < 0.005)


def test_sampler_geometry_and_density_bias():
    W, buf = 100.0, 0.1
    pos, vel, attempts = I.init_perlin(20000, W, 4.0, 2.0, seed=12345, spawn_buffer=buf)
    radius = (1 - 2 * buf) * W / 2
    d = np.linalg.norm(pos[:, :3] - W / 2, axis=1)
    assert d.max() <= radius * (1 + 1e-5)
    assert np.all(pos[:, 3] == 1.0) and np.all(vel[:, 3] == 0)
    assert np.all(np.abs(vel[:, :3]) <= W / 2)
    assert np.mean(attempts < 64) > 0.99
    # accepted points sit where the density field is high
    ns = np.float32(4.0 / W)
    f_acc = I.fbm(*(pos[:, k].astype(np.float32) * ns for k in range(3)), 12345)
    rng = np.random.default_rng(1)
    u = rng.normal(size=(20000, 3))
    u = (u / np.linalg.norm(u, axis=1, keepdims=True)) * (radius * rng.uniform(0, 1, (20000, 1)) ** (1 / 3))
    f_uni = I.fbm(*((u[:, k] + W / 2).astype(np.float32) * ns for k in range(3)), 12345)
    assert f_acc.mean() > f_uni.mean() + 0.05


def test_cube_mode_bounds():
    W, buf = 50.0, 0.2
    pos, _, _ = I.init_perlin(5000, W, 2.0, 1.5, seed=7, spawn_buffer=buf, spawn_shape=0)
    assert pos[:, :3].min() >= buf * W - 1e-4 and pos[:, :3].max() <= (1 - buf) * W + 1e-4


def test_two_galaxies_reference_seeded():
    a = workloads.two_galaxies_reference(4000, 100.0, 2.0, seed=5)
    b = workloads.two_galaxies_reference(4000, 100.0, 2.0, seed=5)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    pos, vel = a
    assert np.all(pos[:, 3] == 2.0)
    c1 = np.median(pos[:2000, :3], axis=0)
    c2 = np.median(pos[2000:, :3], axis=0)
    assert np.allclose(c1, [30, 50, 50], atol=2.0) and np.allclose(c2, [70, 50, 50], atol=2.0)
    assert np.allclose(vel[:2000, :3].mean(0), [0.3, 0.1, 0.0], atol=0.02)
    # spin speed 0.3 sqrt(r/W + 0.01) about the fixed axis
    r = np.linalg.norm(pos[:2000, :3] - [30, 50, 50], axis=1)
    speed = np.linalg.norm(vel[:2000, :3] - [0.3, 0.1, 0.0], axis=1)
    assert np.allclose(speed, 0.3 * np.sqrt(r / 100.0 + 0.01), rtol=1e-4)


def test_snapshot_roundtrip(tmp_path):
    from mog_solver_amd.snapshot import load_snapshot, save_snapshot
    from mog_solver_amd.transforms import compute_transforms

    pos, vel = workloads.uniform_cube(100, 10.0, seed=2)
    off, rot = compute_transforms(4, 8, 10.0)
    cfg = {"n_grids": 4, "grid_size": 8, "world_size": 10.0, "G": 1.0, "dt": "1e-5", "record": True}
    save_snapshot(tmp_path / "a.npz", pos, vel, rot, off, 17, cfg)
    s = load_snapshot(tmp_path / "a.npz")
    assert np.array_equal(s["pos"], pos) and np.array_equal(s["vel"], vel)
    assert np.array_equal(s["rot"], rot) and np.array_equal(s["off"], off)
    assert s["step"] == 17 and s["config"] == cfg
    np.savez(tmp_path / "bad.npz", pos=pos)
    with pytest.raises(Exception):
        load_snapshot(tmp_path / "bad.npz")


def test_plummer_ensemble_shapes_and_ranges():
    from mog_solver_amd import workloads as w

    ens = w.plummer_ensemble(seed=4)
    assert len(ens) == 64
    ns = np.array([c["num_particles"] for c, _, _ in ens])
    assert ns.min() >= 1e4 and ns.max() <= 5e5
    assert {c["n_grids"] for c, _, _ in ens} <= {32, 64, 128, 240}
    assert {c["grid_size"] for c, _, _ in ens} == {16, 32}
    for c, p, v in ens[:5]:
        assert p.shape == (c["num_particles"], 4) and v.shape == p.shape
        assert p.dtype == np.float32 and np.all(p[:, :3] >= 0) and np.all(p[:, :3] < 100)
        r = np.linalg.norm(p[:, :3].astype(np.float64) - 50.0, axis=1)
        # Plummer: half-mass radius a / sqrt(2^(2/3) - 1) ~ 1.305 a
        assert abs(np.median(r) / (5.0 * 1.305) - 1) < 0.1
    again = w.plummer_ensemble(seed=4)
    assert all(np.array_equal(a[1], b[1]) for a, b in zip(ens[:3], again[:3]))
