"""Pins of the closure oracle (no GPU)."""
import numpy as np

from oracle import closure_oracle as C
from oracle import mog_oracle as O


def _batch(n=40, seed=0):
    rng = np.random.default_rng(seed)
    feat = np.zeros((n, 12))
    feat[:, :C.K_IN] = rng.normal(size=(n, C.K_IN))
    feat[:, C.S_COL] = rng.uniform(0.5, 3.0, n)
    mean = rng.normal(size=(n, 3))
    target = mean + rng.normal(size=(n, 3))
    return feat, mean, target


def test_gradient_matches_finite_differences():
    params = C.init_params(1, out_scale=1.0).astype(np.float64)
    feat, mean, target = _batch()
    wnorm = 1.0 / np.sum(feat[:, C.S_COL] ** 2)
    loss, grad = C.loss_and_grad(params, feat, mean, target, wnorm)
    rng = np.random.default_rng(2)
    idx = rng.choice(C.N_PARAMS, 60, replace=False)
    eps = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += eps
        p2[i] -= eps
        fd = (C.loss_and_grad(p1, feat, mean, target, wnorm)[0] - C.loss_and_grad(p2, feat, mean, target, wnorm)[0]) / (2 * eps)
        assert abs(fd - grad[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, grad[i])


def test_zero_network_predicts_the_mean_and_loss_is_relative_error():
    params = np.zeros(C.N_PARAMS)
    feat, mean, target = _batch()
    pred = C.predict(params, feat, mean)
    assert np.allclose(pred, mean)
    wnorm = 1.0 / np.sum(feat[:, C.S_COL] ** 2)
    loss, _ = C.loss_and_grad(params, feat, mean, target, wnorm)
    # loss = sum |target - mean|^2 / sum S^2
    assert np.isclose(loss, np.sum((target - mean) ** 2) * wnorm)


def test_features_ngp_part_is_the_mog_force():
    from mog_solver_amd import workloads

    pos, _ = workloads.uniform_cube(3000, 100.0, seed=4)
    rot, off = O.se3_transforms(4, 8, 100.0, 0)
    rho0 = 3000 / 8**3
    feat, c = C.features(pos, rot, off, 4, 8, 100.0, 100.0, rho0)
    avg, _ = O.forces(pos, rot, off, 4, 8, 100.0, 100.0)
    S = feat[:, C.S_COL:C.S_COL + 1]
    assert np.allclose(feat[:, :3] * S, avg, rtol=1e-12, atol=1e-9 * np.abs(avg).max())
    assert np.allclose(feat[:, 6:9] * S, c)
    assert np.all(feat[:, 3:6] >= 0) and np.all(S > 0) and np.all(feat[:, 11] == 0)
    # S = |c| + |s|
    assert np.allclose(S[:, 0], np.linalg.norm(c, axis=1) + np.linalg.norm(feat[:, 3:6] * S, axis=1))
    # the cloud-in-cell mean tracks the NGP mean (same grids, smoother field)
    assert np.linalg.norm(c - avg) < 0.5 * np.linalg.norm(avg)
    assert C.N_PARAMS == 10 * 64 + 64 + 2 * (64 * 64 + 64) + 64 * 3 + 3


def _mode_grid(g, W, G, m, rho0=50.0, eps=0.3):
    """Milli-mass grid of the cell integrals of rho0 (1 + eps cos(k.l)),
    k = 2 pi m / W, and the analytic force of that density,
    F(l) = 2 pi G A k sin(k.l) / |k|^2 with A = eps rho0 h^3 the mode's
    amplitude in mass per cell (the solve's units: grid / 1000)."""
    h = W / g
    k = 2 * np.pi * np.asarray(m, dtype=np.float64) / W
    idx = np.arange(g) + 0.5
    lz, ly, lx = np.meshgrid(idx * h, idx * h, idx * h, indexing="ij")
    win = np.prod([np.sinc(kd * h / 2 / np.pi) for kd in k])
    cell = rho0 * h**3 * (1 + eps * win * np.cos(k[0] * lx + k[1] * ly + k[2] * lz))
    grid = np.rint(cell * 1000).astype(np.int64)
    A = eps * rho0 * h**3

    def force(l):
        ph = l @ k
        return (2 * np.pi * G * A / (k @ k)) * np.sin(ph)[:, None] * k[None, :]

    return grid, force, A


def test_deconvolved_force_is_the_analytic_force_at_cell_centres():
    """Cell integrals of a single Fourier mode carry the NGP window
    sinc(k_x h/2) sinc(k_y h/2) sinc(k_z h/2); dividing it out recovers the
    force of the continuous density at the cell centres up to the milli-mass
    rounding.  Distinct per-axis modes catch a window taken on the wrong
    axis; the plain solve is off by the window."""
    g, W, G = 16, 100.0, 100.0
    h = W / g
    for m in ((1, 2, 3), (4, 0, 1), (0, 5, 2)):
        grid, force, A = _mode_grid(g, W, G, m)
        D = C.deconvolved_force(grid, g, W, G, compat="physical")
        i = np.arange(g)
        iz, iy, ix = np.meshgrid(i, i, i, indexing="ij")
        l = np.stack([ix.ravel(), iy.ravel(), iz.ravel()], 1) * h + h / 2
        want = force(l).reshape(g, g, g, 3)
        scale = np.abs(want).max()
        assert np.abs(D - want).max() < 1e-6 * scale, m
        plain = O.solve(grid[None], g, W, G, compat="physical")[0][..., :3]
        assert np.abs(plain - want).max() > 1e-3 * scale, m


def test_deconvolved_force_keeps_the_reference_x_alias():
    g, W, G = 16, 100.0, 100.0
    grid, _, _ = _mode_grid(g, W, G, (1, 2, 3))
    D = C.deconvolved_force(grid, g, W, G, compat="reference")
    assert np.array_equal(D[..., 0], D[..., 2])
    D8 = C.deconvolved_force(_mode_grid(8, W, G, (1, 2, 3))[0], 8, W, G, compat="reference")
    assert not np.allclose(D8[..., 0], D8[..., 2])  # log2(8) odd: no alias


def test_cic_interpolation():
    rng = np.random.default_rng(3)
    g = 8
    f = rng.normal(size=(g, g, g, 3))
    cells = rng.integers(0, g, size=(50, 3))
    # cell centres (u integer, also shifted by whole periods) give the cell value
    for shift in (0, g, -2 * g):
        got = C.cic(f, (cells + shift).astype(np.float64), g)
        assert np.array_equal(got, f[cells[:, 2], cells[:, 1], cells[:, 0]])
    # half-way along x: the mean of the two neighbours (periodic)
    u = cells + np.array([0.5, 0.0, 0.0])
    want = 0.5 * (f[cells[:, 2], cells[:, 1], cells[:, 0]] + f[cells[:, 2], cells[:, 1], (cells[:, 0] + 1) % g])
    assert np.allclose(C.cic(f, u, g), want)
    # reproduces any trilinear function exactly inside one period
    a = rng.normal(size=(3, 3))
    i = np.arange(g)
    iz, iy, ix = np.meshgrid(i, i, i, indexing="ij")
    lin = np.stack([ix, iy, iz], -1).astype(np.float64) @ a
    u = rng.uniform(0, g - 1, size=(200, 3))
    assert np.allclose(C.cic(lin, u, g), u @ a)


def test_cell_coords_match_the_ngp_cells():
    """floor(u + 1/2) is the NGP cell on every axis: pins the half-cell
    offset, the frame (rot * c, not its transpose) and the sign of off."""
    from mog_solver_amd import workloads

    g, W = 16, 100.0
    pos, _ = workloads.uniform_cube(20000, W, seed=9)
    rot, off = O.se3_transforms(3, g, W, 5)
    P = O.Params.make(3, g, W)
    c = O.centered32(pos, P)
    for i in range(3):
        u = C.cell_coords(c.astype(np.float64), rot[i], off[i], W, g)
        ngp = np.floor(u + 0.5).astype(np.int64) & (g - 1)
        flat = (ngp[:, 2] * g + ngp[:, 1]) * g + ngp[:, 0]
        idx = O.cell_indices_one(c, rot[i], off[i], P)
        # float32 rounding of the reference's cell arithmetic moves a few
        # particles that sit on a face
        assert np.mean(flat == idx) > 0.999


def test_cic_of_the_deconvolved_force_follows_a_smooth_mode():
    """At arbitrary positions the trilinear interpolation of the deconvolved
    field matches the analytic force of a smooth mode to the interpolation
    error (kh)^2/8; a half-cell offset error would cost ~kh/2."""
    g, W, G = 32, 100.0, 100.0
    h = W / g
    grid, force, _ = _mode_grid(g, W, G, (1, 1, 0))
    D = C.deconvolved_force(grid, g, W, G, compat="physical")
    rng = np.random.default_rng(1)
    l = rng.uniform(0, W, size=(5000, 3))
    got = C.cic(D, l / h - 0.5, g)
    want = force(l)
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    kh = 2 * np.pi * np.sqrt(2) / W * h
    assert err < kh**2 / 8, err
    shifted = C.cic(D, l / h, g)
    assert np.linalg.norm(shifted - want) / np.linalg.norm(want) > 3 * kh**2 / 8
