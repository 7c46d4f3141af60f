"""Pins of the 2D oracle (no GPU): closed forms, brute force, invariants."""
import numpy as np
import pytest

from oracle import mog2d_oracle as O2
from oracle import mog_oracle as O

W, G = 100.0, 100.0


@pytest.mark.parametrize("g,mx,my", [(16, 1, 1), (16, 3, 2), (32, 5, 1), (8, 2, 3)])
def test_single_mode_closed_form(g, mx, my):
    """rho = A cos(kx x) cos(ky y) (mass per cell) -> F_x = 2 pi G A kx /
    (kx^2 + ky^2) sin(kx x) cos(ky y), F_y likewise, at the cell positions."""
    A = 2.5
    h = W / g
    x = np.arange(g) * h
    kx, ky = 2 * np.pi * mx / W, 2 * np.pi * my / W
    rho = A * np.cos(ky * x)[:, None] * np.cos(kx * x)[None, :]  # [y][x]
    F = O2.solve_2d(1000.0 * rho[None], g, W, G)[0]
    c = 2 * np.pi * G * A / (kx * kx + ky * ky)
    fx = c * kx * np.sin(kx * x)[None, :] * np.cos(ky * x)[:, None]
    fy = c * ky * np.cos(kx * x)[None, :] * np.sin(ky * x)[:, None]
    np.testing.assert_allclose(F[..., 0], fx, atol=1e-9 * abs(c * kx))
    np.testing.assert_allclose(F[..., 1], fy, atol=1e-9 * abs(c * ky))


def test_solve_matches_brute_force_dft():
    g = 4
    rng = np.random.default_rng(3)
    grids = rng.integers(0, 9000, (2, g, g)).astype(np.int64)
    F = O2.solve_2d(grids, g, W, G)
    j = np.arange(g)
    E = np.exp(-2j * np.pi * np.outer(j, j) / g)  # forward DFT matrix
    kv = O2.wave_numbers(g, W)
    for i in range(2):
        rho_hat = E @ (grids[i] / 1000.0) @ E.T  # [ky][kx]
        k2 = kv[None, :] ** 2 + kv[:, None] ** 2
        phi = np.where(k2 > 0, -2 * np.pi * G * rho_hat / np.where(k2 > 0, k2, 1), 0)
        Einv = np.conj(E) / g
        fx = (Einv @ (1j * kv[None, :] * phi) @ Einv.T).real
        fy = (Einv @ (1j * kv[:, None] * phi) @ Einv.T).real
        np.testing.assert_allclose(F[i, ..., 0], fx, atol=1e-9 * np.abs(fx).max())
        np.testing.assert_allclose(F[i, ..., 1], fy, atol=1e-9 * np.abs(fy).max())


def test_cell_known_answers_and_rotation():
    g = 16
    h = W / g
    rot = np.zeros(12, np.float32)
    rot[0] = rot[5] = rot[10] = 1.0
    off = np.zeros(4, np.float32)
    pos = np.array([[0.1, 0.1, 0, 1], [h * 3 + 0.5, h * 7 + 0.5, 0, 1], [W - 0.01, 0.5 * h, 0, 1]], np.float32)
    idx = O2.cell_indices_2d(pos, rot, off, g, W)
    assert list(idx) == [0, 7 * g + 3, 0 * g + 15]
    # 90 degrees: rot = [[0, -1], [1, 0]] in the packed block rot[4j+m] = M[j][m]
    r90 = np.zeros(12, np.float32)
    r90[1] = -1.0
    r90[4] = 1.0
    r90[10] = 1.0
    p = np.array([[W / 2 + 3 * h + 0.5, W / 2 + 0.5, 0, 1]], np.float32)
    # local r_m = sum_j rot[4j+m] c_j: r0 = c1, r1 = -c0
    i = O2.cell_indices_2d(p, r90, off, g, W)[0]
    assert i % g == int((0.5 + W / 2) / h) and i // g == int((W / 2 - 3 * h - 0.5) / h)


def test_identity_grid_is_textbook_particle_mesh():
    g = 16
    h = W / g
    rng = np.random.default_rng(0)
    pos = np.zeros((3000, 4), np.float32)
    pos[:, :2] = rng.uniform(0, W, (3000, 2))
    pos[:, 3] = 1.0
    rot = np.zeros((1, 12), np.float32)
    rot[0, 0] = rot[0, 5] = rot[0, 10] = 1.0
    off = np.zeros((1, 4), np.float32)
    f = O2.forces_2d(pos, rot, off, g, W, G)
    ix = np.floor(pos[:, 0].astype(np.float64) / h).astype(int) % g
    iy = np.floor(pos[:, 1].astype(np.float64) / h).astype(int) % g
    rho = np.zeros((g, g))
    np.add.at(rho, (iy, ix), 1.0)
    kv = O2.wave_numbers(g, W)
    k2 = kv[None, :] ** 2 + kv[:, None] ** 2
    rh = np.fft.fft2(rho)
    phi = np.where(k2 > 0, -2 * np.pi * G * rh / np.where(k2 > 0, k2, 1), 0)
    fx = np.fft.ifft2(1j * kv[None, :] * phi).real
    fy = np.fft.ifft2(1j * kv[:, None] * phi).real
    np.testing.assert_allclose(f[:, 0], fx[iy, ix], rtol=1e-9, atol=1e-9 * np.abs(fx).max())
    np.testing.assert_allclose(f[:, 1], fy[iy, ix], rtol=1e-9, atol=1e-9 * np.abs(fy).max())


def test_grid_force_conserves_momentum_and_rotated_field():
    g = 32
    rng = np.random.default_rng(1)
    grids = rng.integers(0, 5000, (1, g, g)).astype(np.int64)
    F = O2.solve_2d(grids, g, W, G)[0]
    rho = grids[0] / 1000.0
    p = np.sum(rho[..., None] * F, axis=(0, 1))
    assert np.max(np.abs(p)) < 1e-9 * np.sum(np.abs(rho[..., None] * F))


def test_transforms_2d_and_package_sampler_agree():
    from mog_solver_amd.mog2d import transforms_2d

    rot, off = O2.transforms_2d(50, 16, W, seed=7)
    r2, o2 = transforms_2d(50, 16, W, seed=7)
    assert np.array_equal(rot, r2) and np.array_equal(off, o2)
    M = np.stack([rot[:, [0, 1]], rot[:, [4, 5]]], axis=1).astype(np.float64)  # M[j][m]
    np.testing.assert_allclose(M @ np.transpose(M, (0, 2, 1)), np.broadcast_to(np.eye(2), M.shape), atol=1e-6)
    np.testing.assert_allclose(np.linalg.det(M), 1.0, atol=1e-6)
    assert np.all(off[:, :2] >= 0) and np.all(off[:, :2] < W / 16) and np.all(off[:, 2:] == 0)


def test_blob_sampler_and_analytic_force():
    """The harness's analytic force: Gauss's law part against a direct 2D
    pair sum (isolated, G h^2 sum m d / |d|^2), and the full periodic form
    (with the neutralising background) against the oracle's periodic solve
    on one fine identity grid."""
    from mog_solver_amd.mse2d import analytic_blob_force, deposited_mass, sample_blob

    pos = sample_blob(128, W, sigma=5.0, radius=20.0, seed=1)
    assert np.all(np.hypot(pos[:, 0] - W / 2, pos[:, 1] - W / 2) < 20.0)
    h = W / 16
    an = analytic_blob_force(pos, W, G, h)
    m = deposited_mass(pos)
    xy = pos[:, :2].astype(np.float64)
    direct = np.zeros_like(xy)
    for a in range(0, xy.shape[0], 1024):
        d = xy[a:a + 1024, None, :] - xy[None, :, :]
        r2 = np.sum(d * d, 2)
        r2[np.arange(d.shape[0]), a + np.arange(d.shape[0])] = np.inf
        direct[a:a + 1024] = G * h * h * np.sum(m[None, :, None] * d / r2[..., None], 1)
    c = xy - W / 2
    r = np.hypot(c[:, 0], c[:, 1])
    bg = G * h * h * np.pi * m.sum() * c / (W * W)
    outer = r > 5.0  # beyond the lattice-discreteness scale near the centre
    err = np.linalg.norm((an + bg - direct)[outer]) / np.linalg.norm(direct[outer])
    assert err < 2e-2, err
    # periodic: one identity grid of 256^2 cells (h' = W/256) -> same law with h'
    g = 256
    rot = np.zeros((1, 12), np.float32)
    rot[0, 0] = rot[0, 5] = rot[0, 10] = 1.0
    f = O2.forces_2d(pos, rot, np.zeros((1, 4), np.float32), g, W, G)
    want = analytic_blob_force(pos, W, G, W / g)
    err = np.linalg.norm((f - want)[outer]) / np.linalg.norm(want[outer])
    assert err < 3e-2, err
