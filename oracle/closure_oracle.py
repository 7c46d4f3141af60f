"""CPU oracle of the learned closure -- TEST INFRASTRUCTURE ONLY.

FP64 NumPy versions of what the closure kernels compute (the closure has no
counterpart in the reference; it is specified by this build):

* ``features``: per particle, over the grids of a small ensemble, the mean m
  and per-axis standard deviation s of the world-frame NGP grid forces
  M_i F_i, the mean c of the trilinear (cloud-in-cell) interpolation of the
  deconvolved world-frame force fields (force spectrum divided by
  sinc(k_x h/2) sinc(k_y h/2) sinc(k_z h/2)) at the continuous cell
  coordinate u = l/h - 1/2, and the mean cell mass rho; x = (m/S, s/S,
  c/S, log(1 + rho/rho0)) with S = |c| + |s|; feat = (x, S, 0).
* ``mlp_forward`` / ``predict``: 10 -> 64 -> 64 -> 64 -> 3 SiLU network,
  F = c + S y.
* ``loss_and_grad``: loss = wnorm * sum_p S_p^2 |y_p - (F_p - c_p)/S_p|^2 and
  its gradient by backpropagation.

Pinning (tests/test_closure_oracle.py): ``loss_and_grad`` against central
finite differences; the NGP part of ``features`` reduces to the plain MOG
force of mog_oracle; the cloud-in-cell part reproduces the grid values at
cell centres and, for a single smooth Fourier mode, the analytic force at
arbitrary positions to interpolation accuracy (the deconvolution divides
out exactly the NGP window of the assignment).
"""
from __future__ import annotations

import numpy as np

from . import mog_oracle as O

K_IN, K_H, K_OUT = 10, 64, 3
S_COL = 10
SHAPES = [("W1", (K_IN, K_H)), ("b1", (K_H,)), ("W2", (K_H, K_H)), ("b2", (K_H,)),
          ("W3", (K_H, K_H)), ("b3", (K_H,)), ("W4", (K_H, K_OUT)), ("b4", (K_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)


def unpack(params):
    params = np.asarray(params, dtype=np.float64)
    out, at = {}, 0
    for name, shape in SHAPES:
        size = int(np.prod(shape))
        out[name] = params[at:at + size].reshape(shape)
        at += size
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SHAPES])


def init_params(seed=0, out_scale=0.1):
    """He-style initial weights (float32 values), zero biases, small last layer."""
    rng = np.random.default_rng(seed)
    d = {}
    for name, shape in SHAPES:
        if name.startswith("W"):
            d[name] = rng.normal(0.0, np.sqrt(2.0 / shape[0]), shape)
        else:
            d[name] = np.zeros(shape)
    d["W4"] *= out_scale
    return pack(d).astype(np.float32)


def sigmoid(z):
    return 1.0 / (1.0 + np.exp(-z))


def silu(z):
    return z * sigmoid(z)


def dsilu(z):
    s = sigmoid(z)
    return s * (1.0 + z * (1.0 - s))


def window(g, W):
    """NGP assignment window per axis, sinc(k h/2) at the Nyquist-positive
    wave numbers."""
    k = O.wave_numbers(g, W)
    x = 0.5 * k * (W / g)
    return np.where(x == 0, 1.0, np.sin(x) / np.where(x == 0, 1.0, x))


def deconvolved_force(grid, g, W, G, compat="reference"):
    """Grid-frame force field (g,g,g,3) [z][y][x] of one milli-mass grid,
    solved as mog_oracle.solve does but with the Green's function divided by
    the NGP window sinc(k_x h/2) sinc(k_y h/2) sinc(k_z h/2)."""
    S = window(g, W)
    win = S[None, None, :] * S[None, :, None] * S[:, None, None]
    rho_hat = np.fft.fftn(np.asarray(grid, dtype=np.float64) / 1000.0)
    phi = rho_hat * O.greens_factor(g, W, G) / win
    kv = O.wave_numbers(g, W)
    out = np.zeros((g, g, g, 3))
    for d, kd in enumerate((kv[None, None, :], kv[None, :, None], kv[:, None, None])):
        out[..., d] = np.fft.ifftn(1j * kd * phi).real
    if compat == "reference" and O.aliases_x_with_z(g):
        out[..., 0] = out[..., 2]
    return out


def cic(field, u, g):
    """Trilinear interpolation of a periodic cell-centred field (g,g,g,3)
    ([z][y][x]) at continuous cell coordinates u (n,3) (x, y, z)."""
    i0 = np.floor(u).astype(np.int64)
    f = u - i0
    out = np.zeros((u.shape[0], 3))
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = (f[:, 0] if dx else 1 - f[:, 0]) * (f[:, 1] if dy else 1 - f[:, 1]) * \
                    (f[:, 2] if dz else 1 - f[:, 2])
                out += w[:, None] * field[(i0[:, 2] + dz) % g, (i0[:, 1] + dy) % g, (i0[:, 0] + dx) % g]
    return out


def cell_coords(c, rot_i, off_i, W, g):
    """FP64 continuous cell coordinate u = l/h - 1/2 (n,3) of centred
    positions c in one grid: l = rot * c + W/2 - off (the frame of
    mog_oracle.cell_indices_one), so that cell i spans u in [i - 1/2, i + 1/2)
    and its centre is u = i."""
    R = np.asarray(rot_i, dtype=np.float64)
    c = np.asarray(c, dtype=np.float64)
    l = np.stack([R[m] * c[:, 0] + R[4 + m] * c[:, 1] + R[8 + m] * c[:, 2] for m in range(3)], 1)
    return (l + W / 2 - np.asarray(off_i, dtype=np.float64)[:3]) / (W / g) - 0.5


def features(pos, rot, off, n_grids, grid_size, world_size, G, rho0, compat="reference"):
    """(feat (n,12), mean (n,3)) in float64: feat = (m/S, s/S, c/S,
    log1p(rho/rho0), S, 0), mean = c."""
    P = O.Params.make(n_grids, grid_size, world_size)
    c = O.centered32(pos, P)
    Ms = O.rotation_matrices(rot).astype(np.float64)
    milli = O.milli_mass(pos)
    n = c.shape[0]
    s1 = np.zeros((n, 3))
    s2 = np.zeros((n, 3))
    cc = np.zeros((n, 3))
    rho = np.zeros(n)
    g = P.g
    c64 = c.astype(np.float64)
    for i in range(P.k):
        idx = O.cell_indices_one(c, rot[i], off[i], P)
        grid = np.zeros(g**3, dtype=np.int64)
        np.add.at(grid, idx, milli)
        F = O.solve(grid.reshape(1, g, g, g), g, world_size, G, compat=compat)[0].reshape(-1, 4)[idx, :3]
        Fw = F @ Ms[i].T
        s1 += Fw
        s2 += Fw * Fw
        rho += grid[idx] / 1000.0
        u = cell_coords(c64, rot[i], off[i], world_size, g)
        D = deconvolved_force(grid.reshape(g, g, g), g, world_size, G, compat)
        cc += cic(D, u, g) @ Ms[i].T
    m = s1 / P.k
    var = np.maximum(s2 / P.k - m * m, 0.0)
    sd = np.sqrt(var)
    cm = cc / P.k
    S = np.linalg.norm(cm, axis=1) + np.sqrt(var.sum(1)) + 1e-30
    feat = np.zeros((n, 12))
    feat[:, 0:3] = m / S[:, None]
    feat[:, 3:6] = sd / S[:, None]
    feat[:, 6:9] = cm / S[:, None]
    feat[:, 9] = np.log1p(rho / P.k / rho0)
    feat[:, S_COL] = S
    return feat, cm


def mlp_forward(params, x, keep=False):
    p = unpack(params)
    z1 = x @ p["W1"] + p["b1"]
    z2 = silu(z1) @ p["W2"] + p["b2"]
    z3 = silu(z2) @ p["W3"] + p["b3"]
    y = silu(z3) @ p["W4"] + p["b4"]
    return (y, (z1, z2, z3)) if keep else y


def predict(params, feat, mean):
    feat = np.asarray(feat, dtype=np.float64)
    y = mlp_forward(params, feat[:, :K_IN])
    return np.asarray(mean, dtype=np.float64)[:, :3] + feat[:, S_COL:S_COL + 1] * y


def loss_and_grad(params, feat, mean, target, wnorm):
    feat = np.asarray(feat, dtype=np.float64)
    x = feat[:, :K_IN]
    S = feat[:, S_COL]
    t = (np.asarray(target, dtype=np.float64)[:, :3] - np.asarray(mean, dtype=np.float64)[:, :3]) / S[:, None]
    w = S * S * wnorm
    p = unpack(params)
    y, (z1, z2, z3) = mlp_forward(params, x, keep=True)
    e = y - t
    loss = float(np.sum(w[:, None] * e * e))
    dy = 2.0 * w[:, None] * e
    g = {}
    a3 = silu(z3)
    g["W4"] = a3.T @ dy
    g["b4"] = dy.sum(0)
    d3 = (dy @ p["W4"].T) * dsilu(z3)
    a2 = silu(z2)
    g["W3"] = a2.T @ d3
    g["b3"] = d3.sum(0)
    d2 = (d3 @ p["W3"].T) * dsilu(z2)
    a1 = silu(z1)
    g["W2"] = a1.T @ d2
    g["b2"] = d2.sum(0)
    d1 = (d2 @ p["W2"].T) * dsilu(z1)
    g["W1"] = x.T @ d1
    g["b1"] = d1.sum(0)
    return loss, pack(g)
