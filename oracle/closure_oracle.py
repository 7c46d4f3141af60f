"""CPU oracle of the learned closure -- TEST INFRASTRUCTURE ONLY.

FP64 NumPy versions of what the closure kernels compute (the closure has no
counterpart in the reference; it is specified by this build):

* ``features``: per particle, over the grids of a small ensemble, the mean m
  and per-axis standard deviation s of the world-frame grid forces M_i F_i,
  the mean cell mass rho; x = (m/S, s/S, log(1 + rho/rho0)) with
  S = |m| + |s|.
* ``mlp_forward`` / ``predict``: 7 -> 64 -> 64 -> 64 -> 3 SiLU network,
  F = m + S y.
* ``loss_and_grad``: loss = wnorm * sum_p S_p^2 |y_p - (F_p - m_p)/S_p|^2 and
  its gradient by backpropagation.

Pinning: ``loss_and_grad`` is checked against central finite differences in
tests/test_closure_oracle.py; ``features`` reduce to the plain MOG force
(mean) of mog_oracle for the same grids.
"""
from __future__ import annotations

import numpy as np

from . import mog_oracle as O

K_IN, K_H, K_OUT = 7, 64, 3
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


def features(pos, rot, off, n_grids, grid_size, world_size, G, rho0, compat="reference"):
    """(feat (n,8), mean (n,3)) in float64: feat = (m/S, s/S, log1p(rho/rho0), S)."""
    P = O.Params.make(n_grids, grid_size, world_size)
    c = O.centered32(pos, P)
    Ms = O.rotation_matrices(rot).astype(np.float64)
    milli = O.milli_mass(pos)
    n = c.shape[0]
    s1 = np.zeros((n, 3))
    s2 = np.zeros((n, 3))
    rho = np.zeros(n)
    g = P.g
    for i in range(P.k):
        idx = O.cell_indices_one(c, rot[i], off[i], P)
        grid = np.zeros(g**3, dtype=np.int64)
        np.add.at(grid, idx, milli)
        F = O.solve(grid.reshape(1, g, g, g), g, world_size, G, compat=compat)[0].reshape(-1, 4)[idx, :3]
        Fw = F @ Ms[i].T
        s1 += Fw
        s2 += Fw * Fw
        rho += grid[idx] / 1000.0
    m = s1 / P.k
    var = np.maximum(s2 / P.k - m * m, 0.0)
    sd = np.sqrt(var)
    S = np.linalg.norm(m, axis=1) + np.sqrt(var.sum(1)) + 1e-30
    feat = np.zeros((n, 8))
    feat[:, 0:3] = m / S[:, None]
    feat[:, 3:6] = sd / S[:, None]
    feat[:, 6] = np.log1p(rho / P.k / rho0)
    feat[:, 7] = S
    return feat, m


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
    return np.asarray(mean, dtype=np.float64)[:, :3] + feat[:, 7:8] * y


def loss_and_grad(params, feat, mean, target, wnorm):
    feat = np.asarray(feat, dtype=np.float64)
    x = feat[:, :K_IN]
    S = feat[:, 7]
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
