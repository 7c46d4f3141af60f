"""CPU oracle of the 2D (DIM == 2) MOG force -- TEST INFRASTRUCTURE ONLY.

The reference keeps 2D branches in its shaders (``update.glsl:19-34``:
``mat2(col0.xy, col1.xy)`` rotations and ``sampleForce2D``;
``init_particles.glsl:151``; ``pack.glsl:8``) but always compiles with
``#define DIM 3`` (``main.py:237``) and has no 2D deposit, so its 2D path
does not run.  This build defines the 2D step from those branches and the
3D pipeline, restricted to two axes:

* cell: ``c = p.xy - W/2``; ``r_m = rot[m]*c0 + rot[4+m]*c1`` (float32, this
  order, no FMA; the 2x2 block of the packed 3D layout); ``l = (r + W/2) -
  off``; per axis ``int(mod(mod(l,W)+W,W)/h) % g``; flat ``y*g + x``;
* deposit: int64 milli-mass ``uint(m*1000)`` per cell;
* solve: ``phi = DFT2(rho/1000) * (-2 pi G / k^2)`` (DC zero), ``F_d =
  Re(IDFT2(i k_d phi))`` with Nyquist-positive wave numbers, in FP64;
* gather: ``world_m = sum_j rot[4m+j] F_j``, averaged over the grids.

Pinning (tests/test_oracle_2d.py): a single Fourier mode of the density has
the closed-form force; brute-force 2D DFT at small g; with one identity
grid the pipeline is textbook 2D particle-mesh; net momentum of the grid
force vanishes; the float32 cell index agrees with FP64 away from cell
faces and with known answers.
"""
from __future__ import annotations

import math

import numpy as np

from .mog_oracle import F32, axis_cell32, milli_mass

__all__ = ["transforms_2d", "cell_indices_2d", "deposit_2d", "solve_2d", "forces_2d", "analytic_mode_force"]


def transforms_2d(n_grids, grid_size, world_size, seed=0):
    """Packed ``(rot (k,12), off (k,4))`` of k random planar rotations
    (angle U[0, 2 pi)) and offsets U[0, h)^2, in the 3D packed layout with
    ``rot[4j+m] = M[j][m]`` for the 2x2 block and the z axis fixed."""
    rng = np.random.default_rng(seed)
    theta = rng.uniform(0.0, 2.0 * np.pi, n_grids)
    h = world_size / grid_size
    o = rng.uniform(0.0, h, (n_grids, 2))
    rot = np.zeros((n_grids, 12), dtype=np.float32)
    rot[:, 0] = np.cos(theta)
    rot[:, 1] = -np.sin(theta)
    rot[:, 4] = np.sin(theta)
    rot[:, 5] = np.cos(theta)
    rot[:, 10] = 1.0
    off = np.zeros((n_grids, 4), dtype=np.float32)
    off[:, :2] = o
    return rot, off


def cell_indices_2d(pos, rot_i, off_i, g, W):
    """Flat cell ``y*g + x`` of every particle in one 2D grid (float32)."""
    Wf = F32(W)
    h = F32(float(W) / g)
    half = F32(Wf * F32(0.5))
    p = np.asarray(pos, dtype=np.float32)
    c0 = (p[:, 0] - half).astype(np.float32)
    c1 = (p[:, 1] - half).astype(np.float32)
    R = np.asarray(rot_i, dtype=np.float32)
    o = np.asarray(off_i, dtype=np.float32)
    idx = np.zeros(p.shape[0], dtype=np.int64)
    lg = int(round(math.log2(g)))
    for m in range(2):
        r = ((R[m] * c0).astype(np.float32) + (R[4 + m] * c1).astype(np.float32)).astype(np.float32)
        l = ((r + half).astype(np.float32) - o[m]).astype(np.float32)
        idx |= axis_cell32(l, Wf, h, g) << (lg * m)
    return idx


def deposit_2d(pos, rot, off, g, W):
    """(k, g, g) int64 milli-mass grids."""
    milli = milli_mass(pos)
    out = np.zeros((rot.shape[0], g * g), dtype=np.int64)
    for i in range(rot.shape[0]):
        np.add.at(out[i], cell_indices_2d(pos, rot[i], off[i], g, W), milli)
    return out.reshape(-1, g, g)


def wave_numbers(g, W):
    s = np.arange(g)
    return np.where(s <= g // 2, s, s - g).astype(np.float64) * (2.0 * np.pi / W)


def solve_2d(grids, g, W, G):
    """(k, g, g, 2) FP64 force grids (Fx, Fy) of milli-mass grids."""
    rho = np.asarray(grids, dtype=np.float64) / 1000.0
    rho_hat = np.fft.fft2(rho, axes=(1, 2))
    kv = wave_numbers(g, W)
    kx = kv[None, None, :]
    ky = kv[None, :, None]
    k2 = kx * kx + ky * ky
    with np.errstate(divide="ignore", invalid="ignore"):
        gf = np.where(k2 > 0, -2.0 * np.pi * G / k2, 0.0)
    phi = rho_hat * gf
    out = np.zeros(rho.shape + (2,))
    out[..., 0] = np.fft.ifft2(1j * kx * phi, axes=(1, 2)).real
    out[..., 1] = np.fft.ifft2(1j * ky * phi, axes=(1, 2)).real
    return out


def forces_2d(pos, rot, off, g, W, G):
    """FP64 average world force (n, 2) over the grids."""
    grids = deposit_2d(pos, rot, off, g, W)
    F = solve_2d(grids, g, W, G).reshape(rot.shape[0], g * g, 2)
    acc = np.zeros((np.asarray(pos).shape[0], 2))
    for i in range(rot.shape[0]):
        idx = cell_indices_2d(pos, rot[i], off[i], g, W)
        f = F[i, idx]
        R = np.asarray(rot[i], dtype=np.float64)
        acc[:, 0] += R[0] * f[:, 0] + R[1] * f[:, 1]
        acc[:, 1] += R[4] * f[:, 0] + R[5] * f[:, 1]
    return acc / rot.shape[0]


def analytic_mode_force(xy, W, G, h, rho_bar, A):
    """Force of the periodic surface density rho_bar (1 + A cos(ax) cos(ay)),
    a = 2 pi / W, with the 2D MOG coupling (del^2 phi = 2 pi G h^2 rho):
    F = pi G h^2 rho_bar A / a * (sin ax cos ay, cos ax sin ay)."""
    a = 2.0 * np.pi / W
    x, y = xy[:, 0].astype(np.float64), xy[:, 1].astype(np.float64)
    c = np.pi * G * h * h * rho_bar * A / a
    return np.stack([c * np.sin(a * x) * np.cos(a * y), c * np.cos(a * x) * np.sin(a * y)], axis=1)
