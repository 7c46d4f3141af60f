"""Per-grid SE(3) transforms: random rotations and sub-voxel offsets.

Same sampling as ``RandomSE3Sampler`` (main.py:9-62) and
``Simulation.compute_transforms`` (main.py:157-172): uniform unit
quaternions by Shoemake's method from ``numpy.random.default_rng(seed)``
(FP64, cast to float32), offsets ``U[0, voxel)`` drawn after them from the
same generator, matrix entries evaluated in float32.  The reference draws
with ``seed=None`` (different grids every run); here the seed is a
parameter, 0 by default.

Packed layouts handed to the kernels (the reference's SSBO layouts):
``rot[i, 4*j + m] = M_i[j][m]`` for j, m < 3 with slots 3, 7, 11 zero, and
``off[i] = (ox, oy, oz, 0)``.  A grid sees the local position
``r_m = sum_j rot[4j+m] c_j`` (M^T c) and returns forces as M F.
"""
from __future__ import annotations

import numpy as np


class RandomSE3Sampler:
    """Uniform random rotations + offsets in [0, voxel) per axis."""

    def __init__(self, grid_size, world_size, n_grids, seed=0):
        self.grid_size = grid_size
        self.world_size = world_size
        self.voxel_size = world_size / grid_size
        self.n_grids = n_grids
        self.seed = seed

    def generate(self):
        """Return ``(rotations (k,3,3) float32 M, offsets (k,3) float32)``."""
        rng = np.random.default_rng(self.seed)
        k = self.n_grids
        u = rng.uniform(0, 1, (k, 3))
        a = np.sqrt(1 - u[:, 0])
        b = np.sqrt(u[:, 0])
        t1 = 2 * np.pi * u[:, 1]
        t2 = 2 * np.pi * u[:, 2]
        quats = np.stack([a * np.sin(t1), a * np.cos(t1), b * np.sin(t2), b * np.cos(t2)], axis=1)
        quats = quats.astype(np.float32)
        offsets = rng.uniform(0, self.voxel_size, (k, 3)).astype(np.float32)
        return self.quat_to_matrix(quats), offsets

    def get_transforms(self, verbose=False):
        if verbose:
            print(f"  Generating {self.n_grids} random SE(3) transforms...")
        return self.generate()

    @staticmethod
    def quat_to_matrix(q):
        """(k,4) float32 quaternions (w, x, y, z) -> (k,3,3) float32 matrices,
        every product and sum rounded to float32 left to right."""
        q = np.asarray(q, dtype=np.float32)
        w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
        one = np.float32(1)
        two = np.float32(2)
        m = np.empty((q.shape[0], 3, 3), dtype=np.float32)
        m[:, 0, 0] = one - two * y * y - two * z * z
        m[:, 0, 1] = two * x * y - two * z * w
        m[:, 0, 2] = two * x * z + two * y * w
        m[:, 1, 0] = two * x * y + two * z * w
        m[:, 1, 1] = one - two * x * x - two * z * z
        m[:, 1, 2] = two * y * z - two * x * w
        m[:, 2, 0] = two * x * z - two * y * w
        m[:, 2, 1] = two * y * z + two * x * w
        m[:, 2, 2] = one - two * x * x - two * y * y
        return m


def pack_transforms(rotations, offsets):
    """(k,3,3) matrices + (k,3) offsets -> packed (k,12) and (k,4) float32."""
    rotations = np.asarray(rotations, dtype=np.float32)
    offsets = np.asarray(offsets, dtype=np.float32)
    k = rotations.shape[0]
    rot = np.zeros((k, 12), dtype=np.float32)
    off = np.zeros((k, 4), dtype=np.float32)
    for j in range(3):
        rot[:, 4 * j : 4 * j + 3] = rotations[:, j, :]
    off[:, :3] = offsets
    return rot, off


def compute_transforms(n_grids, grid_size, world_size, seed=0):
    """Packed ``(offsets (k,4), rotations (k,12))``, the return order of
    ``Simulation.compute_transforms`` (main.py:157-172)."""
    sampler = RandomSE3Sampler(grid_size=grid_size, world_size=world_size, n_grids=n_grids, seed=seed)
    mats, offs = sampler.get_transforms()
    rot, off = pack_transforms(mats, offs)
    return off, rot
