"""2D MOG path: planar grids (the DIM == 2 branches of update.glsl:19-34,
pack.glsl:8, init_particles.glsl:151, which the reference keeps but never
compiles: main.py:237 always defines DIM 3).

``Mog2D`` wraps ``mog2d_*`` of ``libmog.so``: k planar grids of g x g cells
over [0, W)^2 with random planar rotations and offsets within a cell; NGP
int64 milli-mass deposit, per-grid 2D spectral Poisson solve
(-2 pi G / k^2, i k), gather of the average world force, and the
reference's kick/damp/drift/wrap update.  Particles are (n,4) float32
device tensors (x, y, unused, m), as in the 3D path.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _abi
from .solver import _check_particles, _ptr, _stream_ptr


def transforms_2d(n_grids, grid_size, world_size, seed=0):
    """Packed ``(rot (k,12), off (k,4))``: planar rotations by angles
    U[0, 2 pi) in the 2x2 block (``rot[4j+m] = M[j][m]``, z fixed) and
    offsets U[0, h)^2, drawn from ``default_rng(seed)`` (angles first)."""
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


class Mog2D:
    """One 2D MOG context (``mog2d_create`` ... ``mog2d_destroy``)."""

    def __init__(self, n_grids, grid_size, world_size=100.0, G=100.0, rot=None, off=None, seed=0):
        self.lib = _abi.load()
        self.n_grids, self.grid_size = int(n_grids), int(grid_size)
        self.world_size, self.G = float(world_size), float(G)
        if rot is None or off is None:
            rot, off = transforms_2d(self.n_grids, self.grid_size, self.world_size, seed=seed)
        self.rot = np.ascontiguousarray(rot, dtype=np.float32)
        self.off = np.ascontiguousarray(off, dtype=np.float32)
        if self.rot.shape != (self.n_grids, 12) or self.off.shape != (self.n_grids, 4):
            raise ValueError("expected rot (k,12) and off (k,4)")
        h = ctypes.c_void_p()
        _abi.check(self.lib, None, self.lib.mog2d_create(self.n_grids, self.grid_size, self.world_size, self.G,
                                                         self.rot.ctypes.data, self.off.ctypes.data,
                                                         ctypes.byref(h)))
        self._ctx = h

    def close(self):
        if getattr(self, "_ctx", None):
            self.lib.mog2d_destroy(self._ctx)
            self._ctx = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def voxel_size(self):
        return self.world_size / self.grid_size

    def forces(self, pos, out=None, stream=None):
        """Average world force (n,4) = (Fx, Fy, 0, 0) at ``pos``."""
        import torch

        _check_particles(pos, "pos")
        if out is None:
            out = torch.empty_like(pos)
        _abi.check(self.lib, None, self.lib.mog2d_forces(self._ctx, _ptr(pos), pos.shape[0], _ptr(out),
                                                         _stream_ptr(stream)))
        return out

    def deposit(self, pos, out=None, stream=None):
        """Density grids (k, g, g) int64 milli-mass ([grid][y][x]) of ``pos``."""
        import torch

        _check_particles(pos, "pos")
        if out is None:
            out = torch.empty((self.n_grids, self.grid_size, self.grid_size), dtype=torch.int64,
                              device=pos.device)
        _abi.check(self.lib, None, self.lib.mog2d_deposit(self._ctx, _ptr(pos), pos.shape[0], _ptr(out),
                                                          _stream_ptr(stream)))
        return out

    def step(self, pos, vel, dt, damping=1.0, stream=None):
        _check_particles(pos, "pos")
        _check_particles(vel, "vel")
        _abi.check(self.lib, None, self.lib.mog2d_step(self._ctx, _ptr(pos), _ptr(vel), pos.shape[0], float(dt),
                                                       float(damping), _stream_ptr(stream)))
