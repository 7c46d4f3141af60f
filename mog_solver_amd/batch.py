"""Batched ensemble: many independent MOG simulations stepped together.

The reference runs one simulation per process (``Simulation``, main.py:104);
an ensemble of uneven simulations (different n, n_grids, grid_size) is this
build's extension.  ``MogBatch`` wraps ``mog_batch_*`` of ``libmog.so``:
every phase (deposit, the three spectral passes, gather+update) is one
persistent launch over all simulations, and ``run`` replays the whole step
as one hipGraph.  Particles live in packed ``(sum n, 4)`` float32 device
tensors; ``views`` splits them per simulation.  A batch step leaves the same
state, bit for bit, as stepping each simulation alone with the FP32 vector
spectral solve.
"""
from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from . import _abi
from .config import SolverConfig
from .solver import _check_particles, _ptr, _stream_ptr, mog_config
from .transforms import compute_transforms


class MogBatch:
    def __init__(self, cfgs: Sequence[SolverConfig], counts: Sequence[int], transforms=None, seeds=None):
        """``transforms``: per simulation ``(rot (k,12), off (k,4))``; else
        drawn by the reference sampler with ``seeds[s]`` (default s)."""
        if len(cfgs) != len(counts) or not cfgs:
            raise ValueError("need one particle count per config")
        self.lib = _abi.load()
        self.cfgs = list(cfgs)
        self.counts = np.asarray(counts, dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.counts)]).astype(np.int64)
        arr = (_abi.MogConfig * len(cfgs))(*[mog_config(c) for c in cfgs])
        n = np.ascontiguousarray(self.counts)
        h = ctypes.c_void_p()
        _abi.check(self.lib, None, self.lib.mog_batch_create(arr, n.ctypes.data, len(cfgs), ctypes.byref(h)))
        self._b = h
        self.transforms = []
        for s, c in enumerate(self.cfgs):
            if transforms is not None:
                rot, off = transforms[s]
            else:
                off, rot = compute_transforms(c.n_grids, c.grid_size, c.world_size,
                                              seed=int(seeds[s]) if seeds is not None else s)
            self.set_transforms(s, rot, off)

    @property
    def n_total(self) -> int:
        return int(self.offsets[-1])

    def close(self):
        if getattr(self, "_b", None):
            self.lib.mog_batch_destroy(self._b)
            self._b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        _abi.check(self.lib, None, rc)

    def set_transforms(self, sim, rot, off):
        rot = np.ascontiguousarray(rot, dtype=np.float32)
        off = np.ascontiguousarray(off, dtype=np.float32)
        k = self.cfgs[sim].n_grids
        if rot.shape != (k, 12) or off.shape != (k, 4):
            raise ValueError(f"sim {sim}: expected rot (k,12) and off (k,4) with k={k}")
        self._check(self.lib.mog_batch_set_transforms(self._b, sim, rot.ctypes.data, off.ctypes.data))
        if sim < len(self.transforms):
            self.transforms[sim] = (rot, off)
        else:
            self.transforms.append((rot, off))

    def _state(self, pos, vel):
        _check_particles(pos, "pos")
        _check_particles(vel, "vel")
        if pos.shape[0] != self.n_total or vel.shape[0] != self.n_total:
            raise ValueError(f"pos/vel must hold the {self.n_total} particles of the batch")

    def step(self, pos, vel, stream=None):
        self._state(pos, vel)
        self._check(self.lib.mog_batch_step(self._b, _ptr(pos), _ptr(vel), _stream_ptr(stream)))

    def run(self, pos, vel, nsteps, use_graph=True, stream=None):
        self._state(pos, vel)
        self._check(self.lib.mog_batch_run(self._b, _ptr(pos), _ptr(vel), int(nsteps), int(bool(use_graph)),
                                           _stream_ptr(stream)))

    def set_tuning(self, deposit_chunk=None, deposit_threads=None):
        t = self.get_tuning()
        chunk = t["deposit_chunk"] if deposit_chunk is None else int(deposit_chunk)
        threads = t["deposit_threads"] if deposit_threads is None else int(deposit_threads)
        self._check(self.lib.mog_batch_set_tuning(self._b, chunk, threads))

    def get_tuning(self) -> dict:
        chunk, threads, tasks = ctypes.c_int64(), ctypes.c_int32(), ctypes.c_int32()
        self._check(self.lib.mog_batch_get_tuning(self._b, ctypes.byref(chunk), ctypes.byref(threads),
                                                  ctypes.byref(tasks)))
        return {"deposit_chunk": chunk.value, "deposit_threads": threads.value, "n_tasks": tasks.value}

    def views(self, t):
        """Per-simulation row views of a packed (sum n, ...) tensor."""
        return [t[int(a):int(b)] for a, b in zip(self.offsets[:-1], self.offsets[1:])]

    def pack(self, states):
        """Packed device tensors from per-simulation host ``(pos, vel)`` pairs."""
        import torch

        pos = np.concatenate([np.asarray(p, dtype=np.float32) for p, _ in states])
        vel = np.concatenate([np.asarray(v, dtype=np.float32) for _, v in states])
        if pos.shape[0] != self.n_total:
            raise ValueError("particle counts differ from the batch's")
        return torch.from_numpy(pos).cuda(), torch.from_numpy(vel).cuda()


def autotune(batch: MogBatch, pos, vel, steps=5, verbose=False):
    """Pick the deposit task shape by timing graph-replayed steps on a copy
    of the state; leaves the fastest tuning set and returns it."""
    import time

    import torch

    p, v = pos.clone(), vel.clone()
    best = None
    for chunk in (4096, 8192, 16384, 32768, 65536, 131072, 262144):
        for threads in (512, 1024):
            batch.set_tuning(deposit_chunk=chunk, deposit_threads=threads)
            batch.run(p, v, 1)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            batch.run(p, v, steps)
            torch.cuda.synchronize()
            dt = (time.perf_counter() - t0) / steps
            if verbose:
                print(f"[batch autotune] chunk {chunk} threads {threads}: {dt * 1e3:.3f} ms")
            if best is None or dt < best[0]:
                best = (dt, chunk, threads)
    batch.set_tuning(deposit_chunk=best[1], deposit_threads=best[2])
    return batch.get_tuning()
