"""mog_step_noalloc: the force step with its scratch (the gather's visiting
order) in a device-resident array instead of buffers it allocates.

Bar: bit-exact.  Each call must leave the same positions and velocities as
mog_step on the same input (a second context with the same configuration
and transforms); two calls in a row, the second with a smaller input, then
one below the size that uses a visiting order; a size above the maximum is
rejected before any launch.
"""
import ctypes

import numpy as np
import pytest

from mog_solver_amd import _abi
from mog_solver_amd import workloads

from test_gpu_parity import dev, make_solver

pytestmark = pytest.mark.gpu

MAX_PARTICLES = 1 << 20


def _noalloc(lib):
    fn = lib.mog_step_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]
    return fn


def _stream(torch):
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def test_noalloc_step_matches_step(torch_cuda):
    torch = torch_cuda
    k, g, W, G = 24, 16, 100.0, 100.0
    ref, _, _ = make_solver(k, g, W, G)
    na, _, _ = make_solver(k, g, W, G)
    step_na = _noalloc(na.lib)
    for n, seed in ((300_000, 1), (100_000, 2), (5_000, 3)):
        pos, vel = workloads.two_disks(n, W, G, g, seed=seed)
        pa, va = dev(torch, pos), dev(torch, vel)
        pb, vb = dev(torch, pos), dev(torch, vel)
        for _ in range(2):  # the second call reuses the order kept in the scratch
            ref.step(pa, va)
            rc = step_na(na._ctx, ctypes.c_void_p(pb.data_ptr()), ctypes.c_void_p(vb.data_ptr()), n, _stream(torch))
            _abi.check(na.lib, na._ctx, rc)
            torch.cuda.synchronize()
            assert torch.equal(pa, pb) and torch.equal(va, vb), n


def test_noalloc_rejects_too_many_particles(torch_cuda):
    torch = torch_cuda
    s, _, _ = make_solver(4, 16)
    step_na = _noalloc(s.lib)
    n = MAX_PARTICLES + 1
    p = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    v = torch.zeros_like(p)
    rc = step_na(s._ctx, ctypes.c_void_p(p.data_ptr()), ctypes.c_void_p(v.data_ptr()), n, _stream(torch))
    assert rc == _abi.MOG_EINVAL
    assert b"maximum" in s.lib.mog_last_error(s._ctx)
    torch.cuda.synchronize()
    assert int(p.abs().sum()) == 0 and int(v.abs().sum()) == 0
