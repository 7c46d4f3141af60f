"""The step's density grids (`mog_grids_ptr`): after every step, eager or
replayed from a graph, they hold the deposit of the positions that step
started from, and other uses of the context between steps leave the steps
unchanged.  The step deposits into a spare buffer that the previous step's
gather cleared, then swaps the two.
"""
import ctypes

import numpy as np
import pytest

from test_gpu_parity import dev, make_solver

pytestmark = pytest.mark.gpu


def _grids(torch, s):
    """Copy of the context's current density grids (k, g, g, g) int64."""
    torch.cuda.synchronize()
    k, g = s.cfg.n_grids, s.cfg.grid_size
    out = np.empty((k, g, g, g), np.int64)
    hip = ctypes.CDLL("libamdhip64.so")
    hip.hipMemcpy.restype = ctypes.c_int
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    src = s.lib.mog_grids_ptr(s._ctx)
    assert src
    assert hip.hipMemcpy(out.ctypes.data, src, out.nbytes, 2) == 0  # device to host
    return out


def _state(n, seed):
    rng = np.random.default_rng(seed)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0, 100, (n, 3))
    pos[:, 3] = rng.uniform(0.5, 2.0, n)
    vel = np.zeros((n, 4), np.float32)
    vel[:, :3] = rng.normal(0, 30, (n, 3))
    return pos, vel


@pytest.mark.parametrize("use_graph", [False, True])
def test_grids_after_each_step_are_its_deposit(torch_cuda, use_graph):
    torch = torch_cuda
    s, _, _ = make_solver(5, 16, dt=1e-3)
    pos, vel = _state(30000, 1)
    p, v = dev(torch, pos), dev(torch, vel)
    for _ in range(5):
        want = s.deposit(p.clone()).cpu().numpy()
        s.run(p, v, 1, use_graph=use_graph)
        assert np.array_equal(_grids(torch, s), want)


def test_other_calls_between_steps_leave_the_steps_unchanged(torch_cuda):
    torch = torch_cuda
    pos, vel = _state(20000, 2)
    a, _, _ = make_solver(4, 16, dt=1e-3)
    b, _, _ = make_solver(4, 16, dt=1e-3)
    pa, va = dev(torch, pos), dev(torch, vel)
    pb, vb = dev(torch, pos), dev(torch, vel)
    for i in range(4):
        a.run(pa, va, 1, use_graph=bool(i % 2))
        b.run(pb, vb, 1, use_graph=bool(i % 2))
        b.forces(pb)
        b.deposit(pb)
        b.step_profiled(pb.clone(), vb.clone())
    assert torch.equal(pa, pb) and torch.equal(va, vb)
