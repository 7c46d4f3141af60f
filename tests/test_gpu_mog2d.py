"""GPU parity of the 2D path against the FP64 2D oracle.

Bars: average force within 1e-5 relative RMS of the oracle (the pipeline's
force tolerance); the step's update bit-exact given the GPU force; the MSE
harness shows the ensemble convergence the 2D path exists for.
"""
import numpy as np
import pytest

from oracle import mog2d_oracle as O2
from oracle import mog_oracle as O

pytestmark = pytest.mark.gpu

W, G = 100.0, 100.0


def positions(n, g, seed):
    rng = np.random.default_rng(seed)
    h = np.float32(W / g)
    xy = rng.uniform(0, W, (n, 2)).astype(np.float32)
    faces = (np.arange(g + 1, dtype=np.float32) * h)
    edge = np.stack(np.meshgrid(faces, faces[::2]), -1).reshape(-1, 2).astype(np.float32)
    extra = np.array([[0, 0], [W, W / 2], [-1e-7, 3.0], [1.5 * W, -W / 3]], np.float32)
    xy = np.concatenate([xy, edge, extra])
    pos = np.zeros((xy.shape[0], 4), np.float32)
    pos[:, :2] = xy
    pos[:, 2] = rng.uniform(0, W, xy.shape[0]).astype(np.float32)  # unused by the 2D path
    pos[:, 3] = rng.uniform(0.5, 2.0, xy.shape[0]).astype(np.float32)
    return pos


@pytest.mark.parametrize("k,g", [(6, 16), (3, 4), (2, 64), (7, 32), (1, 2), (4, 8)])
def test_deposit_bit_exact(torch_cuda, k, g):
    """Density grids equal the oracle's bit for bit, including particles on
    and next to cell faces and outside the box (the fast cell path with the
    exact float32 sequence as fallback)."""
    from mog_solver_amd.mog2d import Mog2D, transforms_2d

    torch = torch_cuda
    pos = positions(30000, g, seed=k * 7 + g)
    # values one float32 step either side of every face
    h = np.float32(W / g)
    faces = np.arange(g + 1, dtype=np.float32) * h
    near = np.concatenate([np.nextafter(faces, np.float32(-1)), np.nextafter(faces, np.float32(W + 1))])
    grid = np.stack(np.meshgrid(near, near[::3]), -1).reshape(-1, 2).astype(np.float32)
    extra = np.zeros((grid.shape[0], 4), np.float32)
    extra[:, :2] = grid
    extra[:, 3] = 1.0
    pos = np.concatenate([pos, extra])
    rot, off = transforms_2d(k, g, W, seed=g + k)
    m = Mog2D(k, g, W, G, rot=rot, off=off)
    got = m.deposit(torch.from_numpy(pos).cuda()).cpu().numpy()
    want = O2.deposit_2d(pos, rot, off, g, W)
    assert np.array_equal(got, want)
    # identity transforms put grid faces on the lattice: every face value is
    # decided by the exact sequence
    rot_i = np.zeros((1, 12), np.float32)
    rot_i[0, [0, 5, 10]] = 1.0
    m1 = Mog2D(1, g, W, G, rot=rot_i, off=np.zeros((1, 4), np.float32))
    got1 = m1.deposit(torch.from_numpy(pos).cuda()).cpu().numpy()
    assert np.array_equal(got1, O2.deposit_2d(pos, rot_i, np.zeros((1, 4), np.float32), g, W))


@pytest.mark.parametrize("k,g", [(5, 16), (3, 4), (2, 64), (7, 32), (1, 2)])
def test_forces_match_oracle(torch_cuda, k, g):
    from mog_solver_amd.mog2d import Mog2D, transforms_2d

    torch = torch_cuda
    pos = positions(20000, g, seed=k * 10 + g)
    rot, off = transforms_2d(k, g, W, seed=g)
    m = Mog2D(k, g, W, G, rot=rot, off=off)
    got = m.forces(torch.from_numpy(pos).cuda()).cpu().numpy()
    want = O2.forces_2d(pos, rot, off, g, W, G)
    err = O.rel_rms(np.pad(got[:, :2], ((0, 0), (0, 1))), np.pad(want, ((0, 0), (0, 1))))
    assert err < 1e-5, err
    assert np.all(got[:, 2:] == 0)


def test_step_is_force_then_update(torch_cuda):
    from mog_solver_amd.mog2d import Mog2D

    torch = torch_cuda
    k, g = 6, 16
    pos = positions(30000, g, seed=5)
    pos[:, 2] = 0.0
    vel = np.zeros_like(pos)
    vel[:, :2] = np.random.default_rng(2).normal(0, 3, (pos.shape[0], 2)).astype(np.float32)
    m = Mog2D(k, g, W, G, seed=3)
    p, v = torch.from_numpy(pos).cuda(), torch.from_numpy(vel).cuda()
    f = m.forces(p).cpu().numpy()
    m.step(p, v, dt=1e-3, damping=0.99)
    want_p, want_v = O.update32(pos, vel, f, 1e-3, 0.99, W)
    assert np.array_equal(p.cpu().numpy(), want_p) and np.array_equal(v.cpu().numpy(), want_v)
    assert np.all(p.cpu().numpy()[:, 2] == 0)


def test_invalid_arguments(torch_cuda):
    from mog_solver_amd._abi import MogError
    from mog_solver_amd.mog2d import Mog2D

    with pytest.raises(MogError):
        Mog2D(4, 12, W, G)
    with pytest.raises(MogError):
        Mog2D(0, 16, W, G)
    with pytest.raises(MogError):
        Mog2D(4, 128, W, G)


def test_mse_harness_shows_ensemble_convergence(torch_cuda):
    """At fixed grid size the relative force MSE against the analytic field
    falls with the ensemble size."""
    from mog_solver_amd.mse2d import mse_table

    out = mse_table(n_side=512, sizes=(16, 64), ensembles=(1, 32), trials=2)
    t = {(r["G"], r["N"]): r["rel_mse"] for r in out["rows"]}
    print(t)
    assert t[(16, 32)] < 0.5 * t[(16, 1)]
    assert t[(64, 32)] < 0.5 * t[(64, 1)]
