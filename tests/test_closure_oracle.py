"""Pins of the closure oracle (no GPU)."""
import numpy as np

from oracle import closure_oracle as C
from oracle import mog_oracle as O


def _batch(n=40, seed=0):
    rng = np.random.default_rng(seed)
    feat = np.zeros((n, 8))
    feat[:, :7] = rng.normal(size=(n, 7))
    feat[:, 7] = rng.uniform(0.5, 3.0, n)
    mean = rng.normal(size=(n, 3))
    target = mean + rng.normal(size=(n, 3))
    return feat, mean, target


def test_gradient_matches_finite_differences():
    params = C.init_params(1, out_scale=1.0).astype(np.float64)
    feat, mean, target = _batch()
    wnorm = 1.0 / np.sum(feat[:, 7] ** 2)
    loss, grad = C.loss_and_grad(params, feat, mean, target, wnorm)
    rng = np.random.default_rng(2)
    idx = rng.choice(C.N_PARAMS, 60, replace=False)
    eps = 1e-6
    for i in idx:
        p1, p2 = params.copy(), params.copy()
        p1[i] += eps
        p2[i] -= eps
        fd = (C.loss_and_grad(p1, feat, mean, target, wnorm)[0] - C.loss_and_grad(p2, feat, mean, target, wnorm)[0]) / (2 * eps)
        assert abs(fd - grad[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, grad[i])


def test_zero_network_predicts_the_mean_and_loss_is_relative_error():
    params = np.zeros(C.N_PARAMS)
    feat, mean, target = _batch()
    pred = C.predict(params, feat, mean)
    assert np.allclose(pred, mean)
    wnorm = 1.0 / np.sum(feat[:, 7] ** 2)
    loss, _ = C.loss_and_grad(params, feat, mean, target, wnorm)
    # loss = sum |target - mean|^2 / sum S^2
    assert np.isclose(loss, np.sum((target - mean) ** 2) * wnorm)


def test_features_mean_is_the_mog_force():
    from mog_solver_amd import workloads

    pos, _ = workloads.uniform_cube(3000, 100.0, seed=4)
    rot, off = O.se3_transforms(4, 8, 100.0, 0)
    rho0 = 3000 / 8**3
    feat, m = C.features(pos, rot, off, 4, 8, 100.0, 100.0, rho0)
    avg, _ = O.forces(pos, rot, off, 4, 8, 100.0, 100.0)
    assert np.allclose(m, avg, rtol=1e-12, atol=1e-9 * np.abs(avg).max())
    assert np.allclose(feat[:, :3] * feat[:, 7:8], m)
    assert np.all(feat[:, 3:6] >= 0) and np.all(feat[:, 7] > 0)
    assert C.N_PARAMS == 9027
