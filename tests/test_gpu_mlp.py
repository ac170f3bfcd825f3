"""The selector network's GPU kernels (se2p_mlp_*) against the CPU restatement
(oracle/mlp.py, pinned by finite differences): forward pass, loss and
gradient, Adam step, determinism of the gradient sums, and a training run."""
import numpy as np
import pytest

from fasttools2d_amd import selector as S
from oracle import mlp as M

pytestmark = pytest.mark.gpu


def _case(seed, n):
    rng = np.random.default_rng(seed)
    w = (rng.standard_normal(M.N_PARAMS) * 0.3).astype(np.float32)
    x = rng.standard_normal((n, 8)).astype(np.float32)
    t = rng.standard_normal((n, 4)).astype(np.float32)
    return w, x, t


@pytest.mark.parametrize("n", [1, 63, 64, 65, 1000])
def test_mlp_forward_matches_oracle(gpu_ctx, n):
    w, x, _ = _case(n, n)
    got = S.forward(gpu_ctx, w, x)
    ref = M.forward(w.astype(np.float64), x.astype(np.float64))
    assert np.max(np.abs(got - ref)) <= 2e-5 * (1 + np.max(np.abs(ref)))


@pytest.mark.parametrize("n", [5, 64, 130, 2000])
def test_mlp_gradient_matches_oracle(gpu_ctx, n):
    w, x, t = _case(10 + n, n)
    loss, g = S.loss_grad(gpu_ctx, w, x, t)
    rl, rg = M.loss_grad(w.astype(np.float64), x.astype(np.float64), t.astype(np.float64))
    assert abs(loss - rl) <= 1e-5 * rl
    assert np.linalg.norm(g - rg) <= 1e-4 * np.linalg.norm(rg)
    assert np.max(np.abs(g - rg)) <= 1e-4 * np.max(np.abs(rg))
    # fixed summation order: the same bits on a second call
    loss2, g2 = S.loss_grad(gpu_ctx, w, x, t)
    assert loss2 == loss and np.array_equal(g, g2)


def test_mlp_adam_step_matches_oracle(gpu_ctx):
    import ctypes
    from fasttools2d_amd._lib import check, lib
    rng = np.random.default_rng(3)
    w = rng.standard_normal(M.N_PARAMS).astype(np.float32)
    g = rng.standard_normal(M.N_PARAMS).astype(np.float32)
    m = (rng.standard_normal(M.N_PARAMS) * 0.1).astype(np.float32)
    v = (rng.random(M.N_PARAMS) * 0.1).astype(np.float32)
    ptrs = [S._dev(gpu_ctx, a) for a in (w, g, m, v)]
    try:
        check(lib().se2p_mlp_adam(gpu_ctx.handle, ptrs[0], ptrs[1], ptrs[2], ptrs[3], 7, 1e-3,
                                  0.9, 0.999, 1e-8))
        gpu_ctx.sync()
        got_w = S._host(gpu_ctx, ptrs[0], (M.N_PARAMS,))
        got_m = S._host(gpu_ctx, ptrs[2], (M.N_PARAMS,))
    finally:
        for p in ptrs:
            gpu_ctx.free(p)
    rw, rm, _ = M.adam_step(w.astype(np.float64), g.astype(np.float64), m.astype(np.float64),
                            v.astype(np.float64), 7, 1e-3)
    assert np.max(np.abs(got_m - rm)) < 1e-6
    assert np.max(np.abs(got_w - rw)) < 1e-5
    assert ctypes is not None


def test_mlp_training_fits_a_smooth_map(gpu_ctx):
    """Full-batch Adam with the GPU kernels fits 4 smooth functions of 8
    inputs: the loss falls by more than 50x and the oracle's forward pass of the
    trained weights agrees with the GPU's."""
    rng = np.random.default_rng(5)
    X = rng.uniform(-1, 1, (800, 8))
    T = np.stack([np.sin(X[:, 0]) + X[:, 1] ** 2, X[:, 2] * X[:, 3], np.exp(0.5 * X[:, 4]),
                  X[:, 5] - 0.3 * X[:, 6] * X[:, 7]], 1)
    net, hist = S.train(gpu_ctx, X, T, steps=1500, lr=3e-3, seed=1)
    assert hist[-1][1] < hist[0][1] / 50
    xs = ((X - net.x_mean) / net.x_std).astype(np.float32)
    ref = M.forward(net.w.astype(np.float64), xs.astype(np.float64))
    assert np.max(np.abs(S.forward(gpu_ctx, net.w, xs) - ref)) < 1e-4
