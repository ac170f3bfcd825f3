"""The CPU restatement of the selector network (oracle/mlp.py) pinned by
mathematics: central finite differences of the loss for every parameter
block, GELU against its closed form values, and Adam's first step (whose
bias-corrected update is lr * sign(g) for any g)."""
import numpy as np

from oracle import mlp as M


def _net(seed):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(M.N_PARAMS) * 0.3


def test_packed_size_matches_the_library():
    assert M.N_PARAMS == 8 * 64 + 64 + 64 * 64 + 64 + 4 * 64 + 4 == 4996


def test_gelu_values_and_derivative():
    z = np.array([-3.0, -1.0, 0.0, 0.5, 2.0])
    # gelu(z) = z Phi(z): Phi(0) = 1/2, Phi(1) = 0.8413447460685429
    assert M.gelu(np.array([0.0]))[0] == 0.0
    assert abs(M.gelu(np.array([1.0]))[0] - 0.8413447460685429) < 1e-15
    assert abs(M.gelu(np.array([-1.0]))[0] + (1 - 0.8413447460685429)) < 1e-15
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    assert np.max(np.abs(fd - M.gelu_grad(z))) < 1e-9


def test_gradient_against_finite_differences():
    w = _net(1)
    rng = np.random.default_rng(2)
    x = rng.standard_normal((7, 8))
    t = rng.standard_normal((7, 4))
    _, g = M.loss_grad(w, x, t)
    h = 1e-6
    # every parameter of the small blocks, a sample of W2
    idx = list(range(0, 8 * 64, 37)) + list(range(512, 576)) + \
        list(range(576, 576 + 4096, 101)) + list(range(4672, 4996))
    for i in idx:
        e = np.zeros_like(w)
        e[i] = h
        fd = (M.loss_grad(w + e, x, t)[0] - M.loss_grad(w - e, x, t)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 + 1e-6 * abs(g[i]), i


def test_loss_is_the_mean_square():
    w = _net(3)
    x = np.random.default_rng(4).standard_normal((5, 8))
    y = M.forward(w, x)
    loss, _ = M.loss_grad(w, x, y + 0.5)
    assert abs(loss - 0.25) < 1e-14


def test_adam_first_step_is_signed_lr():
    g = np.array([3.0, -0.01, 1e-3, -7.0])
    w, m, v = M.adam_step(np.zeros(4), g, np.zeros(4), np.zeros(4), 1, 0.01, eps=0.0)
    assert np.allclose(w, -0.01 * np.sign(g), rtol=0, atol=1e-15)
