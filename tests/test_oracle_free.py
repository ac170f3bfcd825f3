"""Pins of the oracle's free-space direct sums (the sums FMM/tests/
stokes_{SLP,DLP}_FMM_test.m compare the FMM with): the stresslet identity of
FMM/tests/stresslet_identity_test.m, incompressibility, the Stokes equation
with the textbook Stokeslet pressure p = (r.f)/(2 pi rho^2), and the single
Stokeslet in closed form."""
import math

import numpy as np
import pytest


def _circle(n, r=1.0):
    h = 2 * np.pi / n
    th = h * np.arange(1, n + 1)
    return np.stack([r * np.cos(th), r * np.sin(th)], 1), np.stack([np.cos(th), np.sin(th)], 1), h


@pytest.mark.parametrize("comp", [0, 1])
def test_stresslet_identity(oracle, comp):
    """-u = e_comp inside the unit circle, 0 outside (stresslet_identity_test.m:9-56)."""
    ps, n, h = _circle(512)
    th = 2 * np.pi / 16 * np.arange(1, 17)
    tin = 0.5 * np.stack([np.cos(th), np.sin(th)], 1)
    tout = 1.5 * np.stack([np.cos(th), np.sin(th)], 1)
    p = np.concatenate([ps, tin, tout])
    f = np.zeros((len(p), 2))
    f[:512, comp] = h
    nn = np.concatenate([n, np.zeros((32, 2))])
    u = -oracle.free_dlp(f, nn, p)
    want = np.zeros(2)
    want[comp] = 1
    assert np.abs(u[512:528] - want).max() < 1e-12
    assert np.abs(u[528:]).max() < 1e-12


def test_single_stokeslet_closed_form(oracle):
    f = np.array([[1.0, 2.0]])
    src = np.array([[0.3, -0.2]])
    tar = np.array([[1.3, -0.2], [0.3, 0.8], [2.3, 1.8]])
    u = oracle.free_slp(f, src, tar)
    for k in range(3):
        r = tar[k] - src[0]
        rho2 = r @ r
        want = (-0.5 * math.log(rho2) * f[0] + (r @ f[0]) * r / rho2) / (4 * np.pi)
        assert np.abs(u[k] - want).max() < 1e-15


def _fd_fields(fun, x, h):
    """Central differences of a 2-vector field: J[:, i, j] = d u_i / d x_j, and
    the Laplacian."""
    J = np.zeros((len(x), 2, 2))
    lap = np.zeros((len(x), 2))
    u0 = fun(x)
    for j in range(2):
        e = np.zeros(2)
        e[j] = h
        up, um = fun(x + e), fun(x - e)
        J[:, :, j] = (up - um) / (2 * h)
        lap += (up - 2 * u0 + um) / (h * h)
    return J, lap


def test_slp_is_incompressible_and_solves_stokes(oracle):
    rng = np.random.default_rng(3)
    src = rng.uniform(-1, 1, (40, 2))
    f = rng.standard_normal((40, 2))
    x = rng.uniform(2, 3, (25, 2))  # away from the sources
    J, lap = _fd_fields(lambda t: oracle.free_slp(f, src, t), x, 1e-3)
    scale = np.abs(J).max()
    assert np.abs(J[:, 0, 0] + J[:, 1, 1]).max() < 1e-6 * scale
    # -lap u + grad p = 0 with p = sum_s (r.f_s)/(2 pi rho^2)
    def pres(t):
        r = t[:, None, :] - src[None, :, :]
        return ((r * f[None]).sum(2) / (r ** 2).sum(2)).sum(1) / (2 * np.pi)
    h = 1e-4
    gp = np.stack([(pres(x + [h, 0]) - pres(x - [h, 0])) / (2 * h),
                   (pres(x + [0, h]) - pres(x - [0, h])) / (2 * h)], 1)
    assert np.abs(lap - gp).max() < 1e-5 * np.abs(gp).max()


def test_dlp_is_incompressible(oracle):
    rng = np.random.default_rng(4)
    src = rng.uniform(-1, 1, (40, 2))
    th = rng.uniform(0, 2 * np.pi, 40)
    n = np.stack([np.cos(th), np.sin(th)], 1)
    f = rng.standard_normal((40, 2))
    x = rng.uniform(2, 3, (25, 2))

    def field(t):  # the DLP at free targets: zero-density targets appended to the sources
        p = np.concatenate([src, t])
        ff = np.concatenate([f, np.zeros_like(t)])
        nn = np.concatenate([n, np.zeros_like(t)])
        return oracle.free_dlp(ff, nn, p)[40:]
    J, _ = _fd_fields(field, x, 1e-3)
    assert np.abs(J[:, 0, 0] + J[:, 1, 1]).max() < 1e-6 * np.abs(J).max()


def test_self_terms_are_left_out(oracle):
    rng = np.random.default_rng(5)
    p = rng.uniform(-1, 1, (30, 2))
    f = rng.standard_normal((30, 2))
    u = oracle.free_slp(f, p, p, self=True)
    for k in (0, 17):
        idx = [s for s in range(30) if s != k]
        want = oracle.free_slp(f[idx], p[idx], p[k:k + 1])
        assert np.abs(u[k] - want[0]).max() < 1e-14
