"""The free-space Ewald restatement (oracle/free_ewald.py) pinned on the CPU:
the truncated biharmonic transform against numerical quadrature on both sides
of its series/closed-form switch and at its K -> 0 limit, and the whole sum
against the direct free-space sum for two point sets and two splits (the
result must not depend on xi).  Also the host-side parameter choice of the
library (se2p_free_params, no GPU needed): the box relations the free-space
path relies on."""
import math

import numpy as np
import pytest
from scipy import integrate, special

from oracle import free_ewald as FE


@pytest.mark.parametrize("R", [0.7, 1.0, 3.2, 10.0])
@pytest.mark.parametrize("x", [1e-3, 0.3, 1.0, 1.999, 2.0, 2.5, 7.3, 40.0])
def test_bhat_matches_quadrature(R, x):
    K = x / R
    beta = (2 * math.log(R) - 1) / (16 * math.pi)
    BR = R * R * (math.log(R) - 1) / (8 * math.pi)

    def g(r):
        B = r * r * (math.log(r) - 1) / (8 * math.pi)
        return (B - BR - beta * (r * r - R * R)) * special.j0(K * r) * r

    v, _ = integrate.quad(g, 0, R, limit=500, epsabs=0, epsrel=1e-13)
    got = FE.bhat_trunc(np.array([K]), R)[0]
    assert abs(got - 2 * math.pi * v) <= 1e-11 * abs(2 * math.pi * v)


def test_bhat_limit_at_zero():
    R = 2.3
    assert abs(FE.bhat_trunc(np.array([1e-9]), R)[0] - R ** 4 / 64) < 1e-14 * R ** 4 / 64


@pytest.mark.parametrize("shape,xi", [((1.0, 0.6), 12.0), ((1.0, 0.6), 20.0),
                                      ((0.3, 1.1), 15.0)])
def test_free_ewald_matches_direct_sum(shape, xi):
    rng = np.random.default_rng(int(xi))
    p = rng.random((250, 2)) * shape + [0.3, -2.0]  # anywhere: the sum is translation invariant
    f = rng.standard_normal((250, 2))
    u = FE.slp_free(p, f, xi)
    ref = FE.slp_free_direct(p, f)
    assert np.linalg.norm(u - ref) / np.linalg.norm(ref) < 1e-12


def test_free_params_box_relations():
    import ctypes
    from fasttools2d_amd import _lib
    for n, wx, wy, tol in [(1000, 1.0, 1.0, 1e-8), (200000, 2.0, 0.5, 5e-10), (50, 0.0, 0.0, 1e-3),
                           (10 ** 6, 1.0, 1.0, 5e-13), (3000, 1e-3, 2.0, 5e-4)]:
        prm = _lib.Params()
        Lx, Ly = ctypes.c_double(), ctypes.c_double()
        _lib.check(_lib.lib().se2p_free_params(n, n, wx, wy, tol, ctypes.byref(prm),
                                               ctypes.byref(Lx), ctypes.byref(Ly)))
        Lx, Ly = Lx.value, Ly.value
        h = Lx / prm.Mx
        assert abs(Ly / prm.My - h) <= 1e-12 * h
        assert prm.Mx % 2 == 0 and prm.My % 2 == 0 and max(prm.Mx, prm.My) <= 10240
        W = max(wx, wy) or 1.0
        ex, ey = max(wx, 1e-3 * W), max(wy, 1e-3 * W)
        delta = prm.free_R - math.hypot(ex, ey)
        assert delta > 0 and prm.xi * delta > 3
        assert Lx >= ex + prm.free_R + delta - 1e-12 and Ly >= ey + prm.free_R + delta - 1e-12
        # real-space cells at least the cutoff the rules ask for
        L = math.log(10 / tol)
        assert Lx / prm.nside_x * prm.xi >= math.sqrt(max(L - 3.2, 1.0)) - 1e-9
    with pytest.raises(_lib.Se2pError):
        _lib.check(_lib.lib().se2p_free_params(10, 10, 1.0, 1.0, 1e-15, ctypes.byref(prm),
                                               ctypes.byref(ctypes.c_double()),
                                               ctypes.byref(ctypes.c_double())))


@pytest.mark.parametrize("tol", [5e-3, 5e-4, 5e-7, 5e-10])
@pytest.mark.parametrize("dist", ["uniform", "curves"])
def test_library_parameters_meet_tol(oracle, tol, dist):
    """The library's free-space parameters (se2p_free_params) evaluated the
    library's way on the CPU (oracle spreading and gathering, numpy FFTs, the
    free-space filter) meet tol against the direct sum."""
    import ctypes
    from fasttools2d_amd import _lib
    n = 2000
    rng = np.random.default_rng(3)
    if dist == "uniform":
        p = rng.uniform(-1, 1, (n, 2))
    else:  # two curves, as boundary nodes
        th = rng.random(n) * 2 * np.pi
        r = np.where(np.arange(n) % 2 == 0, 0.4, 0.15)
        p = np.stack([r * np.cos(th) + (np.arange(n) % 2) * 0.7, r * np.sin(th)], 1)
    f = 10 * rng.random((n, 2))
    wx, wy = p.max(0) - p.min(0)
    prm = _lib.Params()
    Lx, Ly = ctypes.c_double(), ctypes.c_double()
    _lib.check(_lib.lib().se2p_free_params(n, n, wx, wy, tol, ctypes.byref(prm),
                                           ctypes.byref(Lx), ctypes.byref(Ly)))
    u = FE.slp_free_grid(p, f, prm.as_dict(), Lx.value, Ly.value, oracle)
    ref = oracle.free_slp(f, p, p, self=True)
    e = np.linalg.norm(u - ref) / np.linalg.norm(ref)
    assert e <= 0.2 * tol
