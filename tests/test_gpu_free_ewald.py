"""The free-space spectral Ewald Stokeslet sum (the fast replacement of the FMM
behind stokesSLPfmm, se2p_params.free_R) against direct sums: the oracle's on
the CPU, and the library's own direct GPU sum at sizes the oracle would take
long for.  The bar is the FMM's tolerance for iprec (cfmm2dpart.f:351-361)."""
import numpy as np
import pytest

import fasttools2d_amd as ft
from fasttools2d_amd import _lib
from fasttools2d_amd import freespace as fs
from fasttools2d_amd.plan import Plan
from fasttools2d_amd.workloads import clusters

pytestmark = pytest.mark.gpu
TOL = {0: 5e-3, 1: 5e-4, 2: 5e-7, 3: 5e-10, 4: 5e-13}


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def ewald_ctx():
    if ft.device_count() < 1:
        pytest.fail("GPU test collected but no GPU is visible")
    c = ft.Context(0)
    c.set_free_method(_lib.SE2P_FREE_EWALD)
    yield c
    c.close()


def _points(kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "uniform":  # stokes_SLP_FMM_test.m:9-20
        return rng.uniform(-1, 1, (n, 2))
    if kind == "clusters":
        return clusters(n, 2.0, 1.0, seed, ncl=12, sigma=0.03) + [5.0, -3.0]
    th = rng.random(n) * 2 * np.pi  # two curves, as boundary nodes
    r = np.where(np.arange(n) % 2 == 0, 0.4, 0.15)
    return np.stack([r * np.cos(th) + (np.arange(n) % 2) * 0.7, r * np.sin(th)], 1)


@pytest.mark.parametrize("iprec", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("dist", ["uniform", "clusters", "curves"])
def test_free_ewald_sources_as_targets(ewald_ctx, oracle, iprec, dist):
    n = 6000
    p = _points(dist, n, 7 + iprec)
    f = 10 * np.random.default_rng(iprec).random((n, 2))
    u1, u2 = ft.stokesSLPfmm(f[:, 0], f[:, 1], p[:, 0], p[:, 1], p[:, 0], p[:, 1], 1, iprec,
                             ctx=ewald_ctx)
    ref = oracle.free_slp(f, p, p, self=True)
    e = _rel(np.stack([u1, u2], 1), ref)
    print(f"{dist} iprec {iprec}: rel err {e:.2e} (tol {TOL[iprec]:.0e})")
    assert e <= TOL[iprec]


@pytest.mark.parametrize("iprec", [1, 3, 4])
def test_free_ewald_separate_targets(ewald_ctx, oracle, iprec):
    rng = np.random.default_rng(20 + iprec)
    src = rng.uniform(-1, 1, (5000, 2))
    tar = np.concatenate([rng.uniform(-1.2, 1.2, (2990, 2)), src[:10]])  # 10 on sources
    f = rng.standard_normal((5000, 2)) + 0.5
    u1, u2 = ft.stokesSLPfmm(f[:, 0], f[:, 1], src[:, 0], src[:, 1], tar[:, 0], tar[:, 1], 0,
                             iprec, ctx=ewald_ctx)
    ref = oracle.free_slp(f, src, tar)
    e = _rel(np.stack([u1, u2], 1), ref)
    print(f"separate targets iprec {iprec}: rel err {e:.2e}")
    assert e <= TOL[iprec]


def test_free_auto_takes_ewald_for_large_sums():
    """The default (automatic) method sums directly below 1e9 pairs and by
    spectral Ewald above; both agree with the forced direct sum."""
    n = 40000
    rng = np.random.default_rng(3)
    p = rng.uniform(-1, 1, (n, 2))
    f = 10 * rng.random((n, 2))
    auto = ft.Context(0)
    direct = ft.Context(0)
    direct.set_free_method(_lib.SE2P_FREE_DIRECT)
    try:
        auto.prof_enable(True)
        auto.prof_reset()
        ua = np.stack(ft.stokesSLPfmm(f[:, 0], f[:, 1], p[:, 0], p[:, 1], p[:, 0], p[:, 1], 1, 3,
                                      ctx=auto), 1)
        auto.sync()
        ran_free = auto.prof_get("free")[1]
        ran_real = auto.prof_get("real")[1]
        ud = np.stack(ft.stokesSLPfmm(f[:, 0], f[:, 1], p[:, 0], p[:, 1], p[:, 0], p[:, 1], 1, 3,
                                      ctx=direct), 1)
    finally:
        auto.close()
        direct.close()
    assert ran_free == 0 and ran_real == 1  # the spectral Ewald stages ran, not the direct sum
    e = _rel(ua, ud)
    print(f"auto (spectral Ewald) vs direct at N={n}: {e:.2e}")
    assert e <= TOL[3]


def test_free_plan_repeated_applies(gpu_ctx, oracle):
    """free_slp_plan: a fixed geometry evaluated for several densities, sources
    as targets (pair-symmetric) and separate targets."""
    rng = np.random.default_rng(5)
    p = rng.uniform(0, 3, (8000, 2))
    tar = rng.uniform(0.5, 2.5, (3000, 2))
    pl, _ = fs.free_slp_plan(gpu_ctx, p, p, 1e-9)
    pt, _ = fs.free_slp_plan(gpu_ctx, p, tar, 1e-9)
    try:
        assert pl.params_in_use()["free_R"] > 0
        for seed in (1, 2):
            f = np.random.default_rng(seed).standard_normal((8000, 2))
            assert _rel(pl.apply(f), oracle.free_slp(f, p, p, self=True)) < 1e-9
            assert _rel(pt.apply(f), oracle.free_slp(f, p, tar)) < 1e-9
    finally:
        pl.close()
        pt.close()


def test_free_ewald_degenerate_extents(ewald_ctx, oracle):
    rng = np.random.default_rng(9)
    line = np.stack([np.full(3000, 0.3), rng.uniform(-1, 1, 3000)], 1)  # zero width
    f = rng.standard_normal((3000, 2))
    u = np.stack(ft.stokesSLPfmm(f[:, 0], f[:, 1], line[:, 0], line[:, 1], line[:, 0],
                                 line[:, 1], 1, 3, ctx=ewald_ctx), 1)
    assert _rel(u, oracle.free_slp(f, line, line, self=True)) <= TOL[3]
    same = np.tile([[0.1, 0.2]], (50, 1))  # coincident points: every pair skipped
    g = rng.standard_normal((50, 2))
    u = np.stack(ft.stokesSLPfmm(g[:, 0], g[:, 1], same[:, 0], same[:, 1], same[:, 0],
                                 same[:, 1], 1, 3, ctx=ewald_ctx), 1)
    assert np.abs(u).max() < 1e-10 * np.abs(g).sum()


def test_free_R_is_stokeslet_only(gpu_ctx):
    p = np.random.default_rng(1).uniform(-0.5, 0.5, (500, 2))
    prm, Lx, Ly = fs.ewald_params(500, 500, 1.0, 1.0, 1e-6)
    n = np.tile([[1.0, 0.0]], (500, 1))
    with pytest.raises(ft.Se2pError):
        Plan(gpu_ctx, "dlp", p, p, Lx, Ly, prm, nrm=n)
