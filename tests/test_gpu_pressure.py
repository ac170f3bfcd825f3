"""Stokeslet and stresslet pressure on the GPU, through the C ABI: the mex
replacements and the drivers against the oracle
(mex_stokes_{slp,dlp}_pressure_{real,kspace}.cpp, Stokes{SLP,DLP}_pressure_ewald_2p.m),
plans and batches bit for bit against the one-shot path."""
import numpy as np
import pytest

import fasttools2d_amd as ft
from fasttools2d_amd.batch import Batch
from fasttools2d_amd.plan import Plan
from tests.inputs import circles, normal_density, uniform_points

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("N,Lx,Ly,nsx,nsy", [(3000, 1.0, 1.0, 8, 8), (2500, 2.0, 1.0, 10, 5)])
def test_pressure_real_matches_oracle(gpu_ctx, oracle, N, Lx, Ly, nsx, nsy):
    ps = uniform_points(N, Lx, Ly, seed=201)
    pt = np.concatenate([uniform_points(N // 2, Lx, Ly, seed=202), ps[:50]])  # coincident
    f = normal_density(N, seed=203)
    xi = 3.5 * nsx / Lx
    got = ft.mex_stokes_slp_pressure_real(ps.T, pt.T, f.T, xi, nsx, nsy, Lx, Ly, ctx=gpu_ctx)
    assert got.shape == (1, len(pt))
    ref = oracle.slp_pressure_real(ps, pt, f, xi, nsx, nsy, Lx, Ly)
    assert _rel(got[0], ref) < 1e-13


@pytest.mark.parametrize("Lx,Ly,tol", [(1.0, 1.0, 1e-12), (2.0, 1.0, 1e-8)])
def test_pressure_kspace_matches_oracle(gpu_ctx, oracle, Lx, Ly, tol):
    N = 4000
    ps = uniform_points(N, Lx, Ly, seed=211)
    pt = uniform_points(3000, Lx, Ly, seed=212)
    f = normal_density(N, seed=213)
    prm = ft.ewald_params("slp_pressure", N, 3000, ft.density_q(f[:, 0], f[:, 1]), Lx, Ly, tol=tol)
    got = ft.mex_stokes_slp_pressure_kspace(ps.T, pt.T, f.T, prm["xi"], prm["eta"], prm["Mx"],
                                            prm["My"], Lx, Ly, prm["w"], prm["P"], ctx=gpu_ctx)
    ref = oracle.slp_pressure_kspace(ps, pt, f, prm["xi"], prm["eta"], prm["Mx"], prm["My"], Lx,
                                     Ly, prm["w"], prm["P"])
    assert _rel(got[0], ref) < 1e-13


def test_pressure_driver_matches_oracle(gpu_ctx, oracle):
    Lx, Ly = 1.0, 2.0
    ps = uniform_points(5000, Lx, Ly, seed=221) + [0.3, -0.7]  # the driver wraps
    pt = uniform_points(3000, Lx, Ly, seed=222)
    f = normal_density(5000, seed=223)
    p, pr, pk, xi = ft.StokesSLP_pressure_ewald_2p(ps[:, 0], ps[:, 1], pt[:, 0], pt[:, 1],
                                                   f[:, 0], f[:, 1], Lx, Ly, tol=1e-12,
                                                   ctx=gpu_ctx)
    np.testing.assert_array_equal(p, pr + pk)
    psw = oracle.wrap(ps, Lx, Ly)
    prm = ft.ewald_params("slp_pressure", 5000, 3000, ft.density_q(f[:, 0], f[:, 1]), Lx, Ly, tol=1e-12)
    assert abs(prm["xi"] - xi) <= 1e-12 * xi  # Q summed in another order
    r_r = oracle.slp_pressure_real(psw, pt, f, xi, prm["nside_x"], prm["nside_y"], Lx, Ly)
    r_k = oracle.slp_pressure_kspace(psw, pt, f, xi, prm["eta"], prm["Mx"], prm["My"], Lx, Ly,
                                     prm["w"], prm["P"]) + oracle.slp_pressure_zero_mode(
                                         psw, f, Lx, Ly)
    assert _rel(pr, r_r) < 1e-13
    assert _rel(pk, r_k) < 1e-13


def test_pressure_plan_graph_and_batch_bitwise(gpu_ctx):
    Lx, Ly = 1.5, 1.0
    ps = uniform_points(6000, Lx, Ly, seed=231)
    pt = uniform_points(4000, Lx, Ly, seed=232)
    prm = ft.ewald_params("slp_pressure", 6000, 4000, 2 * 6000 + 1.0, Lx, Ly, tol=1e-10)
    plan = Plan(gpu_ctx, "slp_pressure", ps, pt, Lx, Ly, prm)
    systems = [dict(kind="slp", psrc=ps, ptar=pt, nrm=None, Lx=Lx, Ly=Ly, prm=prm),
               dict(kind="slp_pressure", psrc=ps, ptar=pt, nrm=None, Lx=Lx, Ly=Ly, prm=prm)]
    b = Batch(gpu_ctx, systems)
    vplan = Plan(gpu_ctx, "slp", ps, pt, Lx, Ly, prm)
    for seed in (1, 2):
        f = normal_density(6000, seed=seed)
        p = plan.apply(f)
        assert p.shape == (4000,)
        for _ in range(2):
            np.testing.assert_array_equal(plan.apply(f, graph=True), p)
        u_b, p_b = b.apply([f, f])
        np.testing.assert_array_equal(p_b, p)
        np.testing.assert_array_equal(u_b, vplan.apply(f))
        # the one-shot driver path gives the same numbers as the plan
        pd, _, _, _ = ft.StokesSLP_pressure_ewald_2p(ps[:, 0], ps[:, 1], pt[:, 0], pt[:, 1],
                                                     f[:, 0], f[:, 1], Lx, Ly, tol=1e-10,
                                                     ctx=gpu_ctx)
        assert _rel(pd, p) < 1e-9
    plan.close()
    vplan.close()
    b.close()


def _unit_normals(N, seed):
    rng = np.random.default_rng(seed)
    th = rng.uniform(0, 2 * np.pi, N)
    return np.stack([np.cos(th), np.sin(th)], 1)


@pytest.mark.parametrize("N,Lx,Ly,nsx,nsy", [(3000, 1.0, 1.0, 8, 8), (2500, 2.0, 1.0, 10, 5)])
def test_dlp_pressure_real_matches_oracle(gpu_ctx, oracle, N, Lx, Ly, nsx, nsy):
    ps = uniform_points(N, Lx, Ly, seed=301)
    pt = np.concatenate([uniform_points(N // 2, Lx, Ly, seed=302), ps[:50]])  # coincident
    f = normal_density(N, seed=303)
    n = _unit_normals(N, seed=304)
    xi = 3.5 * nsx / Lx
    got = ft.mex_stokes_dlp_pressure_real(ps.T, pt.T, f.T, n.T, xi, nsx, nsy, Lx, Ly,
                                          ctx=gpu_ctx)
    assert got.shape == (1, len(pt))
    ref = oracle.dlp_pressure_real(ps, pt, f, n, xi, nsx, nsy, Lx, Ly)
    assert _rel(got[0], ref) < 1e-13


@pytest.mark.parametrize("Lx,Ly,tol", [(1.0, 1.0, 1e-12), (2.0, 1.0, 1e-8)])
def test_dlp_pressure_kspace_matches_oracle(gpu_ctx, oracle, Lx, Ly, tol):
    N = 4000
    ps = uniform_points(N, Lx, Ly, seed=311)
    pt = uniform_points(3000, Lx, Ly, seed=312)
    f = normal_density(N, seed=313)
    n = _unit_normals(N, seed=314)
    prm = ft.ewald_params("dlp_pressure", N, 3000, ft.density_q(f[:, 0], f[:, 1]), Lx, Ly, tol=tol)
    got = ft.mex_stokes_dlp_pressure_kspace(ps.T, pt.T, f.T, n.T, prm["xi"], prm["eta"],
                                            prm["Mx"], prm["My"], Lx, Ly, prm["w"], prm["P"],
                                            ctx=gpu_ctx)
    ref = oracle.dlp_pressure_kspace(ps, pt, f, n, prm["xi"], prm["eta"], prm["Mx"], prm["My"],
                                     Lx, Ly, prm["w"], prm["P"])
    assert _rel(got[0], ref) < 1e-13


def test_dlp_pressure_driver_matches_oracle(gpu_ctx, oracle):
    Lx, Ly = 1.0, 2.0
    ps, n, f = circles(4, 300, Lx, Ly, seed=321)
    ps = ps + [0.3, -0.7]  # the driver wraps
    pt = uniform_points(3000, Lx, Ly, seed=322)
    p, pr, pk, xi = ft.StokesDLP_pressure_ewald_2p(ps[:, 0], ps[:, 1], pt[:, 0], pt[:, 1],
                                                   n[:, 0], n[:, 1], f[:, 0], f[:, 1], Lx, Ly,
                                                   tol=1e-12, ctx=gpu_ctx)
    np.testing.assert_array_equal(p, pr + pk)
    psw = oracle.wrap(ps, Lx, Ly)
    Ns = len(ps)
    prm = ft.ewald_params("dlp_pressure", Ns, 3000, ft.density_q(f[:, 0], f[:, 1]), Lx, Ly, tol=1e-12)
    assert abs(prm["xi"] - xi) <= 1e-12 * xi
    r_r = oracle.dlp_pressure_real(psw, pt, f, n, xi, prm["nside_x"], prm["nside_y"], Lx, Ly)
    r_k = oracle.dlp_pressure_kspace(psw, pt, f, n, xi, prm["eta"], prm["Mx"], prm["My"], Lx, Ly,
                                     prm["w"], prm["P"]) + oracle.dlp_pressure_zero_mode(
                                         f, n, Lx, Ly)
    assert _rel(pr, r_r) < 1e-13
    assert _rel(pk, r_k) < 1e-13


def test_dlp_pressure_plan_graph_and_batch_bitwise(gpu_ctx):
    Lx, Ly = 1.5, 1.0
    ps, n, _ = circles(5, 400, Lx, Ly, seed=331)
    pt = uniform_points(4000, Lx, Ly, seed=332)
    Ns = len(ps)
    prm = ft.ewald_params("dlp_pressure", Ns, 4000, 2 * Ns + 1.0, Lx, Ly, tol=1e-10)
    plan = Plan(gpu_ctx, "dlp_pressure", ps, pt, Lx, Ly, prm, nrm=n)
    systems = [dict(kind="dlp", psrc=ps, ptar=pt, nrm=n, Lx=Lx, Ly=Ly, prm=prm),
               dict(kind="dlp_pressure", psrc=ps, ptar=pt, nrm=n, Lx=Lx, Ly=Ly, prm=prm),
               dict(kind="slp_pressure", psrc=ps, ptar=pt, nrm=None, Lx=Lx, Ly=Ly, prm=prm)]
    b = Batch(gpu_ctx, systems)
    vplan = Plan(gpu_ctx, "dlp", ps, pt, Lx, Ly, prm, nrm=n)
    splan = Plan(gpu_ctx, "slp_pressure", ps, pt, Lx, Ly, prm)
    for seed in (1, 2):
        f = normal_density(Ns, seed=seed)
        p = plan.apply(f)
        assert p.shape == (4000,)
        for _ in range(2):
            np.testing.assert_array_equal(plan.apply(f, graph=True), p)
        u_b, p_b, s_b = b.apply([f, f, f])
        np.testing.assert_array_equal(p_b, p)
        np.testing.assert_array_equal(u_b, vplan.apply(f))
        np.testing.assert_array_equal(s_b, splan.apply(f))
        pd, _, _, _ = ft.StokesDLP_pressure_ewald_2p(ps[:, 0], ps[:, 1], pt[:, 0], pt[:, 1],
                                                     n[:, 0], n[:, 1], f[:, 0], f[:, 1], Lx, Ly,
                                                     tol=1e-10, ctx=gpu_ctx)
        assert _rel(pd, p) < 1e-9
    plan.close()
    vplan.close()
    splan.close()
    b.close()
