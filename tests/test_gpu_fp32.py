"""The FP32 tier (SE2P_FP32C, for tolerances down to 1e-6) against the oracle.

Every stage the tier changes is checked on its own through the stage-level
entry points of a context set to FP32 (spreading, the 2D FFT against numpy,
gathering, the real-space and k-space mex equivalents of both layers), then the
whole sums of plans and batches.  The bar is 1e-6 relative L2, the tier's
stated floor; the FP32 rounding it measures is near 1e-7."""
import math

import numpy as np
import pytest

import fasttools2d_amd as ft
from fasttools2d_amd import _lib
from fasttools2d_amd import periodic_ewald as pe
from fasttools2d_amd.batch import Batch
from fasttools2d_amd.plan import Plan
from fasttools2d_amd.tune import make_params
from tests.inputs import circles, normal_density, oracle_se, rel, uniform_points

pytestmark = pytest.mark.gpu
TOL32 = 1e-6


@pytest.fixture(scope="module")
def ctx32():
    if ft.device_count() < 1:
        pytest.fail("GPU test collected but no GPU is visible")
    c = ft.Context(0)
    c.set_precision(_lib.SE2P_FP32C)
    yield c
    c.close()


def _gauss(xi, P, L, M):
    w = P * L / M / 2
    return w, (2 * xi * w / (0.95 * math.sqrt(math.pi * P))) ** 2


@pytest.mark.parametrize("N,Mx,My,P", [(3000, 64, 64, 8), (5000, 96, 48, 14), (2000, 50, 34, 16),
                                       (1500, 40, 40, 24), (800, 6, 4, 4), (6000, 96, 200, 12)])
def test_fp32_spread_matches_oracle(ctx32, oracle, N, Mx, My, P):
    Lx = 1.0
    Ly = Lx * My / Mx
    xi = 4.0 * Mx / 64 + 2.0
    w, eta = _gauss(xi, P, Lx, Mx)
    p = uniform_points(N, Lx, Ly, seed=N + Mx)
    f = normal_density(N, seed=N + My)
    H1, H2 = pe.Spread(p.T, f.T, Lx, Ly, xi, w, eta, P, Mx, My, ctx=ctx32)
    R1, R2 = oracle.spread(p, f, Lx, Ly, xi, w, eta, P, Mx, My)
    e = max(rel(H1.T, R1), rel(H2.T, R2))
    assert 1e-10 < e < TOL32  # FP32 arithmetic, within the tier's floor


@pytest.mark.parametrize("My,Mx", [(64, 64), (366, 732), (2160, 24), (24, 2560), (3000, 6),
                                   (902, 8), (10240, 2), (4, 6000)])
def test_fp32_fft2_matches_numpy(ctx32, My, Mx):
    rng = np.random.default_rng(My * 7 + Mx)
    Z = rng.standard_normal((My, Mx)) + 1j * rng.standard_normal((My, Mx))
    for inverse in (False, True):
        got = pe.fft2(Z, inverse=inverse, ctx=ctx32)
        ref = np.fft.ifft2(Z) if inverse else np.fft.fft2(Z)
        assert rel(got, ref) < TOL32


@pytest.mark.parametrize("N,Mx,My,P,Lx,Ly", [(700, 64, 64, 24, 1.0, 1.0), (900, 92, 46, 14, 2.0, 1.0)])
def test_fp32_gather_matches_oracle(ctx32, oracle, N, Mx, My, P, Lx, Ly):
    xi = 10.0
    w, eta = _gauss(xi, P, Lx, Mx)
    G = np.random.default_rng(4).standard_normal((My, Mx))
    p = uniform_points(N, Lx, Ly, seed=N)
    got = pe.Gather(G, p.T, Lx, Ly, xi, w, eta, P, ctx=ctx32)
    ref = oracle.gather(np.ascontiguousarray(G.T), p, Lx, Ly, xi, w, eta, P)
    assert rel(got, ref) < TOL32


@pytest.mark.parametrize("Ns,Nt,nsx,nsy,Lx,Ly,xi", [(6000, 3000, 8, 8, 1.0, 1.0, 30.0),
                                                    (8000, 4000, 12, 6, 2.0, 1.0, 25.0)])
def test_fp32_real_space_matches_oracle(ctx32, oracle, Ns, Nt, nsx, nsy, Lx, Ly, xi):
    ps = uniform_points(Ns, Lx, Ly, seed=Ns)
    pt = np.concatenate([uniform_points(Nt - 50, Lx, Ly, seed=Nt), ps[:50]])  # coincident too
    f = normal_density(Ns, seed=3)
    n = normal_density(Ns, seed=4)
    n /= np.linalg.norm(n, axis=1)[:, None]
    got = pe.mex_stokes_slp_real(ps.T, pt.T, f.T, xi, nsx, nsy, Lx, Ly, ctx=ctx32)
    ref = oracle.slp_real(ps, pt, f, xi, nsx, nsy, Lx, Ly)
    assert rel(got.T, ref) < TOL32
    got = pe.mex_stokes_dlp_real(ps.T, pt.T, f.T, n.T, xi, nsx, nsy, Lx, Ly, ctx=ctx32)
    ref = oracle.dlp_real(ps, pt, f, n, xi, nsx, nsy, Lx, Ly)
    assert rel(got.T, ref) < TOL32


# the third: 520 x 260, the Stokeslet on the fused x passes
@pytest.mark.parametrize("Lx,Ly,M,P", [(1.0, 1.0, 128, 12), (2.0, 1.0, 160, 16), (2.0, 1.0, 260, 12)])
def test_fp32_kspace_matches_oracle(ctx32, oracle, Lx, Ly, M, P):
    Ns, Nt, xi = 5000, 2500, 20.0
    Mx, My = int(M * Lx), int(M * Ly)
    w, eta = _gauss(xi, P, Lx, Mx)
    ps = uniform_points(Ns, Lx, Ly, seed=5)
    pt = uniform_points(Nt, Lx, Ly, seed=6)
    f = normal_density(Ns, seed=7)
    n = normal_density(Ns, seed=8)
    n /= np.linalg.norm(n, axis=1)[:, None]
    got = pe.mex_stokes_slp_kspace(ps.T, pt.T, xi, eta, f.T, Mx, My, Lx, Ly, w, P, ctx=ctx32)
    ref = oracle.slp_kspace(ps, pt, xi, eta, f, Mx, My, Lx, Ly, w, P)
    assert rel(got.T, ref) < TOL32
    got = pe.mex_stokes_dlp_kspace(ps.T, pt.T, xi, eta, f.T, n.T, Mx, My, Lx, Ly, w, P, ctx=ctx32)
    ref = oracle.dlp_kspace(ps, pt, xi, eta, f, n, Mx, My, Lx, Ly, w, P)
    assert rel(got.T, ref) < TOL32


@pytest.mark.parametrize("kind,same", [("slp", True), ("slp", False), ("dlp", False)])
def test_fp32_plan_total_within_tolerance(gpu_ctx, oracle, kind, same):
    """Plans at parameters for tol 1e-6 in both precisions: the FP32 plan is
    within 1e-6 of the oracle's accurate sum, and within the FP32 floor of the
    FP64 plan at the same parameters."""
    N, Lx, Ly = 20000, 1.0, 1.0
    p = uniform_points(N, Lx, Ly, seed=21)
    pt = p if same else p.copy()
    f = normal_density(N, seed=22)
    n = None
    if kind == "dlp":
        n = normal_density(N, seed=23)
        n /= np.linalg.norm(n, axis=1)[:, None]
    prm = ft.ewald_params(kind, N, N, ft.density_q(f[:, 0], f[:, 1]), Lx, Ly, P=12, tol=1e-6)
    prm["symmetric"] = 1 if same else 0
    out = {}
    for prec in (0, 1):
        pl = Plan(gpu_ctx, kind, p, pt, Lx, Ly, dict(prm, precision=prec), nrm=n)
        try:
            out[prec] = pl.apply(f)
            used = pl.params_in_use()
        finally:
            pl.close()
        assert used["precision"] == prec and used["symmetric"] == (1 if same else 0)
    sub = np.arange(0, N, 10)
    ref = oracle_se(oracle, kind, p, pt[sub], f, n, Lx, Ly)
    assert rel(out[1][sub], ref) <= TOL32
    assert rel(out[1], out[0]) < 5e-7


def test_fp32_not_applied_to_derived_kinds(gpu_ctx):
    N = 2000
    p = uniform_points(N, 1.0, 1.0, seed=31)
    prm = ft.ewald_params("slp_pressure", N, N, 2.0 * N + 1, 1.0, 1.0, tol=1e-6)
    pl = Plan(gpu_ctx, "slp_pressure", p, p.copy(), 1.0, 1.0, dict(prm, precision=1))
    try:
        assert pl.params_in_use()["precision"] == 0
    finally:
        pl.close()


def test_fp32_mixed_precision_batch_matches_plans(gpu_ctx):
    """A batch whose systems run in different precisions gives each system what
    its own plan gives, bit for bit (fixed-order sums)."""
    systems = []
    for i, (kind, prec, Lx) in enumerate([("slp", 1, 1.0), ("slp", 0, 2.0), ("dlp", 1, 2.0),
                                          ("dlp", 0, 1.0), ("slp", 1, 1.5)]):
        N = 3000 + 1000 * i
        Ly = 1.0
        if kind == "dlp":
            ps, n, f = circles(4, N // 4, Lx, Ly, seed=40 + i)
        else:
            ps, n = uniform_points(N, Lx, Ly, seed=40 + i), None
        pt = uniform_points(N // 2, Lx, Ly, seed=50 + i)
        prm = ft.ewald_params(kind, len(ps), len(pt), 2.0 * N + 1, Lx, Ly, P=12, tol=1e-6)
        systems.append(dict(kind=kind, psrc=ps, ptar=pt, nrm=n, Lx=Lx, Ly=Ly,
                            prm=dict(prm, precision=prec)))
    fs = [normal_density(len(y["psrc"]), seed=60 + i) for i, y in enumerate(systems)]
    b = Batch(gpu_ctx, systems)
    try:
        got = b.apply(fs)
        for i, y in enumerate(systems):
            assert b.params_in_use(i)["precision"] == y["prm"]["precision"]
            pl = Plan(gpu_ctx, y["kind"], y["psrc"], y["ptar"], y["Lx"], y["Ly"], y["prm"],
                      nrm=y["nrm"])
            try:
                np.testing.assert_array_equal(got[i], pl.apply(fs[i]))
            finally:
                pl.close()
    finally:
        b.close()


def test_fp32_c2_stresslet_circles_full_size(oracle, ctx32):
    """Configuration 2's FP32 variant: 20 circles x 2000 nodes, 40 000 targets,
    box 2 x 1, the driver at tol 1e-6 in FP32."""
    Lx, Ly, tol = 2.0, 1.0, 1e-6
    ps, n, f = circles(20, 2000, Lx, Ly, seed=17)
    pt = uniform_points(40000, Lx, Ly, seed=18)
    u1, u2, _, _, _ = ft.StokesDLP_ewald_2p(ps[:, 0], ps[:, 1], pt[:, 0], pt[:, 1], n[:, 0],
                                            n[:, 1], f[:, 0], f[:, 1], Lx, Ly, tol=tol, ctx=ctx32)
    u = np.stack([u1, u2], 1)
    truth = oracle_se(oracle, "dlp", ps, pt, f, n, Lx, Ly, tol=1e-14)
    err, emax = rel(u, truth), float(np.abs(u - truth).max())
    print(f"C2 FP32: rel-L2 error {err:.3e}, max abs {emax:.3e}")
    assert err <= tol and emax <= tol
