"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Integer and index work (binning, start nodes) must be bit-identical; the FP64
sums must agree with the oracle to 1e-13 relative (L2 over all targets), a few
hundred ulps of accumulated rounding, far below the 1e-8 accuracy the method
is run at.
"""
import json
import math
import os

import numpy as np
import pytest

import fasttools2d_amd as ft
from fasttools2d_amd import _lib
from fasttools2d_amd import periodic_ewald as pe
from fasttools2d_amd.workloads import unit_normals
from tests.inputs import circles, normal_density, uniform_points

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
RTOL = 1e-13


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


# ---------------------------------------------------------------- a1, a7
@pytest.mark.parametrize("n,nsx,nsy,Lx,Ly", [(5000, 4, 4, 1.0, 1.0), (20000, 22, 11, 2.0, 1.0),
                                              (3000, 1, 1, 1.0, 1.0), (0, 3, 3, 1.0, 1.0),
                                              (100000, 63, 63, 1.0, 1.0)])
def test_assign_bit_identical(gpu_ctx, oracle, n, nsx, nsy, Lx, Ly):
    p = uniform_points(n, Lx, Ly, seed=n + nsx)
    if n:
        p[:4] = [[-Lx / 2, -Ly / 2], [-Lx / 2 + Lx / nsx, 0.0], [0.0, -Ly / 2 + Ly / nsy],
                 [np.nextafter(Lx / 2, 0), np.nextafter(Ly / 2, 0)]]
    po, bo, nb = pe.Assign(p.T, Lx, Ly, nsx, nsy)
    ro, rb, rn = oracle.assign(p, Lx, Ly, nsx, nsy)
    np.testing.assert_array_equal(po, ro)
    np.testing.assert_array_equal(bo, rb)
    np.testing.assert_array_equal(nb, rn)


def test_assign_large_bucket_path(gpu_ctx, oracle):
    # 9000 points in one cell exceeds the in-LDS bucket sort
    p = uniform_points(12000, 1.0, 1.0, seed=5) * 0.05
    po, bo, nb = pe.Assign(p.T, 1.0, 1.0, 3, 3)
    ro, rb, rn = oracle.assign(p, 1.0, 1.0, 3, 3)
    assert rn.max() > 4096
    np.testing.assert_array_equal(po, ro)
    np.testing.assert_array_equal(bo, rb)


def test_find_closest_node_bit_identical(gpu_ctx, oracle):
    L, M, P = 1.0, 902, 24
    h = L / M
    p = uniform_points(4000, L, L, seed=8)
    k = np.arange(40)
    on = -L / 2 + k * 23 * h
    extra = np.stack([np.concatenate([on, on - 3e-14, on + 3e-14]),
                      np.concatenate([on[::-1], on + 5e-14, on - 5e-14])], 1)
    p = np.concatenate([p, extra])
    mx, my, px, py = pe.FindClosestNode(p.T, L, L, h, P)
    for i in range(len(p)):
        r = oracle.find_closest_node(p[i, 0], p[i, 1], L, L, h, P)
        assert (mx[i], my[i], px[i], py[i]) == r, i


# ---------------------------------------------------------------- a4
def test_e1_against_mpmath(gpu_ctx):
    rows = json.load(open(os.path.join(GOLD, "e1_mpmath.json")))["rows"]
    x = np.array([float(a) for a, _ in rows])
    ref = np.array([float(b) for _, b in rows])
    keep = (ref > 1e-300) & (x < 650)
    x, ref = x[keep], ref[keep]
    got = pe.expint_e1(x, tier=0)
    assert np.max(np.abs(got - ref) / ref) < 4e-16 * 8
    fast = pe.expint_e1(x, tier=1)
    assert np.max(np.abs(fast - ref) / ref) < 2e-12


# ---------------------------------------------------------------- a8, a12, a9
@pytest.mark.parametrize("N,M,P,L", [(500, 64, 24, 1.0), (3000, 90, 16, 1.0), (200, 30, 24, 1.0),
                                     (2000, 128, 8, 2.0)])
def test_spread_matches_oracle(gpu_ctx, oracle, N, M, P, L):
    xi = 8.0 * M / 64 / L
    h = L / M
    w = P * h / 2
    eta = (2 * xi * w / (0.95 * math.sqrt(math.pi * P))) ** 2
    p = uniform_points(N, L, L, seed=N)
    f = normal_density(N, seed=N + 1)
    H1, H2 = pe.Spread(p.T, f.T, L, L, xi, w, eta, P, M, M)
    R1, R2 = oracle.spread(p, f, L, L, xi, w, eta, P, M, M)
    assert _rel(H1.T, R1) < RTOL and _rel(H2.T, R2) < RTOL


@pytest.mark.parametrize("N,Mx,My,P", [(300, 8, 8, 16),    # window over 4 periodic images
                                       (200, 2, 4, 2),     # the smallest grids
                                       (400, 10, 6, 40),   # largest support, 7+ images
                                       (900, 50, 34, 16),  # last blocks partial (not 16 | M)
                                       (1500, 96, 48, 40),
                                       (2000, 130, 66, 24)])
def test_spread_block_edges_match_oracle(gpu_ctx, oracle, N, Mx, My, P):
    """The block spreading's edge cases: grids narrower than a block or than the
    window of start nodes (a point reaches a block through several periodic
    images), partial last blocks, non-square grids, P up to 40."""
    Lx = 1.0
    Ly = Lx * My / Mx
    xi = 4.0 * Mx / 64 + 2.0
    h = Lx / Mx
    w = P * h / 2
    eta = (2 * xi * w / (0.95 * math.sqrt(math.pi * P))) ** 2
    p = uniform_points(N, Lx, Ly, seed=N + Mx)
    f = normal_density(N, seed=N + My)
    H1, H2 = pe.Spread(p.T, f.T, Lx, Ly, xi, w, eta, P, Mx, My)
    R1, R2 = oracle.spread(p, f, Lx, Ly, xi, w, eta, P, Mx, My)
    assert _rel(H1.T, R1) < RTOL and _rel(H2.T, R2) < RTOL


@pytest.mark.parametrize("N,Mx,My,P,dist", [
    (300, 6, 4, 2, "uniform"),       # grid narrower than a strip: many images each way
    (500, 16, 6, 14, "uniform"),     # one strip; rows wrap inside the first segment
    (900, 50, 34, 6, "uniform"),     # partial last strip
    (3000, 96, 200, 14, "uniform"),  # several segments, the last one partial
    (4000, 130, 66, 10, "uniform"),
    (2500, 40, 300, 12, "uniform"),
    (20000, 128, 128, 14, "clusters"),  # dense start rows: chunks spanning many rows
    (20000, 64, 256, 4, "clusters"),
])
def test_spread_ring_edges_match_oracle(gpu_ctx, oracle, N, Mx, My, P, dist):
    """The strip spreading (supports of up to 16 nodes): grids narrower than a
    strip, periodic images in both directions, partial strips and segments,
    and clustered points whose records fill many chunks per ring batch."""
    from fasttools2d_amd.workloads import clusters
    Lx = 1.0
    Ly = Lx * My / Mx
    xi = 4.0 * Mx / 64 + 2.0
    h = Lx / Mx
    w = P * h / 2
    eta = (2 * xi * w / (0.95 * math.sqrt(math.pi * P))) ** 2
    if dist == "clusters":
        p = clusters(N, Lx, Ly, seed=N + My, ncl=6, sigma=0.02)
    else:
        p = uniform_points(N, Lx, Ly, seed=N + Mx)
    f = normal_density(N, seed=N + My)
    H1, H2 = pe.Spread(p.T, f.T, Lx, Ly, xi, w, eta, P, Mx, My)
    R1, R2 = oracle.spread(p, f, Lx, Ly, xi, w, eta, P, Mx, My)
    assert _rel(H1.T, R1) < RTOL and _rel(H2.T, R2) < RTOL
    # a fixed visiting order: the same sums bit for bit on a second call
    G1, G2 = pe.Spread(p.T, f.T, Lx, Ly, xi, w, eta, P, Mx, My)
    assert np.array_equal(G1, H1) and np.array_equal(G2, H2)


@pytest.mark.parametrize("N,Mx,My,P,Lx,Ly", [(700, 64, 64, 24, 1.0, 1.0), (900, 92, 46, 16, 2.0, 1.0),
                                              (300, 26, 26, 24, 1.0, 1.0),
                                              # a grid of the fused x passes (read from Z0)
                                              (800, 520, 260, 12, 2.0, 1.0)])
def test_gather_matches_oracle(gpu_ctx, oracle, N, Mx, My, P, Lx, Ly):
    xi = 10.0
    h = Lx / Mx
    w = P * h / 2
    eta = (2 * xi * w / (0.95 * math.sqrt(math.pi * P))) ** 2
    G = np.random.default_rng(4).standard_normal((My, Mx))
    p = uniform_points(N, Lx, Ly, seed=N)
    got = pe.Gather(G, p.T, Lx, Ly, xi, w, eta, P)
    ref = oracle.gather(np.ascontiguousarray(G.T), p, Lx, Ly, xi, w, eta, P)
    assert _rel(got, ref) < RTOL


@pytest.mark.parametrize("My,Mx", [(8, 8), (64, 64), (902, 902), (366, 732), (41, 82), (2112, 16),
                                   (30, 22), (4096, 4),
                                   # lines of 2049..2560 points run composite radices
                                   # (2160 = 16 15 9, 2560 = 16 16 10, 2250 = 15 15 10,
                                   # 2304 = 16 16 9, 2400 = 16 15 10, 2106 = 6 9 ... 13)
                                   (2160, 24), (24, 2560), (2250, 6), (2304, 10), (2400, 12),
                                   (2106, 4),
                                   # longer lines (512 threads, radices up to 8)
                                   (3000, 6), (8, 4096), (3072, 2), (2592, 8),
                                   # lines above 4096 points (1024 threads, one line per
                                   # CU in up to 160 KiB of LDS; 10240 is the longest)
                                   (4098, 2), (6, 6000), (10240, 2), (2, 10240), (8192, 4),
                                   (4, 10000), (5120, 6), (7, 4374)])
def test_fft2_matches_numpy(gpu_ctx, My, Mx):
    rng = np.random.default_rng(My + Mx)
    Z = rng.standard_normal((My, Mx)) + 1j * rng.standard_normal((My, Mx))
    F = pe.fft2(Z)
    assert _rel(F, np.fft.fft2(Z)) < 1e-14
    assert _rel(pe.fft2(F, inverse=True), Z) < 1e-14


# ---------------------------------------------------------------- a3, a5, a6 (mex real)
REAL_CASES = [
    # Ns, Nt, nsx, nsy, Lx, Ly, xi, coincident targets
    (1000, 1000, 4, 4, 1.0, 1.0, 12.0, True),  # the N = 1000 unit-box configuration
    (3000, 2000, 22, 11, 2.0, 1.0, 54.8, False),
    (500, 300, 1, 1, 1.0, 1.0, 3.0, True),
    (800, 800, 2, 2, 1.0, 1.0, 6.0, True),
    (0, 50, 3, 3, 1.0, 1.0, 9.0, False),
    (50, 0, 3, 3, 1.0, 1.0, 9.0, False),
]


@pytest.mark.parametrize("Ns,Nt,nsx,nsy,Lx,Ly,xi,coinc", REAL_CASES)
def test_slp_real_matches_oracle(gpu_ctx, oracle, Ns, Nt, nsx, nsy, Lx, Ly, xi, coinc):
    ps = uniform_points(Ns, Lx, Ly, seed=Ns + 1)
    pt = ps[:Nt].copy() if coinc else uniform_points(Nt, Lx, Ly, seed=Nt + 2)
    f = normal_density(Ns, seed=3)
    got = pe.mex_stokes_slp_real(ps.T, pt.T, f.T, xi, nsx, nsy, Lx, Ly)
    ref = oracle.slp_real(ps, pt, f, xi, nsx, nsy, Lx, Ly)
    if Nt and Ns:
        assert _rel(got.T, ref) < RTOL
    else:
        assert np.all(got == 0) and got.shape == (2, Nt)


@pytest.mark.parametrize("Ns,Nt,nsx,nsy,Lx,Ly,xi,coinc", REAL_CASES)
def test_dlp_real_matches_oracle(gpu_ctx, oracle, Ns, Nt, nsx, nsy, Lx, Ly, xi, coinc):
    ps = uniform_points(Ns, Lx, Ly, seed=Ns + 1)
    pt = ps[:Nt].copy() if coinc else uniform_points(Nt, Lx, Ly, seed=Nt + 2)
    f = normal_density(Ns, seed=3)
    th = np.random.default_rng(4).random(Ns) * 2 * np.pi
    n = np.stack([np.cos(th), np.sin(th)], 1)
    got = pe.mex_stokes_dlp_real(ps.T, pt.T, f.T, n.T, xi, nsx, nsy, Lx, Ly)
    ref = oracle.dlp_real(ps, pt, f, n, xi, nsx, nsy, Lx, Ly)
    if Nt and Ns:
        assert _rel(got.T, ref) < RTOL
    else:
        assert np.all(got == 0)


# ---------------------------------------------------------------- a8-a12 (mex kspace)
KS_CASES = [
    # Ns, Nt, Mx, My, Lx, Ly, xi, P
    (1000, 1000, 64, 64, 1.0, 1.0, 12.0, 24),  # the N = 1000 unit-box configuration
    (2000, 1500, 120, 60, 2.0, 1.0, 20.0, 16),
    (300, 200, 30, 30, 1.0, 1.0, 6.0, 24),
    (5000, 4000, 256, 256, 1.0, 1.0, 40.0, 12),
    # x lines of 512-2048 points: the fused x passes and filter, reading their rows
    # across the columns (My/2 = 130 row pairs: two jobs past the XCD groups of 16)
    (3000, 2000, 520, 260, 2.0, 1.0, 60.0, 12),
]


def _eta_w(xi, P, Lx, Mx):
    w = P * Lx / Mx / 2
    return (2 * xi * w / (0.95 * math.sqrt(math.pi * P))) ** 2, w


@pytest.mark.parametrize("Ns,Nt,Mx,My,Lx,Ly,xi,P", KS_CASES)
def test_slp_kspace_matches_oracle(gpu_ctx, oracle, Ns, Nt, Mx, My, Lx, Ly, xi, P):
    eta, w = _eta_w(xi, P, Lx, Mx)
    ps = uniform_points(Ns, Lx, Ly, seed=Ns)
    pt = uniform_points(Nt, Lx, Ly, seed=Nt + 7)
    f = normal_density(Ns, seed=9)
    got = pe.mex_stokes_slp_kspace(ps.T, pt.T, xi, eta, f.T, Mx, My, Lx, Ly, w, P)
    ref = oracle.slp_kspace(ps, pt, xi, eta, f, Mx, My, Lx, Ly, w, P)
    assert _rel(got.T, ref) < RTOL


@pytest.mark.parametrize("Lx,Ly", [(4.0, 1.0), (1.0, 4.0)])
def test_kspace_long_lines_match_oracle(gpu_ctx, oracle, Lx, Ly):
    """A 4 x 1 box with a 5120 x 1280 grid (and its transpose): the long side's
    lines run the longest-line class through spreading, both transforms, the
    filter and gathering, for the Stokeslet and the stresslet."""
    Mx, My = (5120, 1280) if Lx > Ly else (1280, 5120)
    xi, P, Ns, Nt = 300.0, 16, 3000, 1500
    eta, w = _eta_w(xi, P, Lx, Mx)
    ps = uniform_points(Ns, Lx, Ly, seed=41)
    pt = uniform_points(Nt, Lx, Ly, seed=42)
    f = normal_density(Ns, seed=43)
    got = pe.mex_stokes_slp_kspace(ps.T, pt.T, xi, eta, f.T, Mx, My, Lx, Ly, w, P)
    ref = oracle.slp_kspace(ps, pt, xi, eta, f, Mx, My, Lx, Ly, w, P)
    assert _rel(got.T, ref) < RTOL
    n = unit_normals(Ns, seed=44)
    got = pe.mex_stokes_dlp_kspace(ps.T, pt.T, xi, eta, f.T, n.T, Mx, My, Lx, Ly, w, P)
    ref = oracle.dlp_kspace(ps, pt, xi, eta, f, n, Mx, My, Lx, Ly, w, P)
    assert _rel(got.T, ref) < RTOL


def test_grid_beyond_longest_line_is_refused(gpu_ctx):
    ps = uniform_points(10, 1.0, 1.0, seed=1).T
    f = normal_density(10, seed=2).T
    with pytest.raises(ft.Se2pError) as e:
        pe.mex_stokes_slp_kspace(ps, ps, 10.0, 0.5, f, 10242, 10242, 1.0, 1.0, 0.1, 24)
    assert e.value.code == _lib.SE2P_ERR_UNSUPPORTED


@pytest.mark.parametrize("Ns,Nt,Mx,My,Lx,Ly,xi,P", KS_CASES)
def test_dlp_kspace_matches_oracle(gpu_ctx, oracle, Ns, Nt, Mx, My, Lx, Ly, xi, P):
    eta, w = _eta_w(xi, P, Lx, Mx)
    ps = uniform_points(Ns, Lx, Ly, seed=Ns)
    pt = uniform_points(Nt, Lx, Ly, seed=Nt + 7)
    f = normal_density(Ns, seed=9)
    th = np.random.default_rng(10).random(Ns) * 2 * np.pi
    n = np.stack([np.cos(th), np.sin(th)], 1)
    got = pe.mex_stokes_dlp_kspace(ps.T, pt.T, xi, eta, f.T, n.T, Mx, My, Lx, Ly, w, P)
    ref = oracle.dlp_kspace(ps, pt, xi, eta, f, n, Mx, My, Lx, Ly, w, P)
    assert _rel(got.T, ref) < RTOL


# ---------------------------------------------------------------- drivers (a13, a14)
@pytest.mark.parametrize("N,Lx,Ly,tol", [(1000, 1.0, 1.0, 1e-16), (4000, 1.0, 2.0, 1e-10),
                                         (20, 1.0, 1.0, 1e-16)])
def test_slp_driver_matches_oracle(gpu_ctx, oracle, N, Lx, Ly, tol):
    rng = np.random.default_rng(N)
    ps = rng.random((N, 2)) * [Lx, Ly]  # unwrapped on purpose: the driver wraps
    pt = rng.random((N // 2 + 1, 2)) * [Lx, Ly]
    f = 10 * rng.random((N, 2))
    u1, u2, ur, uk, xi = ft.StokesSLP_ewald_2p(ps[:, 0], ps[:, 1], pt[:, 0], pt[:, 1], f[:, 0],
                                               f[:, 1], Lx, Ly, tol=tol)
    wps, wpt = oracle.wrap(ps, Lx, Ly), oracle.wrap(pt, Lx, Ly)
    npts = len(ps) + len(pt)
    prm = oracle.params("slp", npts, oracle.density_q(f), Lx, Ly, 24, oracle.default_Nb(npts), tol)
    assert xi == prm["xi"]
    rr = oracle.slp_real(wps, wpt, f, prm["xi"], prm["nside_x"], prm["nside_y"], Lx, Ly)
    rk = oracle.slp_kspace(wps, wpt, prm["xi"], prm["eta"], f, prm["Mx"], prm["My"], Lx, Ly,
                           prm["w"], 24)
    assert _rel(ur.T, rr) < RTOL and _rel(uk.T, rk) < RTOL
    assert _rel(np.stack([u1, u2], 1), rr + rk) < RTOL


def test_dlp_driver_circles_matches_oracle(gpu_ctx, oracle):
    """Stresslet workload at test size: sources on circles, separate targets."""
    Lx, Ly = 2.0, 1.0
    ps, n, f = circles(6, 300, Lx, Ly, seed=17)
    pt = uniform_points(1500, Lx, Ly, seed=18)
    u1, u2, ur, uk, xi = ft.StokesDLP_ewald_2p(ps[:, 0], ps[:, 1], pt[:, 0], pt[:, 1], n[:, 0],
                                               n[:, 1], f[:, 0], f[:, 1], Lx, Ly, tol=1e-10)
    npts = len(ps) + len(pt)
    prm = oracle.params("dlp", npts, oracle.density_q(f), Lx, Ly, 24, oracle.default_Nb(npts),
                        1e-10)
    rr = oracle.dlp_real(ps, pt, f, n, prm["xi"], prm["nside_x"], prm["nside_y"], Lx, Ly)
    rk = oracle.dlp_kspace(ps, pt, prm["xi"], prm["eta"], f, n, prm["Mx"], prm["My"], Lx, Ly,
                           prm["w"], 24) + oracle.dlp_zero_mode(ps, f, n, Lx, Ly)
    assert _rel(ur.T, rr) < RTOL and _rel(uk.T, rk) < RTOL
    assert _rel(np.stack([u1, u2], 1), rr + rk) < RTOL


def test_stresslet_identity_on_gpu(gpu_ctx):
    """tests/stresslet_identity_test.m, evaluated by the GPU driver."""
    r, L, Ns, Nt = 1.0, 4.0, 256, 16
    th = 2 * np.pi * np.arange(1, Ns + 1) / Ns
    tt = 2 * np.pi * np.arange(1, Nt + 1) / Nt
    xt = np.concatenate([0.5 * np.cos(tt), 1.5 * np.cos(tt)])
    yt = np.concatenate([0.5 * np.sin(tt), 1.5 * np.sin(tt)])
    for comp in (0, 1):
        f = np.zeros((Ns, 2))
        f[:, comp] = 2 * np.pi / Ns
        u1, u2, _, _, _ = ft.StokesDLP_ewald_2p(r * np.cos(th), r * np.sin(th), xt, yt,
                                                np.cos(th), np.sin(th), f[:, 0], f[:, 1], L, L)
        inside = u1[:Nt] + 1j * u2[:Nt]
        outside = u1[Nt:] + 1j * u2[Nt:]
        e = 1.0 if comp == 0 else 1j
        assert np.abs(inside - e).max() < 1e-10
        assert np.abs(outside).max() < 1e-10


# ---------------------------------------------------------------- errors
def test_precondition_errors(gpu_ctx):
    ps = uniform_points(10, 1.0, 1.0, seed=1).T
    f = normal_density(10, seed=2).T
    with pytest.raises(ft.Se2pError) as e:
        pe.mex_stokes_slp_kspace(ps, ps, 10.0, 0.5, f, 63, 64, 1.0, 1.0, 0.1, 24)
    assert e.value.code == -1
    with pytest.raises(ft.Se2pError):
        pe.mex_stokes_slp_kspace(ps, ps, 10.0, 0.5, f, 64, 32, 1.0, 1.0, 0.1, 24)  # h differs
    with pytest.raises(ft.Se2pError):
        pe.mex_stokes_slp_kspace(ps, ps, 10.0, 0.5, f, 64, 64, 1.0, 1.0, 0.1, 23)  # odd P
    with pytest.raises(ValueError):
        pe.mex_stokes_slp_kspace(ps, ps, 10.0, 0.5, f[:, :5], 64, 64, 1.0, 1.0, 0.1, 24)
    with pytest.raises(ft.Se2pError):
        pe.mex_stokes_slp_real(ps, ps, f, -1.0, 3, 3, 1.0, 1.0)
