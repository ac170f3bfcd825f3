"""Pins of the CPU oracle (oracle/se2p_oracle.c) against golden vectors, library
routines, brute force and the invariants the reference's own tests check
(tests/direct_sums, tests/consistency_checks, tests/stresslet_identity_test.m).

CPU only; sized to run in well under a minute.
"""
import json
import math
import os
from fractions import Fraction

import numpy as np
import pytest
from scipy.special import exp1

from tests.inputs import normal_density, uniform_points

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b))


# ---------------------------------------------------------------- a4: E1
def test_e1_matches_mpmath_table(oracle):
    rows = json.load(open(os.path.join(GOLD, "e1_mpmath.json")))["rows"]
    worst = 0.0
    for xs, ref in rows:
        x = float(xs)
        r = float(ref)
        if r == 0.0:  # underflows double
            continue
        worst = max(worst, abs(oracle.e1(x) - r) / r)
    assert worst < 3e-16, worst


def test_e1_derivative_identity(oracle):
    # E1'(x) = -e^{-x}/x, checked by central differences on a smooth range
    for x in (0.3, 0.9, 1.0, 1.1, 3.0, 15.9, 16.1, 40.0):
        h = 1e-4 * min(x, 1.0 / x)
        d = (oracle.e1(x + h) - oracle.e1(x - h)) / (2 * h)
        assert abs(d + math.exp(-x) / x) < 1e-6 * math.exp(-x) / x


# ---------------------------------------------------------------- a1: Assign
@pytest.mark.parametrize("nsx,nsy,Lx,Ly", [(4, 4, 1.0, 1.0), (22, 11, 2.0, 1.0), (1, 1, 1.0, 1.0),
                                           (7, 3, 1.0, 0.5)])
def test_assign_is_stable_counting_sort(oracle, nsx, nsy, Lx, Ly):
    p = uniform_points(3000, Lx, Ly, seed=nsx * 100 + nsy)
    # exact cell edges and the box ends are included on purpose
    edges = np.array([[-Lx / 2, -Ly / 2], [-Lx / 2 + Lx / nsx, 0.0], [0.0, -Ly / 2 + Ly / nsy],
                      [np.nextafter(Lx / 2, 0), np.nextafter(Ly / 2, 0)]])
    p = np.concatenate([p, edges, p[:50]])  # duplicates keep index order inside a cell
    po, bo, nb = oracle.assign(p, Lx, Ly, nsx, nsy)
    bx = np.clip(np.floor(nsx * (p[:, 0] / Lx + 0.5)).astype(int), 0, nsx - 1)
    by = np.clip(np.floor(nsy * (p[:, 1] / Ly + 0.5)).astype(int), 0, nsy - 1)
    key = by * nsx + bx
    np.testing.assert_array_equal(po, np.argsort(key, kind="stable"))
    np.testing.assert_array_equal(nb, np.bincount(key, minlength=nsx * nsy))
    np.testing.assert_array_equal(bo, np.concatenate([[0], np.cumsum(nb)]))
    # geometry: a point of cell (bx, by) lies inside that cell
    cx = (po_cells := key[po]) % nsx
    cy = po_cells // nsx
    x, y = p[po, 0], p[po, 1]
    assert np.all(x >= -Lx / 2 + cx * Lx / nsx - 1e-12)
    assert np.all(x <= -Lx / 2 + (cx + 1) * Lx / nsx + 1e-12)
    assert np.all(y >= -Ly / 2 + cy * Ly / nsy - 1e-12)
    assert np.all(y <= -Ly / 2 + (cy + 1) * Ly / nsy + 1e-12)


# ---------------------------------------------------------------- a7: FindClosestNode
def test_find_closest_node_geometry(oracle):
    L, M, P = 1.0, 64, 24
    h = L / M
    rng = np.random.default_rng(3)
    for x, y in rng.random((500, 2)) - 0.5:
        mx, my, px, py = oracle.find_closest_node(x, y, L, L, h, P)
        assert 0 <= px < h and 0 <= py < h
        # the node P/2 after the start node is the grid node at or below the point
        assert abs((-L / 2 + (mx + P // 2) * h) + px - x) < 1e-13
        assert abs((-L / 2 + (my + P // 2) * h) + py - y) < 1e-13


def test_find_closest_node_snaps_onto_nodes(oracle):
    L, M, P = 1.0, 64, 24
    h = L / M
    for k in (0, 1, 17, 63):
        on = -L / 2 + k * h
        for x in (on, on - 4e-14, on + 4e-14):
            mx, _, px, _ = oracle.find_closest_node(x, 0.0, L, L, h, P)
            assert px == 0.0
            assert mx + P // 2 == k


# ---------------------------------------------------------------- a8/a12: Spread, Gather
def _gauss_grid(p, f, L, M, P, xi, eta, oracle):
    h = L / M
    H = np.zeros((2, M, M))
    for (x, y), fk in zip(p, f):
        mx, my, _, _ = oracle.find_closest_node(x, y, L, L, h, P)
        for i in range(P + 1):
            dx = -L / 2 + (mx + i) * h - x
            for j in range(P + 1):
                dy = -L / 2 + (my + j) * h - y
                g = math.exp(-2 * xi * xi * (dx * dx + dy * dy) / eta)
                H[:, (mx + i) % M, (my + j) % M] += g * fk
    return H


def test_spread_equals_direct_gaussians(oracle):
    L, M, P, xi = 1.0, 32, 8, 10.0
    h = L / M
    w = P * h / 2
    eta = (2 * xi * w / (0.95 * math.sqrt(math.pi * P))) ** 2
    p = uniform_points(25, L, L, seed=11)
    p[0] = [-L / 2 + 3 * h, -L / 2]  # on a node, at the box corner
    f = normal_density(25, seed=12)
    H1, H2 = oracle.spread(p, f, L, L, xi, w, eta, P, M, M)
    ref = _gauss_grid(p, f, L, M, P, xi, eta, oracle)
    assert _rel(H1, ref[0]) < 1e-13
    assert _rel(H2, ref[1]) < 1e-13


def test_spread_mass_and_gather_adjoint(oracle):
    L, M, P, xi = 1.0, 64, 24, 12.0
    h = L / M
    w = P * h / 2
    eta = (2 * xi * w / (0.95 * math.sqrt(math.pi * P))) ** 2
    p = uniform_points(200, L, L, seed=5)
    f = normal_density(200, seed=6)
    H1, H2 = oracle.spread(p, f, L, L, xi, w, eta, P, M, M)
    # a resolved Gaussian integrates to pi eta / (2 xi^2)
    mass = math.pi * eta / (2 * xi * xi)
    assert abs(h * h * H1.sum() - mass * f[:, 0].sum()) < 1e-12 * mass * np.abs(f).sum()
    assert abs(h * h * H2.sum() - mass * f[:, 1].sum()) < 1e-12 * mass * np.abs(f).sum()
    # <Spread f, G> = <f, Gather G> / scale
    G = np.random.default_rng(9).standard_normal((M, M))
    s = (4 * xi * xi / eta) ** 2 * h * h / math.pi / (4 * math.pi)
    g = oracle.gather(G, p, L, L, xi, w, eta, P)
    assert abs((H1 * G).sum() - (f[:, 0] * g).sum() / s) < 1e-12 * np.abs(H1 * G).sum()


# ---------------------------------------------------------------- a9: FFT
@pytest.mark.parametrize("shape", [(8, 8), (6, 10), (30, 22), (82, 22), (902, 4), (41, 2)])
def test_fft2_matches_numpy(oracle, shape):
    rng = np.random.default_rng(sum(shape))
    Z = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    F = oracle.fft2(Z)
    assert _rel(F, np.fft.fft2(Z)) < 1e-14
    assert _rel(oracle.fft2(F, inverse=True), Z) < 1e-14


# ---------------------------------------------------------------- a3/a5/a6: real space
def _brute_real(kind, ps, pt, f, n, xi, nsx, nsy, Lx, Ly):
    """Numpy restatement of the cell-list pair set with scipy's E1."""
    out = np.zeros((len(pt), 2))
    cut2 = Lx * Ly / nsx / nsy

    def cell(p):
        bx = np.clip(np.floor(nsx * (p[:, 0] / Lx + 0.5)).astype(int), 0, nsx - 1)
        by = np.clip(np.floor(nsy * (p[:, 1] / Ly + 0.5)).astype(int), 0, nsy - 1)
        return bx, by

    sx, sy = cell(ps)
    tx, ty = cell(pt)
    for t in range(len(pt)):
        ddx = (sx - tx[t] + 1) % nsx - 1
        ddy = (sy - ty[t] + 1) % nsy - 1
        inside = (np.abs(ddx) <= 1) & (np.abs(ddy) <= 1)
        # image of the source cell next to the target cell
        shx = (tx[t] + ddx - sx) * Lx / nsx
        shy = (ty[t] + ddy - sy) * Ly / nsy
        rx = pt[t, 0] - (ps[:, 0] + shx)
        ry = pt[t, 1] - (ps[:, 1] + shy)
        r2 = rx * rx + ry * ry
        same = (ddx == 0) & (ddy == 0)
        take = inside & (same | (r2 < cut2))
        if kind == "slp":
            selfp = take & same & (r2 < 1e-15)
            pair = take & ~selfp
            x = xi * xi * r2[pair]
            E = exp1(x)
            e = np.exp(-x)
            fx, fy = f[pair, 0], f[pair, 1]
            rdf = rx[pair] * fx + ry[pair] * fy
            out[t, 0] = (0.5 * E * fx + e * (rdf * rx[pair] / r2[pair] - fx)).sum()
            out[t, 1] = (0.5 * E * fy + e * (rdf * ry[pair] / r2[pair] - fy)).sum()
            c = -(np.euler_gamma / 2 + math.log(xi) + 1)
            out[t] += c * f[selfp].sum(axis=0)
        else:
            pair = take & (r2 != 0)
            a, b, q = rx[pair], ry[pair], r2[pair]
            fx, fy, nx, ny = f[pair, 0], f[pair, 1], n[pair, 0], n[pair, 1]
            e = np.exp(-xi * xi * q)
            rf, rn, fn = a * fx + b * fy, a * nx + b * ny, fx * nx + fy * ny
            k = -4 * rf * rn / q ** 2 * (1 + xi * xi * q)
            out[t, 0] = (e * (k * a + 2 * xi * xi * (fx * rn + nx * rf + fn * a))).sum()
            out[t, 1] = (e * (k * b + 2 * xi * xi * (fy * rn + ny * rf + fn * b))).sum()
    return out / (4 * math.pi)


@pytest.mark.parametrize("nsx,nsy,Lx,Ly", [(4, 4, 1.0, 1.0), (6, 3, 2.0, 1.0)])
def test_real_space_matches_numpy_restatement(oracle, nsx, nsy, Lx, Ly):
    ps = uniform_points(400, Lx, Ly, seed=21)
    pt = np.concatenate([uniform_points(200, Lx, Ly, seed=22), ps[:30]])  # coincident points
    f = normal_density(400, seed=23)
    th = np.random.default_rng(24).random(400) * 2 * np.pi
    n = np.stack([np.cos(th), np.sin(th)], 1)
    xi = 3.2 * nsx / Lx
    u = oracle.slp_real(ps, pt, f, xi, nsx, nsy, Lx, Ly)
    assert _rel(u, _brute_real("slp", ps, pt, f, None, xi, nsx, nsy, Lx, Ly)) < 1e-13
    u = oracle.dlp_real(ps, pt, f, n, xi, nsx, nsy, Lx, Ly)
    assert _rel(u, _brute_real("dlp", ps, pt, f, n, xi, nsx, nsy, Lx, Ly)) < 1e-13


# ---------------------------------------------------------------- whole sums
def _ewald(oracle, kind, ps, pt, f, n, Lx, Ly, tol, Nb=None, P=24):
    npts = len(ps) + len(pt)
    Q = (f ** 2).sum() + 1
    prm = oracle.params(kind, npts, Q, Lx, Ly, P, Nb if Nb else oracle.default_Nb(npts, P), tol)
    if kind == "slp":
        ur = oracle.slp_real(ps, pt, f, prm["xi"], prm["nside_x"], prm["nside_y"], Lx, Ly)
        uk = oracle.slp_kspace(ps, pt, prm["xi"], prm["eta"], f, prm["Mx"], prm["My"], Lx, Ly,
                               prm["w"], P)
    else:
        ur = oracle.dlp_real(ps, pt, f, n, prm["xi"], prm["nside_x"], prm["nside_y"], Lx, Ly)
        uk = oracle.dlp_kspace(ps, pt, prm["xi"], prm["eta"], f, n, prm["Mx"], prm["My"], Lx, Ly,
                               prm["w"], P) + oracle.dlp_zero_mode(ps, f, n, Lx, Ly)
    return ur, uk, prm


def test_direct_sums_like_reference(oracle):
    """tests/direct_sums/direct_sums_velocity_test.m: 20 sources, 20 targets, tol 1e-16."""
    Lx = Ly = 1.0
    rng = np.random.default_rng(31)
    f = 10 * rng.random((20, 2))
    n1 = rng.random(20)
    n = np.stack([n1, np.sqrt(1 - n1 ** 2)], 1)
    ps = oracle.wrap(rng.random((20, 2)), Lx, Ly)
    pt = oracle.wrap(rng.random((20, 2)), Lx, Ly)
    ur, uk, prm = _ewald(oracle, "slp", ps, pt, f, None, Lx, Ly, 1e-16)
    assert np.abs(ur - oracle.slp_real_ds(ps, pt, f, Lx, Ly, prm["xi"])).max() < 1e-13
    assert np.abs(uk - oracle.slp_kspace_ds(ps, pt, f, Lx, Ly, prm["xi"], 70)).max() < 1e-11
    ur, uk, prm = _ewald(oracle, "dlp", ps, pt, f, n, Lx, Ly, 1e-16)
    assert np.abs(ur - oracle.dlp_real_ds(ps, pt, f, n, Lx, Ly, prm["xi"])).max() < 1e-12
    assert np.abs(uk - oracle.dlp_kspace_ds(ps, pt, f, n, Lx, Ly, prm["xi"], 70)).max() < 1e-10


@pytest.mark.parametrize("kind", ["slp", "dlp"])
def test_total_independent_of_ewald_split(oracle, kind):
    """tests/consistency_checks/consistency_test_velocity.m:54-76: Nb in {3, 9, 27}."""
    Lx, Ly = 1.0, 2.0
    ps = uniform_points(300, Lx, Ly, seed=41)
    pt = uniform_points(150, Lx, Ly, seed=42)
    f = 10 * np.random.default_rng(43).random((300, 2))
    n1 = np.random.default_rng(44).random(300)
    n = np.stack([n1, np.sqrt(1 - n1 ** 2)], 1)
    us = []
    for Nb in (3, 9, 27):
        ur, uk, _ = _ewald(oracle, kind, ps, pt, f, n, Lx, Ly, 1e-14, Nb=Nb)
        us.append(ur + uk)
    scale = np.abs(us[0]).max()
    for u in us[1:]:
        assert np.abs(u - us[0]).max() < 1e-11 * scale


@pytest.mark.parametrize("kind", ["slp", "dlp"])
def test_periodic_replication(oracle, kind):
    """consistency_test_velocity.m:80-94 / :158-194: Lx -> 2 Lx with copied sources."""
    Lx, Ly = 1.0, 2.0
    ps = uniform_points(200, Lx, Ly, seed=51)
    pt = uniform_points(100, Lx, Ly, seed=52)
    f = normal_density(200, seed=53)
    th = np.random.default_rng(54).random(200) * 2 * np.pi
    n = np.stack([np.cos(th), np.sin(th)], 1)
    ur, uk, _ = _ewald(oracle, kind, ps, pt, f, n, Lx, Ly, 1e-14)
    u1 = ur + uk
    ps2 = np.concatenate([ps, ps + [Lx, 0]])
    ps2 = oracle.wrap(ps2, 2 * Lx, Ly)
    f2, n2 = np.concatenate([f, f]), np.concatenate([n, n])
    ur, uk, _ = _ewald(oracle, kind, ps2, pt, f2, n2, 2 * Lx, Ly, 1e-14)
    u2 = ur + uk
    if kind == "dlp":
        u1 = u1 - oracle.dlp_zero_mode(ps, f, n, Lx, Ly)
        u2 = u2 - oracle.dlp_zero_mode(ps2, f2, n2, 2 * Lx, Ly)
    assert np.abs(u1 - u2).max() < 1e-10 * np.abs(u1).max()


def test_stresslet_identity(oracle):
    """tests/stresslet_identity_test.m: DLP of a constant density on a circle is e_i
    inside and 0 outside."""
    r, L, Ns, Nt = 1.0, 4.0, 256, 16
    th = 2 * np.pi * np.arange(1, Ns + 1) / Ns
    ps = np.stack([r * np.cos(th), r * np.sin(th)], 1)
    n = np.stack([np.cos(th), np.sin(th)], 1)
    tt = 2 * np.pi * np.arange(1, Nt + 1) / Nt
    pt = np.concatenate([0.5 * np.stack([np.cos(tt), np.sin(tt)], 1),
                         1.5 * np.stack([np.cos(tt), np.sin(tt)], 1)])
    for comp in (0, 1):
        f = np.zeros((Ns, 2))
        f[:, comp] = 2 * np.pi / Ns
        ur, uk, _ = _ewald(oracle, "dlp", oracle.wrap(ps, L, L), oracle.wrap(pt, L, L), f, n, L,
                           L, 1e-16)
        u = ur + uk
        e = np.zeros(2)
        e[comp] = 1
        assert np.abs(u[:Nt] - e).max() < 1e-10
        assert np.abs(u[Nt:]).max() < 1e-10


# ---------------------------------------------------------------- a14: parameters
def test_rat_matches_continued_fractions(oracle):
    assert oracle.rat(2.0) == (2, 1)
    assert oracle.rat(0.5) == (1, 2)
    assert oracle.rat(1.0) == (1, 1)
    assert oracle.rat(1 / 3) == (1, 3)
    assert oracle.rat(math.pi) == (355, 113)
    for x in (1.25, 0.75, 2.5, 4.0, 1.6):
        N, D = oracle.rat(x)
        assert Fraction(N, D) == Fraction(x).limit_denominator(1000)


def test_params_reference_recipe_values(oracle):
    # N = 1e6 sources = targets, unit box, tol 1e-8, f ~ N(0,1): Q ~ 2e6
    p = oracle.params("slp", 2_000_000, 2e6 + 1, 1.0, 1.0, 24, oracle.default_Nb(2e6), 1e-8)
    assert (p["nside_x"], p["nside_y"], p["Mx"], p["My"]) == (63, 63, 902, 902)
    assert abs(p["xi"] - 297.88) < 0.05
    assert abs(p["w"] - 0.013304) < 1e-6 and abs(p["eta"] - 0.92319) < 1e-4
    # stresslet, 40k + 40k points, box 2 x 1, tol 1e-10 and 1e-6
    Q = 2 * 40000 * 1e-3 + 1
    p = oracle.params("dlp", 80000, Q, 2.0, 1.0, 24, oracle.default_Nb(80000), 1e-10)
    assert (p["nside_x"], p["nside_y"]) == (22, 11)
    assert abs(p["rc"] - 2 / 22) < 1e-15
    # N = 4e6: nside 121, tol 1e-10
    p = oracle.params("slp", 8_000_000, 8e6 + 1, 1.0, 1.0, 24, oracle.default_Nb(8e6), 1e-10)
    assert (p["nside_x"], p["Mx"]) == (121, 2112)
    assert abs(p["xi"] - 632) < 1.0


def test_density_q_sums_columns_in_order(oracle):
    """Q = sum(sum(f.^2)) + 1 for f = [f1'; f2']: MATLAB's sum of a 2 x N matrix
    adds down each column, then the outer sum adds the N column sums in order.
    A hand-written sequential loop pins it; for this density the row-first
    order differs in the last bits and moves the recipe's xi."""
    rng = np.random.default_rng(16)
    f = rng.standard_normal((1000, 2)) * np.exp(rng.uniform(-20, 20, (1000, 2)))
    q = 0.0
    for i in range(len(f)):
        col = f[i, 0] * f[i, 0] + f[i, 1] * f[i, 1]
        q = q + col
    assert oracle.density_q(f) == q + 1.0
    import fasttools2d_amd.periodic_ewald as pe
    assert pe.density_q(f[:, 0], f[:, 1]) == q + 1.0
    q1 = q2 = 0.0
    for i in range(len(f)):
        q1 += f[i, 0] * f[i, 0]
        q2 += f[i, 1] * f[i, 1]
    assert (q1 + q2) + 1 != q + 1.0
    a = oracle.params("slp", 2000, q + 1.0, 1.0, 1.0, 24, oracle.default_Nb(2000), 1e-10)
    b = oracle.params("slp", 2000, (q1 + q2) + 1, 1.0, 1.0, 24, oracle.default_Nb(2000), 1e-10)
    assert a["xi"] != b["xi"]


def test_refcpu_build_matches_the_checker(oracle):
    """The CPU baseline build (rational-fit E1, -O3 with contraction) computes
    the same sums as the checker to rounding."""
    p = uniform_points(3000, 1.0, 1.0, seed=77)
    f = normal_density(3000, seed=78)
    R = oracle.refcpu()
    a = oracle.slp_real(p, p, f, 40.0, 7, 7, 1.0, 1.0, L=R)
    b = oracle.slp_real(p, p, f, 40.0, 7, 7, 1.0, 1.0)
    assert np.linalg.norm(a - b) <= 1e-13 * np.linalg.norm(b)
    a = oracle.slp_kspace(p, p, 40.0, 0.5, f, 96, 96, 1.0, 1.0, 24 / 96 / 2, 24, L=R)
    b = oracle.slp_kspace(p, p, 40.0, 0.5, f, 96, 96, 1.0, 1.0, 24 / 96 / 2, 24)
    assert np.linalg.norm(a - b) <= 1e-13 * np.linalg.norm(b)
