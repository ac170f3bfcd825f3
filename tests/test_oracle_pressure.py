"""Pins of the oracle's SLP pressure (mex_stokes_slp_pressure_{real,kspace}.cpp,
StokesSLP_pressure_ewald_2p.m) against what fixes it independently: a numpy
restatement of the real-space pair set, the reference's direct sums
(tests/direct_sums/direct_sums_pressure_test.m with matlab_sums/
stokes_slp_pressure_{real,kspace}_ds.m), independence of the total from the
Ewald split, and periodic replication."""
import math

import numpy as np
import pytest

from tests.inputs import normal_density, uniform_points


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _brute_pressure_real(ps, pt, f, xi, nsx, nsy, Lx, Ly):
    """Own cell: every source but coincident ones; the 8 neighbour cells (periodic
    images): r^2 < Lx Ly/(nsx nsy); pair (r.f) e^{-xi^2 r^2}/r^2 / (2 pi)."""
    cut2 = Lx * Ly / nsx / nsy

    def cell(p):
        bx = np.clip(np.floor(nsx * (p[:, 0] / Lx + 0.5)).astype(int), 0, nsx - 1)
        by = np.clip(np.floor(nsy * (p[:, 1] / Ly + 0.5)).astype(int), 0, nsy - 1)
        return bx, by

    sx, sy = cell(ps)
    tx, ty = cell(pt)
    out = np.zeros(len(pt))
    for t in range(len(pt)):
        ddx = (sx - tx[t] + 1) % nsx - 1
        ddy = (sy - ty[t] + 1) % nsy - 1
        inside = (np.abs(ddx) <= 1) & (np.abs(ddy) <= 1)
        shx = (tx[t] + ddx - sx) * Lx / nsx
        shy = (ty[t] + ddy - sy) * Ly / nsy
        rx = pt[t, 0] - (ps[:, 0] + shx)
        ry = pt[t, 1] - (ps[:, 1] + shy)
        r2 = rx * rx + ry * ry
        same = (ddx == 0) & (ddy == 0)
        take = inside & np.where(same, r2 != 0, r2 < cut2)
        rf = rx[take] * f[take, 0] + ry[take] * f[take, 1]
        out[t] = (rf * np.exp(-xi * xi * r2[take]) / r2[take]).sum()
    return out / (2 * math.pi)


def _pressure(oracle, ps, pt, f, Lx, Ly, tol, Nb=None, P=24):
    npts = len(ps) + len(pt)
    Q = (f ** 2).sum() + 1
    prm = oracle.params("slp", npts, Q, Lx, Ly, P, Nb if Nb else oracle.default_Nb(npts, P), tol)
    pr = oracle.slp_pressure_real(ps, pt, f, prm["xi"], prm["nside_x"], prm["nside_y"], Lx, Ly)
    pk = oracle.slp_pressure_kspace(ps, pt, f, prm["xi"], prm["eta"], prm["Mx"], prm["My"], Lx,
                                    Ly, prm["w"], P)
    return pr, pk + oracle.slp_pressure_zero_mode(ps, f, Lx, Ly), prm


@pytest.mark.parametrize("nsx,nsy,Lx,Ly", [(4, 4, 1.0, 1.0), (6, 3, 2.0, 1.0)])
def test_pressure_real_matches_numpy_restatement(oracle, nsx, nsy, Lx, Ly):
    ps = uniform_points(400, Lx, Ly, seed=81)
    pt = np.concatenate([uniform_points(200, Lx, Ly, seed=82), ps[:30]])  # coincident points
    f = normal_density(400, seed=83)
    xi = 3.2 * nsx / Lx
    p = oracle.slp_pressure_real(ps, pt, f, xi, nsx, nsy, Lx, Ly)
    assert _rel(p, _brute_pressure_real(ps, pt, f, xi, nsx, nsy, Lx, Ly)) < 1e-13


def test_pressure_direct_sums_like_reference(oracle):
    """direct_sums_pressure_test.m: 20 sources, 20 targets, tol 1e-16; the direct
    k-space sum includes the zero mode, as stokes_slp_pressure_kspace_ds.m does."""
    Lx = Ly = 1.0
    rng = np.random.default_rng(91)
    f = 10 * rng.random((20, 2))
    ps = oracle.wrap(rng.random((20, 2)), Lx, Ly)
    pt = oracle.wrap(rng.random((20, 2)), Lx, Ly)
    pr, pk, prm = _pressure(oracle, ps, pt, f, Lx, Ly, 1e-16)
    assert np.abs(pr - oracle.slp_pressure_real_ds(ps, pt, f, Lx, Ly, prm["xi"])).max() < 1e-12
    assert np.abs(pk - oracle.slp_pressure_kspace_ds(ps, pt, f, Lx, Ly, prm["xi"], 70)).max() < 1e-10


def test_pressure_independent_of_ewald_split(oracle):
    """consistency_test_velocity.m:54-76 applied to the pressure: Nb in {3, 9, 27}."""
    Lx, Ly = 1.0, 2.0
    ps = uniform_points(300, Lx, Ly, seed=95)
    pt = uniform_points(150, Lx, Ly, seed=96)
    f = 10 * np.random.default_rng(97).random((300, 2))
    ps_ = []
    for Nb in (3, 9, 27):
        pr, pk, _ = _pressure(oracle, ps, pt, f, Lx, Ly, 1e-14, Nb=Nb)
        ps_.append(pr + pk)
    scale = np.abs(ps_[0]).max()
    for p in ps_[1:]:
        assert np.abs(p - ps_[0]).max() < 1e-11 * scale


def test_pressure_periodic_replication(oracle):
    """Doubling the box in x with a copy of every source leaves the pressure
    unchanged (consistency_test_velocity.m:80-94, for the pressure)."""
    Lx, Ly = 1.0, 1.0
    ps = uniform_points(200, Lx, Ly, seed=98)
    pt = uniform_points(100, Lx, Ly, seed=99)
    f = normal_density(200, seed=100)
    pr, pk, _ = _pressure(oracle, ps, pt, f, Lx, Ly, 1e-14)
    ps2 = np.concatenate([ps, ps + [Lx, 0]])
    ps2 = oracle.wrap(ps2 - [Lx / 2, 0], 2 * Lx, Ly)
    pt2 = oracle.wrap(pt - [Lx / 2, 0], 2 * Lx, Ly)
    f2 = np.concatenate([f, f])
    pr2, pk2, _ = _pressure(oracle, ps2, pt2, f2, 2 * Lx, Ly, 1e-14)
    p, p2 = pr + pk, pr2 + pk2
    assert np.abs(p2 - p).max() < 1e-10 * np.abs(p).max()


def test_pressure_filter_closed_form(oracle):
    """The filter on one mode: -i (k.q) e^{-(1-eta) K^2/(4 xi^2)}/K^2, k = 0 -> 0."""
    Mx, My, Lx, Ly, xi, eta = 8, 6, 2.0, 1.5, 3.0, 0.4
    rng = np.random.default_rng(7)
    H1 = rng.standard_normal((Mx, My)) + 1j * rng.standard_normal((Mx, My))
    H2 = rng.standard_normal((Mx, My)) + 1j * rng.standard_normal((Mx, My))
    out = oracle.slp_pressure_filter(H1, H2, Lx, Ly, xi, eta)
    for ix in range(Mx):
        for iy in range(My):
            k1 = 2 * np.pi / Lx * (ix if ix <= Mx // 2 else ix - Mx)
            k2 = 2 * np.pi / Ly * (iy if iy <= My // 2 else iy - My)
            K2 = k1 * k1 + k2 * k2
            want = 0 if K2 == 0 else -1j * (k1 * H1[ix, iy] + k2 * H2[ix, iy]) * \
                np.exp(-(1 - eta) * K2 / (4 * xi * xi)) / K2
            assert abs(out[ix, iy] - want) <= 1e-14 * max(1.0, abs(want))


# ------------------------------------------------------------ DLP pressure
def _brute_dlp_pressure_real(ps, pt, f, n, xi, nsx, nsy, Lx, Ly):
    """As _brute_pressure_real with the stresslet pressure pair term
    -e^{-xi^2 r^2} ((f.n - 2 xi^2 (r.n)(r.f))/r^2 - 2 (r.n)(r.f)/r^4) / (2 pi)."""
    cut2 = Lx * Ly / nsx / nsy

    def cell(p):
        bx = np.clip(np.floor(nsx * (p[:, 0] / Lx + 0.5)).astype(int), 0, nsx - 1)
        by = np.clip(np.floor(nsy * (p[:, 1] / Ly + 0.5)).astype(int), 0, nsy - 1)
        return bx, by

    sx, sy = cell(ps)
    tx, ty = cell(pt)
    out = np.zeros(len(pt))
    for t in range(len(pt)):
        ddx = (sx - tx[t] + 1) % nsx - 1
        ddy = (sy - ty[t] + 1) % nsy - 1
        inside = (np.abs(ddx) <= 1) & (np.abs(ddy) <= 1)
        shx = (tx[t] + ddx - sx) * Lx / nsx
        shy = (ty[t] + ddy - sy) * Ly / nsy
        rx = pt[t, 0] - (ps[:, 0] + shx)
        ry = pt[t, 1] - (ps[:, 1] + shy)
        r2 = rx * rx + ry * ry
        same = (ddx == 0) & (ddy == 0)
        k = inside & np.where(same, r2 != 0, r2 < cut2)
        rf = rx[k] * f[k, 0] + ry[k] * f[k, 1]
        rn = rx[k] * n[k, 0] + ry[k] * n[k, 1]
        fn = f[k, 0] * n[k, 0] + f[k, 1] * n[k, 1]
        q = r2[k]
        out[t] = (-np.exp(-xi * xi * q) * ((fn - 2 * xi * xi * rn * rf) / q - 2 * rn * rf / q ** 2)).sum()
    return out / (2 * math.pi)


def _normals(n, seed):
    th = np.random.default_rng(seed).random(n) * 2 * np.pi
    return np.stack([np.cos(th), np.sin(th)], 1)


def _dlp_pressure(oracle, ps, pt, f, n, Lx, Ly, tol, Nb=None, P=24):
    npts = len(ps) + len(pt)
    Q = (f ** 2).sum() + 1
    prm = oracle.params("dlp", npts, Q, Lx, Ly, P, Nb if Nb else oracle.default_Nb(npts, P), tol)
    pr = oracle.dlp_pressure_real(ps, pt, f, n, prm["xi"], prm["nside_x"], prm["nside_y"], Lx, Ly)
    pk = oracle.dlp_pressure_kspace(ps, pt, f, n, prm["xi"], prm["eta"], prm["Mx"], prm["My"],
                                    Lx, Ly, prm["w"], P)
    return pr, pk + oracle.dlp_pressure_zero_mode(f, n, Lx, Ly), prm


@pytest.mark.parametrize("nsx,nsy,Lx,Ly", [(4, 4, 1.0, 1.0), (6, 3, 2.0, 1.0)])
def test_dlp_pressure_real_matches_numpy_restatement(oracle, nsx, nsy, Lx, Ly):
    ps = uniform_points(400, Lx, Ly, seed=181)
    pt = np.concatenate([uniform_points(200, Lx, Ly, seed=182), ps[:30]])
    f = normal_density(400, seed=183)
    n = _normals(400, 184)
    xi = 3.2 * nsx / Lx
    p = oracle.dlp_pressure_real(ps, pt, f, n, xi, nsx, nsy, Lx, Ly)
    assert _rel(p, _brute_dlp_pressure_real(ps, pt, f, n, xi, nsx, nsy, Lx, Ly)) < 1e-13


def test_dlp_pressure_direct_sums_like_reference(oracle):
    """direct_sums_pressure_test.m (stresslet part): 20 sources, 20 targets, tol 1e-16."""
    Lx = Ly = 1.0
    rng = np.random.default_rng(191)
    f = 10 * rng.random((20, 2))
    n1 = rng.random(20)
    n = np.stack([n1, np.sqrt(1 - n1 ** 2)], 1)
    ps = oracle.wrap(rng.random((20, 2)), Lx, Ly)
    pt = oracle.wrap(rng.random((20, 2)), Lx, Ly)
    pr, pk, prm = _dlp_pressure(oracle, ps, pt, f, n, Lx, Ly, 1e-16)
    assert np.abs(pr - oracle.dlp_pressure_real_ds(ps, pt, f, n, Lx, Ly, prm["xi"])).max() < 1e-11
    assert np.abs(pk - oracle.dlp_pressure_kspace_ds(ps, pt, f, n, Lx, Ly, prm["xi"], 70)).max() < 1e-9


def test_dlp_pressure_independent_of_ewald_split(oracle):
    Lx, Ly = 1.0, 2.0
    ps = uniform_points(300, Lx, Ly, seed=195)
    pt = uniform_points(150, Lx, Ly, seed=196)
    f = 10 * np.random.default_rng(197).random((300, 2))
    n = _normals(300, 198)
    tot = []
    for Nb in (3, 9, 27):
        pr, pk, _ = _dlp_pressure(oracle, ps, pt, f, n, Lx, Ly, 1e-14, Nb=Nb)
        tot.append(pr + pk)
    scale = np.abs(tot[0]).max()
    for p in tot[1:]:
        assert np.abs(p - tot[0]).max() < 1e-11 * scale


def test_dlp_pressure_filter_closed_form(oracle):
    Mx, My, Lx, Ly, xi, eta = 8, 6, 2.0, 1.5, 3.0, 0.4
    rng = np.random.default_rng(8)
    H = [rng.standard_normal((Mx, My)) + 1j * rng.standard_normal((Mx, My)) for _ in range(4)]
    out = oracle.dlp_pressure_filter(*H, Lx, Ly, xi, eta)
    for ix in range(Mx):
        for iy in range(My):
            k1 = 2 * np.pi / Lx * (ix if ix <= Mx // 2 else ix - Mx)
            k2 = 2 * np.pi / Ly * (iy if iy <= My // 2 else iy - My)
            K2 = k1 * k1 + k2 * k2
            kqk = k1 * k1 * H[0][ix, iy] + k1 * k2 * (H[1][ix, iy] + H[2][ix, iy]) + \
                k2 * k2 * H[3][ix, iy]
            want = 0 if K2 == 0 else -kqk * np.exp(-(1 - eta) * K2 / (4 * xi * xi)) / K2
            assert abs(out[ix, iy] - want) <= 1e-14 * max(1.0, abs(want))
