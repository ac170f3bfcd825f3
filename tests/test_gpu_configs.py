"""The benchmark configurations at their stated sizes, on the GPU, scored
against the CPU oracle (tests/inputs.oracle_se: the oracle's spectral Ewald
sum at tolerance 1e-13, all sources).

C1  Stokeslet, N = 1000, unit box, xi 12, M 64, P 24, rc 0.25: parity with the
    oracle at those parameters, and its true error against direct Ewald sums.
C2  stresslet, 20 circles x 2000 nodes, 40 000 separate targets, box 2 x 1,
    the driver at tol 1e-10: parity of its parts with the oracle, and its error.
C3  Stokeslet + stresslet, N = 4e6 points from 64 Gaussian clusters, FP64 at
    1e-10 with parameters from the autotuner (grid at most 2048 per side),
    scored on 1e4 targets.
C4  512 independent Stokeslet systems in one batch, 20 graph-replayed matvecs
    with fresh densities; sampled systems scored against the oracle.
"""
import numpy as np
import pytest

import fasttools2d_amd as ft
from fasttools2d_amd.batch import Batch
from fasttools2d_amd.plan import Plan
from fasttools2d_amd.tune import autotune, make_params, rule_params
from tests.inputs import (circles, clusters, normal_density, oracle_se, rel, serving_systems,
                          uniform_points, unit_normals)

pytestmark = pytest.mark.gpu


def _apply(ctx, kind, ps, pt, Lx, Ly, prm, f, nrm=None):
    pl = Plan(ctx, kind, ps, pt, Lx, Ly, prm, nrm=nrm)
    try:
        return pl.apply(f), pl.params_in_use()
    finally:
        pl.close()


# ---------------------------------------------------------------- C1
def test_c1_parity_and_true_error(gpu_ctx, oracle):
    N = 1000
    p = uniform_points(N, 1.0, 1.0, seed=101)
    f = normal_density(N, seed=102)
    prm = make_params(12.0, 4, 4, 64, 64, 24, 1.0)
    assert abs(prm["w"] - 0.1875) < 1e-15 and abs(prm["eta"] - 0.29759) < 1e-5
    u, _ = _apply(gpu_ctx, "slp", p, p, 1.0, 1.0, prm, f)
    ref = (oracle.slp_real(p, p, f, 12.0, 4, 4, 1.0, 1.0) +
           oracle.slp_kspace(p, p, 12.0, prm["eta"], f, 64, 64, 1.0, 1.0, prm["w"], 24))
    assert rel(u, ref) < 1e-13
    # truth: direct Ewald sums (the reference's stokes_slp_{real,kspace}_ds.m) at
    # two splits, whose agreement bounds their own error
    truth = [oracle.slp_real_ds(p, p, f, 1.0, 1.0, xi) +
             oracle.slp_kspace_ds(p, p, f, 1.0, 1.0, xi, kinf)
             for xi, kinf in ((8.0, 20), (9.0, 22))]
    assert rel(truth[0], truth[1]) < 1e-12
    err = rel(u, truth[0])
    print(f"C1 true rel-L2 error {err:.3e}, max abs {np.abs(u - truth[0]).max():.3e}")
    # not the "about 1e-10" the configuration is quoted at: xi 12 with rc 0.25
    # truncates the real space at e^{-9}; the oracle at these parameters is
    # 2.3e-5 from the truth on this input (the recipe's estimate, 1.4e-3, is a
    # bound)
    assert 1e-6 < err < 1e-3


# ---------------------------------------------------------------- C2
def test_c2_stresslet_circles_full_size(gpu_ctx, oracle):
    Lx, Ly, tol = 2.0, 1.0, 1e-10
    ps, n, f = circles(20, 2000, Lx, Ly, seed=17)
    pt = uniform_points(40000, Lx, Ly, seed=18)
    assert len(ps) == 40000
    u1, u2, ur, uk, xi = ft.StokesDLP_ewald_2p(ps[:, 0], ps[:, 1], pt[:, 0], pt[:, 1], n[:, 0],
                                               n[:, 1], f[:, 0], f[:, 1], Lx, Ly, tol=tol)
    npts = len(ps) + len(pt)
    prm = oracle.params("dlp", npts, oracle.density_q(f), Lx, Ly, 24, oracle.default_Nb(npts),
                        tol)
    assert xi == prm["xi"] and (prm["nside_x"], prm["Mx"], prm["My"]) == (22, 732, 366)
    rr = oracle.dlp_real(ps, pt, f, n, prm["xi"], prm["nside_x"], prm["nside_y"], Lx, Ly)
    rk = (oracle.dlp_kspace(ps, pt, prm["xi"], prm["eta"], f, n, prm["Mx"], prm["My"], Lx, Ly,
                            prm["w"], 24) + oracle.dlp_zero_mode(ps, f, n, Lx, Ly))
    assert rel(ur.T, rr) < 1e-13 and rel(uk.T, rk) < 1e-13
    u = np.stack([u1, u2], 1)
    truth = oracle_se(oracle, "dlp", ps, pt, f, n, Lx, Ly, tol=1e-14)
    err, emax = rel(u, truth), float(np.abs(u - truth).max())
    print(f"C2 rel-L2 error {err:.3e}, max abs {emax:.3e}")
    assert emax <= tol and err <= tol


# ---------------------------------------------------------------- C3
_C3 = {}


@pytest.fixture(scope="module")
def c3_points():
    N = 4_000_000
    p = clusters(N, 1.0, 1.0, seed=3)
    return p, normal_density(N, seed=4), normal_density(N, seed=5), unit_normals(N, seed=6)


@pytest.mark.parametrize("kind", ["slp", "dlp"])
def test_c3_clustered_fp64_autotuned(gpu_ctx, oracle, c3_points, kind):
    p, f, g, n = c3_points
    tol = 1e-10
    dens, nrm = (f, None) if kind == "slp" else (g, n)
    prm, rep = autotune(gpu_ctx, kind, p, p, dens, 1.0, 1.0, tol, nrm=nrm, reps=2,
                        final_rounds=2, max_grid=2048, nside_scales=(1.0, 1.4),
                        log=lambda m: print(m, flush=True))
    print(f"C3 {kind}: tuned {prm} ms {rep['ms']:.2f} err {rep['err']:.2e} "
          f"held-out {rep['err_heldout']}")
    assert prm["Mx"] <= 2048 and rep["err"] <= 0.8 * tol
    u, used = _apply(gpu_ctx, kind, p, p, 1.0, 1.0, prm, dens, nrm)
    sub = np.sort(np.random.default_rng(7).choice(len(p), 10_000, replace=False))
    ref = oracle_se(oracle, kind, p, p[sub], dens, nrm, 1.0, 1.0)
    err = rel(u[sub], ref)
    print(f"C3 {kind}: rel-L2 vs oracle on 1e4 targets {err:.3e}, symmetric {used['symmetric']}")
    assert err <= tol
    _C3[kind] = (u[sub], ref)


@pytest.mark.parametrize("kind", ["slp", "dlp"])
def test_c3_clustered_fp32_autotuned(gpu_ctx, oracle, c3_points, kind):
    """Configuration 3's FP32 run: tolerance 1e-6, the FP32 tier."""
    p, f, g, n = c3_points
    tol = 1e-6
    dens, nrm = (f, None) if kind == "slp" else (g, n)
    prm, rep = autotune(gpu_ctx, kind, p, p, dens, 1.0, 1.0, tol, nrm=nrm, reps=2,
                        final_rounds=2, max_grid=2048, nside_scales=(1.0, 1.4),
                        log=lambda m: print(m, flush=True))
    assert prm["precision"] == 1 and rep["err"] <= 0.8 * tol
    u, used = _apply(gpu_ctx, kind, p, p, 1.0, 1.0, prm, dens, nrm)
    assert used["precision"] == 1
    sub = np.sort(np.random.default_rng(7).choice(len(p), 10_000, replace=False))
    err = rel(u[sub], oracle_se(oracle, kind, p, p[sub], dens, nrm, 1.0, 1.0))
    print(f"C3 FP32 {kind}: {rep['ms']:.2f} ms, rel-L2 vs oracle on 1e4 targets {err:.3e}")
    assert err <= tol


def test_c3_combined(gpu_ctx):
    if set(_C3) != {"slp", "dlp"}:
        pytest.skip("needs both C3 parts from this session")
    u = _C3["slp"][0] + _C3["dlp"][0]
    ref = _C3["slp"][1] + _C3["dlp"][1]
    assert rel(u, ref) <= 1e-10


# ---------------------------------------------------------------- C4
def test_c4_serving_batch_512_systems(gpu_ctx, oracle):
    systems = serving_systems(512)
    for y in systems:
        n = len(y["psrc"])
        # FP32 or FP64 per system by its tolerance
        y["prm"] = rule_params("slp", n, n, y["Lx"], y["Ly"], y["tol"], symmetric=1,
                               precision=None)
    b = Batch(gpu_ctx, systems)
    try:
        assert sum(len(y["psrc"]) for y in systems) > 1e7
        precs = [b.params_in_use(i)["precision"] for i in range(len(systems))]
        assert precs == [int(y["tol"] >= 1e-6) for y in systems]
        assert 0 < sum(precs) < len(systems)
        rng = np.random.default_rng(77)
        keep = {}
        for k in range(20):
            fk = rng.standard_normal(b.nf)
            b.f.upload(fk)
            b.apply_device(b.f.ptr, b.u.ptr, graph=True)
            if k in (0, 19):
                keep[k] = (fk, b.u.download())
        # the replayed graph against stream launches on the last density
        b.apply_device(b.f.ptr, b.u.ptr, graph=False)
        assert rel(b.u.download(), keep[19][1]) < 1e-12
        idx = sorted(set(np.linspace(0, len(systems) - 1, 12).astype(int)) |
                     {int(np.argmin([len(y["psrc"]) for y in systems])),
                      int(np.argmax([len(y["psrc"]) for y in systems]))})
        # and at least two systems of each precision
        idx = sorted(set(idx) | set([i for i in range(len(systems)) if precs[i]][:2]) |
                     set([i for i in range(len(systems)) if not precs[i]][:2]))
        worst = 0.0
        for i in idx:
            y = systems[i]
            ps = y["psrc"]
            sub = np.arange(len(ps)) if len(ps) <= 20000 else np.sort(
                np.random.default_rng(i).choice(len(ps), 2000, replace=False))
            for k, (fk, uk) in keep.items():
                fi = fk[b.f_off[i]:b.f_off[i + 1]].reshape(-1, 2)
                ui = uk[b.u_off[i]:b.u_off[i + 1]].reshape(-1, 2)
                e = rel(ui[sub], oracle_se(oracle, "slp", ps, ps[sub], fi, None, y["Lx"], y["Ly"]))
                worst = max(worst, e / y["tol"])
        print(f"C4: worst sampled error / tol {worst:.3f} over systems {idx}")
        assert worst <= 1.0
    finally:
        b.close()


# ---------------------------------------------------------------- driver Q
def test_driver_density_q_column_order(gpu_ctx, oracle):
    """Q = sum(sum(f.^2)) + 1 adds the columns f1(i)^2 + f2(i)^2 in order; for
    this density the row-first order rounds differently and moves xi."""
    rng = np.random.default_rng(16)
    f = rng.standard_normal((1000, 2)) * np.exp(rng.uniform(-20, 20, (1000, 2)))
    p = uniform_points(1000, 1.0, 1.0, seed=161)
    _, _, _, _, xi = ft.StokesSLP_ewald_2p(p[:, 0], p[:, 1], p[:, 0], p[:, 1], f[:, 0], f[:, 1],
                                           1.0, 1.0, tol=1e-10)
    q_col = oracle.density_q(f)
    q_row = (sum(f[:, 0] ** 2) + sum(f[:, 1] ** 2)) + 1  # each row first
    assert q_col == ft.density_q(f[:, 0], f[:, 1])
    prm = oracle.params("slp", 2000, q_col, 1.0, 1.0, 24, oracle.default_Nb(2000), 1e-10)
    assert xi == prm["xi"]
    assert prm["xi"] != oracle.params("slp", 2000, q_row, 1.0, 1.0, 24,
                                      oracle.default_Nb(2000), 1e-10)["xi"]
