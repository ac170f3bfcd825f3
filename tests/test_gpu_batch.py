"""Batches of independent systems: every system of a batch gives, bit for bit,
what its own plan gives; graph replays with fresh densities match eager
batches; small systems agree with the oracle.  The systems follow the serving
workload (box aspect 1-4, N log-uniform, tolerance log-uniform in 1e-10..1e-4)
at test sizes, mixed with stresslet systems on circles."""
import numpy as np
import pytest

import fasttools2d_amd as ft
from fasttools2d_amd.batch import Batch
from fasttools2d_amd.plan import Plan
from tests.inputs import circles, normal_density, uniform_points

pytestmark = pytest.mark.gpu

ASPECTS = (1.0, 1.5, 2.0, 2.5, 3.0, 4.0)


def _systems(nsys, seed, nmin=500, nmax=12000, with_dlp=True):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(nsys):
        Lx, Ly = float(rng.choice(ASPECTS)), 1.0
        tol = 10 ** rng.uniform(-10, -4)
        if with_dlp and i % 4 == 3:
            ps, nrm, _ = circles(3, int(rng.integers(100, 400)), Lx, Ly, seed=seed + i)
            pt = uniform_points(int(rng.integers(nmin, nmax // 4)), Lx, Ly, seed=seed + 100 + i)
            prm = ft.ewald_params("dlp", len(ps), len(pt), 2.0 * len(ps) + 1, Lx, Ly, tol=tol)
            out.append(dict(kind="dlp", psrc=ps, ptar=pt, nrm=nrm, Lx=Lx, Ly=Ly, prm=prm))
        else:
            n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
            ps = uniform_points(n, Lx, Ly, seed=seed + i)
            prm = ft.ewald_params("slp", n, n, 2.0 * n + 1, Lx, Ly, tol=tol)
            out.append(dict(kind="slp", psrc=ps, ptar=ps, nrm=None, Lx=Lx, Ly=Ly, prm=prm))
    return out


def _densities(systems, seed):
    return [normal_density(len(y["psrc"]), seed=seed + 7 * i) for i, y in enumerate(systems)]


def test_batch_matches_plans_bitwise(gpu_ctx):
    systems = _systems(12, seed=400)
    b = Batch(gpu_ctx, systems)
    fs = _densities(systems, seed=1)
    us = b.apply(fs)
    for y, f, u in zip(systems, fs, us):
        pl = Plan(gpu_ctx, y["kind"], y["psrc"], y["ptar"], y["Lx"], y["Ly"], y["prm"],
                  nrm=y["nrm"])
        np.testing.assert_array_equal(u, pl.apply(f))
        pl.close()
    b.close()


def test_batch_graph_replays_fresh_densities(gpu_ctx):
    systems = _systems(9, seed=500)
    b = Batch(gpu_ctx, systems)
    for it in range(4):  # repeated matvecs on fixed geometry, new densities each time
        fs = _densities(systems, seed=10 + it)
        eager = b.apply(fs)
        replay = b.apply(fs, graph=True)
        for u, v in zip(eager, replay):
            np.testing.assert_array_equal(u, v)
    b.close()


def test_batch_small_systems_match_oracle(gpu_ctx, oracle):
    Lx, Ly = 2.0, 1.0
    ps = uniform_points(700, Lx, Ly, seed=71)
    f = normal_density(700, seed=72)
    prm = ft.ewald_params("slp", 700, 700, 2.0 * 700 + 1, Lx, Ly, tol=1e-12)
    cps, cn, cf = circles(2, 150, 1.0, 1.0, seed=73)
    cpt = uniform_points(400, 1.0, 1.0, seed=74)
    cprm = ft.ewald_params("dlp", len(cps), 400, 2.0 * len(cps) + 1, 1.0, 1.0, tol=1e-12)
    systems = [dict(kind="slp", psrc=ps, ptar=ps, nrm=None, Lx=Lx, Ly=Ly, prm=prm),
               dict(kind="dlp", psrc=cps, ptar=cpt, nrm=cn, Lx=1.0, Ly=1.0, prm=cprm)]
    b = Batch(gpu_ctx, systems)
    u_slp, u_dlp = b.apply([f, cf])
    r = oracle.slp_real(ps, ps, f, prm["xi"], prm["nside_x"], prm["nside_y"], Lx, Ly)
    r += oracle.slp_kspace(ps, ps, prm["xi"], prm["eta"], f, prm["Mx"], prm["My"], Lx, Ly,
                           prm["w"], prm["P"])
    assert np.linalg.norm(u_slp - r) / np.linalg.norm(r) < 1e-13
    q = cprm
    r = oracle.dlp_real(cps, cpt, cf, cn, q["xi"], q["nside_x"], q["nside_y"], 1.0, 1.0)
    r += oracle.dlp_kspace(cps, cpt, q["xi"], q["eta"], cf, cn, q["Mx"], q["My"], 1.0, 1.0,
                           q["w"], q["P"]) + oracle.dlp_zero_mode(cps, cf, cn, 1.0, 1.0)
    assert np.linalg.norm(u_dlp - r) / np.linalg.norm(r) < 1e-13
    b.close()


def test_empty_batch_and_bad_system(gpu_ctx):
    b = Batch(gpu_ctx, [])
    assert b.apply([]) == []
    b.close()
    bad = _systems(1, seed=600, with_dlp=False)[0]
    bad = dict(bad, prm=dict(bad["prm"], Mx=bad["prm"]["Mx"] + 1))  # odd M: rejected
    with pytest.raises(ft.Se2pError):
        Batch(gpu_ctx, [bad])
