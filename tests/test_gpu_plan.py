"""Plans and graph replay: same results as the one-shot path, bit for bit."""
import numpy as np
import pytest

import fasttools2d_amd as ft
from fasttools2d_amd.plan import DeviceArray, Plan, params_struct
from fasttools2d_amd import _lib
from tests.inputs import normal_density, uniform_points

pytestmark = pytest.mark.gpu


def _eval_device(ctx, kind, ps, pt, f, n, Lx, Ly, prm):
    d = [DeviceArray.from_host(ctx, a) for a in (ps, pt, f)]
    dn = DeviceArray.from_host(ctx, n) if n is not None else None
    du = DeviceArray(ctx, (len(pt), 2))
    s = params_struct(prm)
    _lib.check(_lib.lib().se2p_eval_device(ctx.handle, kind, d[0].ptr, d[1].ptr, d[2].ptr,
                                           dn.ptr if dn else None, len(ps), len(pt), Lx, Ly,
                                           __import__("ctypes").byref(s), du.ptr))
    ctx.sync()
    u = du.download()
    for a in d + [du] + ([dn] if dn else []):
        a.free()
    return u


@pytest.mark.parametrize("kind", ["slp", "dlp"])
def test_plan_matches_one_shot_and_graph_replays(gpu_ctx, oracle, kind):
    Lx, Ly = 1.0, 2.0
    ps = uniform_points(6000, Lx, Ly, seed=61)
    pt = uniform_points(4000, Lx, Ly, seed=62)
    th = np.random.default_rng(63).random(6000) * 2 * np.pi
    n = np.stack([np.cos(th), np.sin(th)], 1) if kind == "dlp" else None
    prm = ft.ewald_params(kind, 6000, 4000, 2 * 6000 + 1.0, Lx, Ly, tol=1e-12)
    k = _lib.SE2P_SLP if kind == "slp" else _lib.SE2P_DLP
    plan = Plan(gpu_ctx, kind, ps, pt, Lx, Ly, prm, nrm=n)
    for seed in (1, 2, 3):
        f = normal_density(6000, seed=seed)
        ref = _eval_device(gpu_ctx, k, ps, pt, f, n, Lx, Ly, prm)
        u_eager = plan.apply(f)
        np.testing.assert_array_equal(u_eager, ref)
        # replay the captured graph several times with eager work in between
        for _ in range(3):
            u_graph = plan.apply(f, graph=True)
            np.testing.assert_array_equal(u_graph, ref)
            np.testing.assert_array_equal(plan.apply(f), ref)
    # and the one-shot result agrees with the oracle
    f = normal_density(6000, seed=3)
    if kind == "slp":
        r = oracle.slp_real(ps, pt, f, prm["xi"], prm["nside_x"], prm["nside_y"], Lx, Ly)
        r += oracle.slp_kspace(ps, pt, prm["xi"], prm["eta"], f, prm["Mx"], prm["My"], Lx, Ly,
                               prm["w"], prm["P"])
    else:
        r = oracle.dlp_real(ps, pt, f, n, prm["xi"], prm["nside_x"], prm["nside_y"], Lx, Ly)
        r += oracle.dlp_kspace(ps, pt, prm["xi"], prm["eta"], f, n, prm["Mx"], prm["My"], Lx,
                               Ly, prm["w"], prm["P"]) + oracle.dlp_zero_mode(ps, f, n, Lx, Ly)
    got = plan.apply(f, graph=True)
    assert np.linalg.norm(got - r) / np.linalg.norm(r) < 1e-13
    plan.close()


@pytest.mark.parametrize("kind", ["slp", "dlp", "slp_vorticity", "slp_stress", "dlp_stress"])
def test_two_stream_apply_matches_one_stream_order(gpu_ctx, kind):
    """An apply runs the real-space sum and the k-space grid passes on two streams
    joined before the gather; with one-stream stage profiling they run on one
    stream in order.  Both orders give the same bits (kinds with a zero mode and
    with a second output grid included), eager and graph-replayed, call after
    call, and so does an apply timed on both streams."""
    Lx, Ly = 1.0, 1.5
    ps = uniform_points(5000, Lx, Ly, seed=91)
    pt = uniform_points(3000, Lx, Ly, seed=92)
    th = np.random.default_rng(93).random(5000) * 2 * np.pi
    dlp = kind.startswith("dlp")
    n = np.stack([np.cos(th), np.sin(th)], 1) if dlp else None
    prm = ft.ewald_params("dlp" if dlp else "slp", 5000, 3000, 2 * 5000 + 1.0, Lx, Ly, tol=1e-10)
    plan = Plan(gpu_ctx, kind, ps, pt, Lx, Ly, prm, nrm=n)
    try:
        for seed in (4, 5):
            f = normal_density(5000, seed=seed)
            gpu_ctx.prof_enable(True, serial=True)
            try:
                one = plan.apply(f)
            finally:
                gpu_ctx.prof_enable(False)
            for _ in range(2):
                np.testing.assert_array_equal(plan.apply(f), one)
                np.testing.assert_array_equal(plan.apply(f, graph=True), one)
            gpu_ctx.prof_enable(True)
            try:
                np.testing.assert_array_equal(plan.apply(f), one)
            finally:
                gpu_ctx.prof_enable(False)
    finally:
        plan.close()
