"""Velocity gradient, pressure gradient, vorticity and stress of both layers on
the GPU, through the C ABI: the real-space and k-space mex replacements against
the oracle (mex_stokes_{slp,dlp}_{gradient,pressure_grad,vorticity,stress}_
{real,kspace}.cpp), plans, graph replays and mixed batches bit for bit, and the
drivers (Stokes{SLP,DLP}_*_ewald_2p.m) against the oracle's parts assembled as
each driver assembles them."""
import numpy as np
import pytest

import fasttools2d_amd as ft
from fasttools2d_amd.batch import Batch
from fasttools2d_amd.plan import Plan
from tests.inputs import circles, normal_density, uniform_points

pytestmark = pytest.mark.gpu

KINDS = ["slp_gradient", "slp_pressure_grad", "slp_vorticity", "slp_stress", "dlp_gradient",
         "dlp_pressure_grad", "dlp_vorticity", "dlp_stress"]
CODE = {k: 4 + i for i, k in enumerate(KINDS)}


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _normals(N, seed):
    th = np.random.default_rng(seed).uniform(0, 2 * np.pi, N)
    return np.stack([np.cos(th), np.sin(th)], 1)


def _mex(kind, part):
    layer, q = kind.split("_", 1)
    return getattr(ft, f"mex_stokes_{layer}_{q}_{part}")


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("N,Lx,Ly,nsx,nsy", [(3000, 1.0, 1.0, 8, 8), (2500, 2.0, 1.0, 10, 5)])
def test_derived_real_matches_oracle(gpu_ctx, oracle, kind, N, Lx, Ly, nsx, nsy):
    ps = uniform_points(N, Lx, Ly, seed=401)
    pt = np.concatenate([uniform_points(N // 2, Lx, Ly, seed=402), ps[:50]])  # coincident
    f = normal_density(N, seed=403)
    n = _normals(N, seed=404) if kind.startswith("dlp") else None
    xi = 3.5 * nsx / Lx
    args = (ps.T, pt.T, f.T) + ((n.T,) if n is not None else ()) + (xi, nsx, nsy, Lx, Ly)
    got = _mex(kind, "real")(*args, ctx=gpu_ctx)
    ref = oracle.derived_real(CODE[kind], ps, pt, f, n, xi, nsx, nsy, Lx, Ly)
    assert got.shape == (ref.shape[1], len(pt))
    assert _rel(got.T, ref) < 1e-13


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("Lx,Ly,tol", [(1.0, 1.0, 1e-12), (2.0, 1.0, 1e-8)])
def test_derived_kspace_matches_oracle(gpu_ctx, oracle, kind, Lx, Ly, tol):
    N = 4000
    ps = uniform_points(N, Lx, Ly, seed=411)
    pt = uniform_points(3000, Lx, Ly, seed=412)
    f = normal_density(N, seed=413)
    dlp = kind.startswith("dlp")
    n = _normals(N, seed=414) if dlp else None
    p = ft.ewald_params(kind, N, 3000, ft.density_q(f[:, 0], f[:, 1]), Lx, Ly, tol=tol)
    fn = _mex(kind, "kspace")
    tail = (p["Mx"], p["My"], Lx, Ly, p["w"], p["P"])
    if kind in ("slp_pressure_grad", "dlp_pressure_grad", "dlp_vorticity"):  # f (n) before xi
        got = fn(ps.T, pt.T, f.T, *((n.T,) if dlp else ()), p["xi"], p["eta"], *tail, ctx=gpu_ctx)
    else:
        got = fn(ps.T, pt.T, p["xi"], p["eta"], f.T, *((n.T,) if dlp else ()), *tail,
                 ctx=gpu_ctx)
    ref = oracle.derived_kspace(CODE[kind], ps, pt, f, n, p["xi"], p["eta"], p["Mx"], p["My"],
                                Lx, Ly, p["w"], p["P"])
    assert _rel(got.T, ref) < 1e-13


def test_derived_plans_graphs_and_batch_bitwise(gpu_ctx, oracle):
    """Every derived kind as a plan (eager and graph-replayed) and all of them in
    one batch with a velocity system; the plan equals real + k-space of the
    oracle (+ the SLP vorticity's zero mode, StokesSLP_vorticity_ewald_2p.m:145)."""
    Lx, Ly = 1.5, 1.0
    ps, n, _ = circles(4, 300, Lx, Ly, seed=421)
    pt = uniform_points(2500, Lx, Ly, seed=422)
    Ns = len(ps)
    systems, plans = [], []
    for kind in KINDS + ["slp"]:
        prm = ft.ewald_params(kind, Ns, len(pt), 2.0 * Ns + 1, Lx, Ly, tol=1e-11)
        nrm = n if kind.startswith("dlp") else None
        systems.append(dict(kind=kind, psrc=ps, ptar=pt, nrm=nrm, Lx=Lx, Ly=Ly, prm=prm))
        plans.append(Plan(gpu_ctx, kind, ps, pt, Lx, Ly, prm, nrm=nrm))
    b = Batch(gpu_ctx, systems)
    for seed in (1, 2):
        f = normal_density(Ns, seed=seed)
        outs = b.apply([f] * len(systems))
        for kind, pl, y, ub in zip(KINDS + ["slp"], plans, systems, outs):
            u = pl.apply(f)
            np.testing.assert_array_equal(ub, u)
            np.testing.assert_array_equal(pl.apply(f, graph=True), u)
            if kind == "slp" or seed == 2:
                continue
            q, code = y["prm"], CODE[kind]
            nrm = y["nrm"]
            ref = oracle.derived_real(code, ps, pt, f, nrm, q["xi"], q["nside_x"], q["nside_y"],
                                      Lx, Ly)
            ref += oracle.derived_kspace(code, ps, pt, f, nrm, q["xi"], q["eta"], q["Mx"],
                                         q["My"], Lx, Ly, q["w"], q["P"])
            if kind == "slp_vorticity":
                ref -= (f[:, 0] * ps[:, 1] - f[:, 1] * ps[:, 0]).sum() / (2 * Lx * Ly)
            assert _rel(u.reshape(ref.shape), ref) < 1e-12
    for pl in plans:
        pl.close()
    b.close()


# ----------------------------------------------------------------- drivers
def _oracle_parts(oracle, kind, ps, pt, f, n, Lx, Ly, tol):
    p = ft.ewald_params(kind, len(ps), len(pt), ft.density_q(f[:, 0], f[:, 1]), Lx, Ly, tol=tol)
    psw, ptw = oracle.wrap(ps, Lx, Ly), oracle.wrap(pt, Lx, Ly)
    code = CODE[kind]
    R = oracle.derived_real(code, psw, ptw, f, n, p["xi"], p["nside_x"], p["nside_y"], Lx, Ly)
    K = oracle.derived_kspace(code, psw, ptw, f, n, p["xi"], p["eta"], p["Mx"], p["My"], Lx, Ly,
                              p["w"], p["P"])
    return psw, p, R, K


def _proj(T, b, order=(0, 2, 1, 3)):
    a, c, d, e = order
    return np.stack([T[:, a] * b[:, 0] + T[:, c] * b[:, 1], T[:, d] * b[:, 0] + T[:, e] * b[:, 1]])


@pytest.mark.parametrize("layer", ["slp", "dlp"])
def test_gradient_and_stress_drivers(gpu_ctx, oracle, layer):
    """Targets are the sources (the DLP drivers' self terms apply at every target)."""
    Lx, Ly, tol = 1.0, 2.0, 1e-12
    ps, n, f = circles(3, 200, Lx, Ly, seed=431)
    ps = ps + [0.3, -0.7]  # the drivers wrap
    b = _normals(len(ps), seed=432)
    dlp = layer == "dlp"
    nargs = (n[:, 0], n[:, 1]) if dlp else ()
    for q in ("gradient", "stress"):
        kind = f"{layer}_{q}"
        drv = getattr(ft, f"Stokes{layer.upper()}_{q}_ewald_2p")
        v1, v2, vr, vk, xi = drv(ps[:, 0], ps[:, 1], ps[:, 0], ps[:, 1], *nargs, f[:, 0],
                                 f[:, 1], b[:, 0], b[:, 1], Lx, Ly, tol=tol, ctx=gpu_ctx)
        psw, p, R, K = _oracle_parts(oracle, kind, ps, ps, f, n if dlp else None, Lx, Ly, tol)
        assert abs(xi - p["xi"]) <= 1e-12 * xi
        er = _proj(R, b, (0, 1, 2, 3) if kind == "dlp_stress" else (0, 2, 1, 3))
        ek = _proj(K, b)
        if kind == "slp_stress":
            ek_zm = (f * psw).sum() / (2 * Lx * Ly)
        elif kind == "dlp_stress":
            ek = ek + (n * f).sum() / (2 * Lx * Ly)
            ek_zm = 0.0
        else:
            ek_zm = 0.0
        assert _rel(vr, er) < 1e-12
        assert _rel(vk, ek) < 1e-12
        tot = er + ek + ek_zm
        if dlp:  # self terms at coincident points
            fn = (f * n).sum(1)
            ug = xi ** 2 * (f.T * (b * n).sum(1) + n.T * (f * b).sum(1) + b.T * fn) / (2 * np.pi)
            if q == "gradient":
                tot = tot + ug
            else:
                tot = tot + 2 * ug
                tot[0] -= xi ** 2 * b[:, 0] * fn / (2 * np.pi)
        assert _rel(np.stack([v1, v2]), tot) < 1e-12


@pytest.mark.parametrize("layer", ["slp", "dlp"])
def test_pressure_gradient_and_vorticity_drivers(gpu_ctx, oracle, layer):
    Lx, Ly, tol = 2.0, 1.0, 1e-12
    ps, n, f = circles(3, 200, Lx, Ly, seed=441)
    pt = np.concatenate([uniform_points(500, Lx, Ly, seed=442), ps[:40]])
    dlp = layer == "dlp"
    nargs = (n[:, 0], n[:, 1]) if dlp else ()
    # pressure gradient
    kind = f"{layer}_pressure_grad"
    drv = getattr(ft, f"Stokes{layer.upper()}_pressure_grad_ewald_2p")
    g, gr, gk, xi = drv(ps[:, 0], ps[:, 1], pt[:, 0], pt[:, 1], *nargs, f[:, 0], f[:, 1], Lx, Ly,
                        tol=tol, ctx=gpu_ctx)
    _, p, R, K = _oracle_parts(oracle, kind, ps, pt, f, n if dlp else None, Lx, Ly, tol)
    assert _rel(gr, R.T) < 1e-12 and _rel(gk, K.T) < 1e-12
    if dlp:  # StokesDLP_pressure_grad_ewald_2p.m:154
        want = -2 * (R.T + K.T)
    else:  # StokesSLP_pressure_grad_ewald_2p.m:151-156, pself at the 40 coincident targets
        want = -(R.T + K.T)
        want[:, 500:] -= xi ** 2 * f[:40].T / (2 * np.pi)
    assert _rel(g, want) < 1e-12
    # vorticity
    kind = f"{layer}_vorticity"
    drv = getattr(ft, f"Stokes{layer.upper()}_vorticity_ewald_2p")
    om, omr, omk, xi = drv(ps[:, 0], ps[:, 1], pt[:, 0], pt[:, 1], *nargs, f[:, 0], f[:, 1], Lx,
                           Ly, tol=tol, ctx=gpu_ctx)
    psw, p, R, K = _oracle_parts(oracle, kind, ps, pt, f, n if dlp else None, Lx, Ly, tol)
    assert _rel(omr, R[:, 0]) < 1e-12 and _rel(omk, K[:, 0]) < 1e-12
    want = R[:, 0] + K[:, 0]
    if not dlp:  # StokesSLP_vorticity_ewald_2p.m:145
        want = want - (f[:, 0] * psw[:, 1] - f[:, 1] * psw[:, 0]).sum() / (2 * Lx * Ly)
    assert _rel(om, want) < 1e-12


def test_dlp_pressure_driver_self_term(gpu_ctx, oracle):
    """StokesDLP_pressure_ewald_2p.m:165-172: at targets that are sources p gains
    xi^2 (f.n)/(2 pi)."""
    Lx, Ly = 1.0, 1.0
    ps, n, f = circles(3, 200, Lx, Ly, seed=451)
    pt = np.concatenate([uniform_points(300, Lx, Ly, seed=452), ps[:25]])
    p, pr, pk, xi = ft.StokesDLP_pressure_ewald_2p(ps[:, 0], ps[:, 1], pt[:, 0], pt[:, 1],
                                                   n[:, 0], n[:, 1], f[:, 0], f[:, 1], Lx, Ly,
                                                   tol=1e-12, ctx=gpu_ctx)
    want = pr + pk
    want[300:] += xi ** 2 * (f[:25] * n[:25]).sum(1) / (2 * np.pi)
    np.testing.assert_allclose(p, want, rtol=0, atol=1e-13 * np.abs(want).max())
