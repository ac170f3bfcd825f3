"""Pins of the oracle's derived quantities (velocity gradient, pressure gradient,
vorticity, stress of both layers; mex_stokes_{slp,dlp}_{gradient,pressure_grad,
vorticity,stress}_{real,kspace}.cpp) against what fixes them independently:

* the reference's direct sums (tests/direct_sums/direct_sums_{gradient,
  pressure_grad,vorticity,stress}_test.m with matlab_sums/stokes_*_ds.m), which
  state each kernel in other files and in another form (projected on a
  per-target vector b; the k-space part as an explicit mode sum);
* calculus on the already pinned velocity and pressure: the gradient is the
  derivative of the velocity, the vorticity its curl, the stress
  -p I + grad u + grad u^T (mu = 1), the pressure gradient the derivative of
  the pressure (with the sign each mex file uses);
* independence of every total from the Ewald split.
"""
import numpy as np
import pytest

from tests.inputs import normal_density, uniform_points

KINDS = {"slp_gradient": 4, "slp_pressure_grad": 5, "slp_vorticity": 6, "slp_stress": 7,
         "dlp_gradient": 8, "dlp_pressure_grad": 9, "dlp_vorticity": 10, "dlp_stress": 11}


def _recipe(kind):
    # StokesDLP_vorticity_ewald_2p.m uses the SLP recipe; other stresslet drivers the DLP one
    return "dlp" if kind >= 8 and kind != 10 else "slp"


def _params(oracle, kind, npts, Q, Lx, Ly, tol, Nb=None, P=24):
    Nb = Nb if Nb else oracle.default_Nb(npts, P)
    return oracle.params(_recipe(kind), npts, Q, Lx, Ly, P, Nb, tol)


def _project(T, b):
    """The drivers' projection u_j = sum_p b_p T[2p + j] (e.g.
    StokesSLP_gradient_ewald_2p.m:133-135)."""
    return np.stack([b[:, 0] * T[:, 0] + b[:, 1] * T[:, 2],
                     b[:, 0] * T[:, 1] + b[:, 1] * T[:, 3]], 1)


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("name", sorted(KINDS, key=KINDS.get))
def test_derived_parts_match_direct_sums_like_reference(oracle, name):
    """direct_sums_*_test.m: 10 sources, 10 targets, tol 1e-16, kinf = 70, the
    gradient and stress probed with b = (1, 0) and (0, 1)."""
    kind = KINDS[name]
    rng = np.random.default_rng(300 + kind)
    Lx, Ly = (1.0, 2.0) if kind in (4, 8) else (1.0, 1.0)
    f = 10 * rng.random((10, 2))
    n1 = rng.random(10)
    nrm = np.stack([n1, np.sqrt(1 - n1 ** 2)], 1) if kind >= 8 else None
    ps = oracle.wrap(np.stack([Lx * rng.random(10), Ly * rng.random(10)], 1), Lx, Ly)
    pt = oracle.wrap(np.stack([Lx * rng.random(10), Ly * rng.random(10)], 1), Lx, Ly)
    prm = _params(oracle, kind, 20, (f ** 2).sum() + 1, Lx, Ly, 1e-16)
    xi = prm["xi"]
    R = oracle.derived_real(kind, ps, pt, f, nrm, xi, prm["nside_x"], prm["nside_y"], Lx, Ly)
    K = oracle.derived_kspace(kind, ps, pt, f, nrm, xi, prm["eta"], prm["Mx"], prm["My"], Lx,
                              Ly, prm["w"], 24)
    probes = ([1.0, 0.0], [0.0, 1.0]) if R.shape[1] == 4 else (None,)
    for bb in probes:
        b = None if bb is None else np.tile(bb, (10, 1))
        rds = oracle.derived_real_ds(kind, ps, pt, f, nrm, b, Lx, Ly, xi)
        kds = oracle.derived_kspace_ds(kind, ps, pt, f, nrm, b, Lx, Ly, xi, 70)
        if b is not None:
            r, k = _project(R, b), _project(K, b)
            if kind == 11:  # StokesDLP_stress_ewald_2p.m:167-169 (also in the direct sum)
                k = k + (nrm * f).sum() / (2 * Lx * Ly)
        else:
            r, k = (R[:, 0], K[:, 0]) if R.shape[1] == 1 else (R, K)
        assert _rel(r, rds) < 1e-13
        assert _rel(k, kds) < 2e-11


# ---------------------------------------------------------------- calculus
LX, LY = 1.0, 1.0


@pytest.fixture(scope="module")
def setup(oracle):
    rng = np.random.default_rng(77)
    ps = oracle.wrap(uniform_points(200, LX, LY, seed=78), LX, LY)
    pt = uniform_points(30, 0.8 * LX, 0.8 * LY, seed=79)
    f = normal_density(200, seed=80)
    th = rng.uniform(0, 2 * np.pi, 200)
    nrm = np.stack([np.cos(th), np.sin(th)], 1)
    return ps, pt, f, nrm


def _total(oracle, setup, kind, pts, Nb=None):
    """Real + k-space part (no zero mode) of kind 0..11 at targets pts, tol 1e-14."""
    ps, _, f, nrm = setup
    dlp = kind in (1, 3) or kind >= 8
    npts = len(ps) + 30
    p = _params(oracle, kind if kind >= 4 else (8 if dlp else 4), npts, (f ** 2).sum() + 1, LX,
                LY, 1e-14, Nb)
    a = (p["xi"], p["nside_x"], p["nside_y"], LX, LY)
    k = (p["Mx"], p["My"], LX, LY, p["w"], 24)
    if kind == 0:
        return oracle.slp_real(ps, pts, f, *a) + oracle.slp_kspace(ps, pts, p["xi"], p["eta"], f,
                                                                   *k)
    if kind == 1:
        return oracle.dlp_real(ps, pts, f, nrm, *a) + oracle.dlp_kspace(
            ps, pts, p["xi"], p["eta"], f, nrm, *k)
    if kind == 2:
        return oracle.slp_pressure_real(ps, pts, f, *a) + oracle.slp_pressure_kspace(
            ps, pts, f, p["xi"], p["eta"], *k)
    if kind == 3:
        return oracle.dlp_pressure_real(ps, pts, f, nrm, *a) + oracle.dlp_pressure_kspace(
            ps, pts, f, nrm, p["xi"], p["eta"], *k)
    n = nrm if dlp else None
    return oracle.derived_real(kind, ps, pts, f, n, *a) + oracle.derived_kspace(
        kind, ps, pts, f, n, p["xi"], p["eta"], *k)


def _fd(oracle, setup, kind, h=1e-5):
    pt = setup[1]
    gx = (_total(oracle, setup, kind, pt + [h, 0]) - _total(oracle, setup, kind, pt - [h, 0]))
    gy = (_total(oracle, setup, kind, pt + [0, h]) - _total(oracle, setup, kind, pt - [0, h]))
    return gx / (2 * h), gy / (2 * h)


@pytest.mark.parametrize("vel,grad", [(0, 4), (1, 8)])
def test_gradient_is_derivative_of_velocity(oracle, setup, vel, grad):
    """Slots [d/dx u1, d/dx u2, d/dy u1, d/dy u2] against central differences of
    the velocity sum (step 1e-5: the difference error is ~1e-5 relative)."""
    gx, gy = _fd(oracle, setup, vel)
    G = _total(oracle, setup, grad, setup[1])
    assert _rel(G, np.concatenate([gx, gy], 1)) < 3e-5


@pytest.mark.parametrize("pres,pgrad,sign", [(2, 5, -1), (3, 9, 1)])
def test_pressure_gradient_is_derivative_of_pressure(oracle, setup, pres, pgrad, sign):
    """The Stokeslet mex sums -grad p (its real part accumulates with -=,
    mex_stokes_slp_pressure_grad_real.cpp:120-121; the driver negates,
    StokesSLP_pressure_grad_ewald_2p.m:156); the stresslet's sum +grad p."""
    gx, gy = _fd(oracle, setup, pres)
    PG = _total(oracle, setup, pgrad, setup[1])
    assert _rel(PG, sign * np.stack([gx, gy], 1)) < 3e-5


@pytest.mark.parametrize("grad,vort", [(4, 6), (8, 10)])
def test_vorticity_is_curl_of_gradient(oracle, setup, grad, vort):
    G = _total(oracle, setup, grad, setup[1])
    om = _total(oracle, setup, vort, setup[1])
    assert _rel(om[:, 0], G[:, 1] - G[:, 2]) < 1e-11


@pytest.mark.parametrize("grad,pres,stress", [(4, 2, 7), (8, 3, 11)])
def test_stress_is_pressure_and_rate_of_strain(oracle, setup, grad, pres, stress):
    """sigma = -p I + grad u + grad u^T with mu = 1."""
    G = _total(oracle, setup, grad, setup[1])
    p = _total(oracle, setup, pres, setup[1])
    S = _total(oracle, setup, stress, setup[1])
    want = -p[:, None] * np.array([1.0, 0, 0, 1.0]) + G + G[:, [0, 2, 1, 3]]
    assert _rel(S, want) < 1e-10


@pytest.mark.parametrize("kind", sorted(KINDS.values()))
def test_derived_totals_independent_of_ewald_split(oracle, setup, kind):
    """consistency_test_velocity.m:54-76 applied to each quantity: Nb 9 vs 27."""
    a = _total(oracle, setup, kind, setup[1], Nb=9)
    b = _total(oracle, setup, kind, setup[1], Nb=27)
    assert _rel(a, b) < 1e-10
