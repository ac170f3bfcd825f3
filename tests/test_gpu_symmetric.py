"""Pair-symmetric real space of the velocity sums (sources = targets,
se2p_params.symmetric): the Stokeslet, and the stresslet at the end.

Each pair is evaluated once and added to both points with atomics, so results
are compared to tolerance, not bit for bit: against the fixed-order sums of the
same plan with symmetric = 0 (which the parity suite pins to the oracle) and
against the oracle itself.  A pair taken twice or missed moves the result by
about 1e-2 relative; the bound is 1e-13.  The cases cover the orders the pair
rule has to break: exact duplicates (the self term across coincident points),
equal y keys inside a cell and across neighbouring cells, boxes one and two
cells wide (a cell's own periodic images as neighbours), windows staged in
several rounds and blocks of more than 256 targets (a dense cluster), one-cell
blocks, and inputs the y order cannot handle (fallback to the fixed order).
"""
import numpy as np
import pytest

import fasttools2d_amd as ft
from fasttools2d_amd.batch import Batch
from fasttools2d_amd.plan import Plan
from tests.inputs import normal_density, uniform_points

pytestmark = pytest.mark.gpu

REL = 1e-13


def _prm(N, Lx, Ly, nsx, nsy, tol=1e-10):
    prm = dict(ft.ewald_params("slp", N, N, 2 * N + 1.0, Lx, Ly, tol=tol))
    prm["nside_x"], prm["nside_y"] = nsx, nsy
    prm["rc"] = Lx / nsx
    return prm


def _apply(ctx, p, f, Lx, Ly, prm, sym, graph=False, reps=1):
    q = dict(prm, symmetric=1 if sym else 0)
    plan = Plan(ctx, "slp", p, p, Lx, Ly, q)
    try:
        return [plan.apply(f, graph=graph) for _ in range(reps)]
    finally:
        plan.close()


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _check(ctx, p, f, Lx, Ly, prm, oracle=None):
    ref = _apply(ctx, p, f, Lx, Ly, prm, False)[0]
    for u in _apply(ctx, p, f, Lx, Ly, prm, True, reps=2):
        assert _rel(u, ref) < REL
        assert np.max(np.abs(u - ref)) < REL * np.max(np.abs(ref)) * 10
        # the pairs were summed in the other order (the symmetric path ran)
        assert np.any(u != ref)
    if oracle is not None:
        r = oracle.slp_real(p, p, f, prm["xi"], prm["nside_x"], prm["nside_y"], Lx, Ly)
        r += oracle.slp_kspace(p, p, prm["xi"], prm["eta"], f, prm["Mx"], prm["My"], Lx, Ly,
                               prm["w"], prm["P"])
        assert _rel(_apply(ctx, p, f, Lx, Ly, prm, True)[0], r) < REL


@pytest.mark.parametrize("N,nside", [(6000, 18), (20000, 33), (3000, 4)])
def test_uniform_matches_fixed_order_and_oracle(gpu_ctx, oracle, N, nside):
    p = uniform_points(N, 1.0, 1.0, seed=N)
    f = normal_density(N, seed=N + 1)
    _check(gpu_ctx, p, f, 1.0, 1.0, _prm(N, 1.0, 1.0, nside, nside), oracle)


def test_duplicates_and_equal_keys(gpu_ctx, oracle):
    """Exact duplicates, a lattice with shared y values (equal keys within and
    across cells) and points on cell edges."""
    Lx = Ly = 1.0
    rng = np.random.default_rng(5)
    base = uniform_points(2500, Lx, Ly, seed=6)
    dup = base[rng.choice(2500, 300, replace=False)]
    g = (np.arange(24) + 0.5) / 24 - 0.5
    lat = np.stack(np.meshgrid(g, g), -1).reshape(-1, 2)
    edges = np.stack([np.arange(-6, 6) / 12.0, np.full(12, 0.25)], 1)
    p = np.concatenate([base, dup, lat, edges])
    f = normal_density(len(p), seed=7)
    _check(gpu_ctx, p, f, Lx, Ly, _prm(len(p), Lx, Ly, 12, 12), oracle)


@pytest.mark.parametrize("Lx,Ly,nsx,nsy", [(1.0, 1.0, 3, 3), (2.0, 1.0, 2, 1), (1.0, 3.0, 1, 3),
                                           (4.0, 1.0, 4, 1)])
def test_narrow_boxes(gpu_ctx, oracle, Lx, Ly, nsx, nsy):
    N = 1500
    p = uniform_points(N, Lx, Ly, seed=11)
    f = normal_density(N, seed=12)
    prm = ft.ewald_params("slp", N, N, 2 * N + 1.0, Lx, Ly, tol=1e-10)
    prm = dict(prm, nside_x=nsx, nside_y=nsy)
    _check(gpu_ctx, p, f, Lx, Ly, prm, oracle)


def test_cluster_rounds_and_wide_blocks(gpu_ctx, oracle):
    """A dense cluster in a sparse background: windows beyond the staging buffer
    (several rounds) and blocks of more than 256 targets."""
    Lx = Ly = 1.0
    rng = np.random.default_rng(21)
    bg = uniform_points(4000, Lx, Ly, seed=22)
    cl = np.clip(rng.normal(0.1, 0.01, (2500, 2)), -0.49, 0.49)
    p = np.concatenate([bg, cl])
    f = normal_density(len(p), seed=23)
    _check(gpu_ctx, p, f, Lx, Ly, _prm(len(p), Lx, Ly, 20, 20), oracle)


def test_one_cell_blocks(gpu_ctx):
    N = 30000
    p = uniform_points(N, 1.0, 1.0, seed=31)
    f = normal_density(N, seed=32)
    _check(gpu_ctx, p, f, 1.0, 1.0, _prm(N, 1.0, 1.0, 20, 20))


def test_unordered_input_falls_back_to_fixed_order(gpu_ctx):
    """A point outside the box leaves its cell without a y order: the plan then
    sums in the fixed order, bit for bit."""
    N = 4000
    p = uniform_points(N, 1.0, 1.0, seed=41)
    p[7] = [0.1, 0.7]
    f = normal_density(N, seed=42)
    prm = _prm(N, 1.0, 1.0, 14, 14)
    np.testing.assert_array_equal(_apply(gpu_ctx, p, f, 1.0, 1.0, prm, True)[0],
                                  _apply(gpu_ctx, p, f, 1.0, 1.0, prm, False)[0])


def test_graph_replays_and_distinct_arrays(gpu_ctx):
    """Graph replays agree with eager applies; separate source and target arrays
    (equal values, different pointers) keep the fixed order."""
    N = 8000
    p = uniform_points(N, 1.0, 1.0, seed=51)
    f = normal_density(N, seed=52)
    prm = _prm(N, 1.0, 1.0, 21, 21)
    ref = _apply(gpu_ctx, p, f, 1.0, 1.0, prm, False)[0]
    for u in _apply(gpu_ctx, p, f, 1.0, 1.0, prm, True, graph=True, reps=4):
        assert _rel(u, ref) < REL
    plan = Plan(gpu_ctx, "slp", p, p.copy(), 1.0, 1.0, dict(prm, symmetric=1))
    try:
        np.testing.assert_array_equal(plan.apply(f), ref)
    finally:
        plan.close()


def test_batch_mixes_symmetric_and_fixed_order_systems(gpu_ctx):
    systems, fs = [], []
    for i, (N, Lx, Ly, ns) in enumerate([(5000, 1.0, 1.0, 16), (3000, 2.0, 1.0, 12)]):
        p = uniform_points(N, Lx, Ly, seed=60 + i)
        prm = _prm(N, Lx, Ly, ns * (2 if Lx > Ly else 1), ns)
        systems.append({"kind": "slp", "psrc": p, "ptar": p, "Lx": Lx, "Ly": Ly,
                        "prm": dict(prm, symmetric=1 - i)})
        fs.append(normal_density(N, seed=70 + i))
    b = Batch(gpu_ctx, systems)
    try:
        got = b.apply(fs)
        got_g = b.apply(fs, graph=True)
    finally:
        b.close()
    for y, f, u, ug in zip(systems, fs, got, got_g):
        ref = _apply(gpu_ctx, y["psrc"], f, y["Lx"], y["Ly"], y["prm"], False)[0]
        assert _rel(u, ref) < REL and _rel(ug, ref) < REL
        if not y["prm"]["symmetric"]:
            np.testing.assert_array_equal(u, ref)


def _dlp_apply(ctx, p, f, n, Lx, Ly, prm, sym, graph=False, reps=1):
    plan = Plan(ctx, "dlp", p, p, Lx, Ly, dict(prm, symmetric=1 if sym else 0), nrm=n)
    try:
        used = plan.params_in_use()["symmetric"]
        return [plan.apply(f, graph=graph) for _ in range(reps)], used
    finally:
        plan.close()


def _normals(N, seed):
    th = np.random.default_rng(seed).random(N) * 2 * np.pi
    return np.stack([np.cos(th), np.sin(th)], 1)


@pytest.mark.parametrize("case", ["uniform", "duplicates", "narrow", "cluster"])
def test_stresslet_pairs_once_for_both_points(gpu_ctx, oracle, case):
    """The stresslet's pair-symmetric sums (its kernel is odd in r, so a pair's
    term for the source is the target's expression negated, with the target's
    density and normal) against its fixed-order sums and the oracle: uniform
    points, exact duplicates (coincident pairs are skipped both ways), a box
    one cell wide, and a dense cluster (windows in several rounds, blocks of
    more than 256 targets)."""
    Lx, Ly, nsx, nsy = 1.0, 1.0, 18, 18
    if case == "uniform":
        p = uniform_points(6000, Lx, Ly, seed=81)
    elif case == "duplicates":
        q = uniform_points(2500, Lx, Ly, seed=82)
        p = np.concatenate([q, q[:500]])
    elif case == "narrow":
        Lx, Ly, nsx, nsy = 3.0, 1.0, 3, 1
        p = uniform_points(2000, Lx, Ly, seed=83)
    else:
        rng = np.random.default_rng(84)
        p = np.concatenate([uniform_points(3000, Lx, Ly, seed=85),
                            np.clip(rng.normal(-0.2, 0.01, (2500, 2)), -0.49, 0.49)])
    N = len(p)
    n = _normals(N, 86)
    f = normal_density(N, seed=87)
    prm = dict(ft.ewald_params("dlp", N, N, 2 * N + 1.0, Lx, Ly, tol=1e-10))
    prm["nside_x"], prm["nside_y"], prm["rc"] = nsx, nsy, Lx / nsx
    (ref,), used0 = _dlp_apply(gpu_ctx, p, f, n, Lx, Ly, prm, False)
    got, used1 = _dlp_apply(gpu_ctx, p, f, n, Lx, Ly, prm, True, reps=2)
    assert (used0, used1) == (0, 1)
    for u in got:
        assert _rel(u, ref) < REL
        assert np.any(u != ref)  # the pairs were summed in the other order
    r = oracle.dlp_real(p, p, f, n, prm["xi"], nsx, nsy, Lx, Ly)
    r += oracle.dlp_kspace(p, p, prm["xi"], prm["eta"], f, n, prm["Mx"], prm["My"], Lx, Ly,
                           prm["w"], prm["P"])
    r += oracle.dlp_zero_mode(p, f, n, Lx, Ly)
    assert _rel(got[0], r) < REL


def test_stresslet_fp32_tier_and_graph_replays(gpu_ctx, oracle):
    """The stresslet's pair-symmetric sums in the FP32 tier within 1e-6 of the
    oracle, and FP64 graph replays equal to their eager applies to rounding."""
    N = 8000
    p = uniform_points(N, 1.0, 1.0, seed=88)
    n = _normals(N, 89)
    f = normal_density(N, seed=90)
    prm = dict(ft.ewald_params("dlp", N, N, 2 * N + 1.0, 1.0, 1.0, tol=1e-10))
    prm["nside_x"] = prm["nside_y"] = 21
    prm["rc"] = 1.0 / 21
    r = oracle.dlp_real(p, p, f, n, prm["xi"], 21, 21, 1.0, 1.0)
    r += oracle.dlp_kspace(p, p, prm["xi"], prm["eta"], f, n, prm["Mx"], prm["My"], 1.0, 1.0,
                           prm["w"], prm["P"])
    r += oracle.dlp_zero_mode(p, f, n, 1.0, 1.0)
    (u32,), used = _dlp_apply(gpu_ctx, p, f, n, 1.0, 1.0, dict(prm, precision=1), True)
    assert used == 1 and _rel(u32, r) < 1e-6
    (eager,), _ = _dlp_apply(gpu_ctx, p, f, n, 1.0, 1.0, prm, True)
    replays, _ = _dlp_apply(gpu_ctx, p, f, n, 1.0, 1.0, prm, True, graph=True, reps=3)
    for u in replays:
        assert _rel(u, eager) < REL and _rel(u, r) < REL


def test_full_size_tuned_split_matches_fixed_order(gpu_ctx):
    """The benchmark's configuration at its full size (N = 1e6 sources = targets,
    fast E1 tier, the tuner's split): pair-symmetric and fixed-order sums agree to
    rounding, eagerly and replayed from a graph."""
    from fasttools2d_amd.tune import make_params
    N = 1_000_000
    rng = np.random.default_rng(1)
    p = rng.random((N, 2)) - 0.5
    f = rng.standard_normal((N, 2))
    prm = make_params(3.8 * 201, 201, 201, 2048, 2048, 14, 1.0, 1)
    ref = _apply(gpu_ctx, p, f, 1.0, 1.0, prm, False)[0]
    for g in (False, True):
        u = _apply(gpu_ctx, p, f, 1.0, 1.0, prm, True, graph=g, reps=2)
        for v in u:
            assert _rel(v, ref) < REL


@pytest.mark.parametrize("sym", [False, True])
def test_fast_e1_tier_within_its_bound(gpu_ctx, oracle, sym):
    """The fast E1 tier (library log, exponential and reciprocals, E1 fits with
    relative error below 1e-12) against the full tier and the oracle on the same
    split: the sums agree to 1e-12, near (xi r < 1) and far pairs both present."""
    N = 20000
    p = uniform_points(N, 1.0, 1.0, seed=91)
    f = normal_density(N, seed=92)
    prm = _prm(N, 1.0, 1.0, 33, 33)
    full = _apply(gpu_ctx, p, f, 1.0, 1.0, prm, sym)[0]
    fast = _apply(gpu_ctx, p, f, 1.0, 1.0, dict(prm, e1_tier=1), sym)[0]
    assert _rel(fast, full) < 1e-12
    r = oracle.slp_real(p, p, f, prm["xi"], 33, 33, 1.0, 1.0)
    r += oracle.slp_kspace(p, p, prm["xi"], prm["eta"], f, prm["Mx"], prm["My"], 1.0, 1.0,
                           prm["w"], prm["P"])
    assert _rel(fast, r) < 1e-12


def _cluster_cell(n_bg, n_cl, center, sigma, seed):
    rng = np.random.default_rng(seed)
    bg = uniform_points(n_bg, 1.0, 1.0, seed=seed + 1)
    cl = np.clip(rng.normal(center, sigma, (n_cl, 2)), -0.499, 0.499)
    return np.concatenate([bg, cl])


def test_cells_beyond_the_in_lds_sort(gpu_ctx, oracle):
    """A cell of more than 4096 points (a tight cluster) is put in y order by the
    global radix sort, so the pair-symmetric sums still run, and the fixed-order
    sums trim their stencil columns by y: both match the oracle."""
    p = _cluster_cell(3000, 6000, [0.2, -0.1], 0.004, 51)
    f = normal_density(len(p), seed=53)
    prm = _prm(len(p), 1.0, 1.0, 12, 12)
    cells = np.floor((p + 0.5) * 12).astype(int)
    assert np.max(np.bincount(cells[:, 1] * 12 + cells[:, 0])) > 4096
    plan = Plan(gpu_ctx, "slp", p, p, 1.0, 1.0, dict(prm, symmetric=1))
    try:
        assert plan.params_in_use()["symmetric"] == 1
    finally:
        plan.close()
    _check(gpu_ctx, p, f, 1.0, 1.0, prm, oracle)
    r = oracle.slp_real(p, p, f, prm["xi"], 12, 12, 1.0, 1.0)
    r += oracle.slp_kspace(p, p, prm["xi"], prm["eta"], f, prm["Mx"], prm["My"], 1.0, 1.0,
                           prm["w"], prm["P"])
    assert _rel(_apply(gpu_ctx, p, f, 1.0, 1.0, prm, False)[0], r) < REL


def test_big_cell_with_a_point_outside_falls_back(gpu_ctx):
    """A point outside the box in a cell of more than 4096 points: that cell has
    no y order, so the plan sums each target's pairs in the fixed order.  Its
    targets keep the sources' order, so its waves, and with them the chunking of
    long candidate lists, differ from a plan asked for the fixed order: the two
    agree to rounding, not bit for bit."""
    p = _cluster_cell(3000, 6000, [0.0, 0.46], 0.004, 61)
    p[-1] = [0.0, 0.52]
    f = normal_density(len(p), seed=63)
    prm = _prm(len(p), 1.0, 1.0, 12, 12)
    plan = Plan(gpu_ctx, "slp", p, p, 1.0, 1.0, dict(prm, symmetric=1))
    try:
        assert plan.params_in_use()["symmetric"] == 0
    finally:
        plan.close()
    u = _apply(gpu_ctx, p, f, 1.0, 1.0, prm, True)
    assert _rel(u[0], _apply(gpu_ctx, p, f, 1.0, 1.0, prm, False)[0]) < REL
