"""FP32 far pairs in the fast E1 tier of the Stokeslet (se2p_params.far32).

Pairs with xi^2 r^2 >= far32 are evaluated in FP32 and folded into the FP64
sums per chunk.  The tests bound what that changes: the whole sum moves by far
less than the tolerance (tune.far32_for puts it near 5e-4 tol), it does move
(the FP32 path runs), the result still meets the tolerance against the
oracle, plans and batches agree, and a negative threshold is refused."""
import numpy as np
import pytest

from fasttools2d_amd import _lib
from fasttools2d_amd.batch import Batch
from fasttools2d_amd.plan import Plan
from fasttools2d_amd.tune import far32_for, rule_params
from tests.inputs import normal_density, oracle_se, rel, uniform_points

pytestmark = pytest.mark.gpu


def _prm(n, Lx, Ly, tol, sym, far32):
    prm = rule_params("slp", n, n, Lx, Ly, tol, symmetric=sym)
    assert prm["e1_tier"] == 1 and prm["precision"] == 0
    prm["far32"] = far32
    return prm


@pytest.mark.parametrize("sym", [0, 1])
@pytest.mark.parametrize("tol", [1e-6 * 0.9, 1e-8, 1e-10])
def test_far32_moves_the_sum_far_below_tol(gpu_ctx, sym, tol):
    n, Lx, Ly = 120_000, 2.0, 1.0
    p = np.ascontiguousarray(uniform_points(n, Lx, Ly, seed=31))
    f = normal_density(n, seed=32)
    out = {}
    for far in (0.0, far32_for(tol)):
        pl = Plan(gpu_ctx, "slp", p, p, Lx, Ly, _prm(n, Lx, Ly, tol, sym, far))
        try:
            assert pl.params_in_use()["far32"] == far
            out[far] = pl.apply(f)
        finally:
            pl.close()
    far = far32_for(tol)
    assert far > 0
    d = rel(out[far], out[0.0])
    print(f"far32 {far:.2f} tol {tol:.1e} sym {sym}: change {d:.2e} ({d / tol:.1e} tol)")
    assert 1e-15 < d < 1e-2 * tol


def test_far32_meets_tol_against_oracle(gpu_ctx, oracle):
    n, Lx, Ly, tol = 6000, 1.0, 1.0, 1e-8
    p = np.ascontiguousarray(uniform_points(n, Lx, Ly, seed=41))
    f = normal_density(n, seed=42)
    prm = _prm(n, Lx, Ly, tol, 1, far32_for(tol))
    pl = Plan(gpu_ctx, "slp", p, p, Lx, Ly, prm)
    try:
        u = pl.apply(f)
    finally:
        pl.close()
    ref = oracle_se(oracle, "slp", p, p, f, None, Lx, Ly)
    e = rel(u, ref)
    print(f"far32 {prm['far32']:.2f}: rel err vs oracle {e:.2e}")
    assert e < tol


def test_far32_batch_equals_plans(gpu_ctx):
    """Batch systems with and without far32, each bitwise equal to its own plan
    (fixed-order sums)."""
    systems, refs = [], []
    for i, (n, tol, far) in enumerate([(9000, 1e-8, True), (15000, 1e-9, False),
                                       (4000, 3e-7, True)]):
        Lx = 1.0 + i * 0.5
        p = np.ascontiguousarray(uniform_points(n, Lx, 1.0, seed=50 + i))
        prm = _prm(n, Lx, 1.0, tol, 0, far32_for(tol) if far else 0.0)
        systems.append({"psrc": p, "ptar": p, "Lx": Lx, "Ly": 1.0, "prm": prm})
    fs = [normal_density(len(y["psrc"]), seed=60 + i) for i, y in enumerate(systems)]
    for y, f in zip(systems, fs):
        pl = Plan(gpu_ctx, "slp", y["psrc"], y["ptar"], y["Lx"], y["Ly"], y["prm"])
        try:
            refs.append(pl.apply(f))
        finally:
            pl.close()
    b = Batch(gpu_ctx, systems)
    try:
        us = b.apply(fs)
        for i in range(len(systems)):
            assert b.params_in_use(i)["far32"] == systems[i]["prm"]["far32"]
    finally:
        b.close()
    for u, r in zip(us, refs):
        assert np.array_equal(u, r)


def test_far32_negative_is_refused(gpu_ctx):
    n = 2000
    p = np.ascontiguousarray(uniform_points(n, 1.0, 1.0, seed=70))
    with pytest.raises(_lib.Se2pError):
        Plan(gpu_ctx, "slp", p, p, 1.0, 1.0, _prm(n, 1.0, 1.0, 1e-8, 0, -1.0))
