"""The autotuner (fasttools2d_amd/tune.py) on a small system: its parameters meet
the tolerance against an independent accurate solution, and it takes
pair-symmetric sums exactly when the targets are the sources (the same array)."""
import numpy as np
import pytest

import fasttools2d_amd as ft
from fasttools2d_amd.plan import Plan
from fasttools2d_amd.tune import autotune, fft_alternatives, fft_stages, rel_err
from tests.inputs import normal_density, uniform_points



def _solve(ctx, p, pt, f, prm):
    plan = Plan(ctx, "slp", p, pt, 1.0, 1.0, prm)
    try:
        return plan.apply(f)
    finally:
        plan.close()


@pytest.mark.gpu
@pytest.mark.parametrize("same", [True, False])
def test_autotune_meets_tolerance(gpu_ctx, same):
    N, tol = 20000, 1e-8
    p = uniform_points(N, 1.0, 1.0, seed=5)
    pt = p if same else p.copy()
    f = normal_density(N, seed=6)
    prm, rep = autotune(gpu_ctx, "slp", p, pt, f, 1.0, 1.0, tol, reps=2)
    assert prm["symmetric"] == (1 if same else 0)
    assert rep["err"] <= 0.8 * tol
    # an accurate solution from the reference recipe with other constants
    ref = ft.ewald_params("slp", N, N, 2 * N + 1.0, 1.0, 1.0, P=24, tol=1e-14)
    u_ref = _solve(gpu_ctx, p, p.copy(), f, ref)
    assert rel_err(_solve(gpu_ctx, p, pt, f, prm), u_ref) <= tol


def test_fft_grid_alternatives():
    # fewer stages only, each alternative shorter than the one before
    assert fft_stages(2000) == 5 and fft_stages(2048) == 4 and fft_stages(2160) == 3
    assert fft_alternatives(2000) == [2048, 2160]
    assert fft_alternatives(2160) == []
    assert fft_stages(2002) is None  # 7 and 11 are no radices
