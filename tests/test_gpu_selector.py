"""The learned parameter selector (fasttools2d_amd/selector.py, with the
committed weights it was trained to) on held-out random configurations it
never saw: every chosen parameter set meets its tolerance against an FP64
solution at tol 1e-13, and it is faster than the analytic recipe it falls
back to (the method without the network)."""
import numpy as np
import pytest

from fasttools2d_amd import selector as S

pytestmark = pytest.mark.gpu


def test_selector_weights_trained_on_sweeps():
    import json
    w = json.load(open(S.WEIGHTS_PATH))
    assert len(w["w"]) == S.N_PARAMS and w["samples"] >= 500
    assert w["loss"][-1][1] < w["loss"][0][1] / 20  # the GPU training converged
    sweeps = json.load(open(S.SWEEPS_PATH))["rows"]
    assert len(sweeps) >= w["samples"] and sum(r["runs"] for r in sweeps) > 5000


def test_selector_on_held_out_configurations(gpu_ctx):
    net = S.load()
    rows = S.evaluate(gpu_ctx, net, S.random_configs(16, 4242, nmin=1e3, nmax=2e5))
    summ = S.summarize(rows)
    print(summ)
    assert summ["worst_err_over_tol_selector"] <= 1.0
    assert summ["speedup_vs_analytic_gmean"] > 1.5
    assert summ["fallbacks"] < len(rows)


def test_selector_falls_back_when_the_check_fails(gpu_ctx):
    """A network whose outputs are far off (all weights zero: the outputs are the
    training means, whatever the system) still yields parameters that meet the
    tolerance, through the safer variant or the analytic recipe."""
    net = S.load()
    net.w = np.zeros_like(net.w)
    c = S.random_configs(1, 77, nmin=2e4, nmax=3e4, tmin=1e-9, tmax=1e-9)[0]
    prm, rep = S.select(gpu_ctx, net, c["kind"], c["points"], c["Lx"], c["Ly"], c["tol"],
                        nrm=c["nrm"])
    ok = rep["check_err"] <= 0.5 * c["tol"]
    assert ok or rep["safer"] or rep["fallback"]
    if rep["fallback"]:
        assert prm["nside_x"] == S.analytic_params(c["kind"], c["N"], c["tol"], c["Lx"], c["Ly"],
                                                   c["f"])["nside_x"]
