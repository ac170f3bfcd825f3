"""se2p_eval_device_noalloc: the allocation-free variant of se2p_eval_device,
whose workspace is carved from the library's fixed per-GPU scratch array.

Its results must equal se2p_eval_device's bit for bit (same kernels, same
inputs, fixed-order sums), on two calls in a row, the second smaller than the
first (the scratch keeps the first call's contents), for a Stokeslet and a
stresslet system, with the recipe's support (block spreading) and a narrower
one (strip spreading, whose records take more scratch); sizes beyond the
scratch are rejected with an error."""
import ctypes

import numpy as np
import pytest

import fasttools2d_amd as ft
from fasttools2d_amd import _lib
from fasttools2d_amd.plan import DeviceArray, params_struct
from tests.inputs import circles, normal_density, uniform_points

pytestmark = pytest.mark.gpu


def _fn(name):
    return getattr(_lib.lib(), name)


def _eval(ctx, name, kind, ps, pt, f, nrm, Lx, Ly, prm):
    dps = DeviceArray.from_host(ctx, ps)
    dpt = DeviceArray.from_host(ctx, pt)
    df = DeviceArray.from_host(ctx, f)
    dn = DeviceArray.from_host(ctx, nrm) if nrm is not None else None
    du = DeviceArray(ctx, (len(pt), 2))
    try:
        s = params_struct(prm)
        _lib.check(_fn(name)(ctx.handle, _lib.kind_code(kind), dps.ptr, dpt.ptr, df.ptr,
                             dn.ptr if dn else None, len(ps), len(pt), float(Lx), float(Ly),
                             ctypes.byref(s), du.ptr))
        ctx.sync()
        return du.download()
    finally:
        for a in (dps, dpt, df, dn, du):
            if a is not None:
                a.free()


def _case(kind, n, seed, P=None):
    if kind == "slp":
        Lx, Ly = 1.0, 1.0
        ps = uniform_points(n, Lx, Ly, seed)
        pt = uniform_points(n // 2 + 7, Lx, Ly, seed + 1)
        f, nrm = normal_density(n, seed + 2), None
    else:
        Lx, Ly = 2.0, 1.0
        ps, nrm, f = circles(8, n // 8, Lx, Ly, seed)
        pt = uniform_points(n // 2 + 7, Lx, Ly, seed + 1)
    npts = len(ps) + len(pt)
    prm = ft.ewald_params(kind, len(ps), len(pt), ft.density_q(f[:, 0], f[:, 1]), Lx, Ly,
                          tol=1e-10)
    prm["symmetric"] = 0
    if P is not None:  # a narrower support: strip spreading instead of blocks
        prm["P"] = P
    assert npts > 0
    return ps, pt, f, nrm, Lx, Ly, prm


@pytest.mark.parametrize("kind,P", [("slp", None), ("dlp", None), ("slp", 12), ("dlp", 10)])
def test_noalloc_matches_eval_device_twice(gpu_ctx, kind, P):
    for n, seed in ((120_000, 1), (9_000, 5)):  # the second call smaller
        ps, pt, f, nrm, Lx, Ly, prm = _case(kind, n, seed, P)
        a = _eval(gpu_ctx, "se2p_eval_device_noalloc", kind, ps, pt, f, nrm, Lx, Ly, prm)
        b = _eval(gpu_ctx, "se2p_eval_device", kind, ps, pt, f, nrm, Lx, Ly, prm)
        assert np.all(np.isfinite(a))
        np.testing.assert_array_equal(a, b)


def test_noalloc_rejects_sizes_beyond_scratch(gpu_ctx):
    ps, pt, f, nrm, Lx, Ly, prm = _case("slp", 2000, 9)
    big = dict(prm, Mx=4096, My=4096, nside_x=8, nside_y=8)
    s = params_struct(big)
    # only the host-side size check runs: the pointers are never read
    rc = _fn("se2p_eval_device_noalloc")(gpu_ctx.handle, 0, None, None, None, None, 3_000_000,
                                          3_000_000, 1.0, 1.0, ctypes.byref(s), None)
    assert rc == _lib.SE2P_ERR_UNSUPPORTED
    assert b"scratch" in _lib.lib().se2p_last_error()
