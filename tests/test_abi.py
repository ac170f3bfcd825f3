"""The C-ABI library loads, exports what include/se2p.h declares, and its
host-only functions behave (no GPU needed)."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "se2p.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(se2p_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from fasttools2d_amd import _lib
    L = _lib.lib()
    declared = _declared()
    assert len(declared) >= 28
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r" T (se2p_\w+)", out))
    missing = [d for d in declared if d not in exported]
    assert not missing, missing
    for d in declared:
        getattr(L, d)
    assert set(_lib.EXPORTS) == set(declared)


def test_library_is_gfx950_code_object():
    from fasttools2d_amd import _lib
    blob = open(_lib.LIB_PATH, "rb").read()
    assert b"gfx950" in blob


def test_no_gpu_context_is_an_error_not_a_crash():
    from fasttools2d_amd import _lib
    if _lib.device_count() > 0:
        pytest.skip("a GPU is visible")
    h = ctypes.c_void_p()
    rc = _lib.lib().se2p_ctx_create(0, ctypes.byref(h))
    assert rc == _lib.SE2P_ERR_NODEV
    assert b"no HIP device" in _lib.lib().se2p_last_error()


def test_null_context_rejected():
    from fasttools2d_amd import _lib
    L = _lib.lib()
    rc = L.se2p_slp_real(None, None, None, None, 0, 0, 1.0, 1, 1, 1.0, 1.0, None)
    assert rc == _lib.SE2P_ERR_ARG


@pytest.mark.parametrize("kind,N,Lx,Ly,tol,P,Nb", [
    ("slp", 1000, 1.0, 1.0, 1e-16, 24, None),
    ("slp", 1_000_000, 1.0, 1.0, 1e-8, 24, None),
    ("slp", 20000, 1.0, 2.0, 1e-10, 16, 9),
    ("dlp", 40000, 2.0, 1.0, 1e-10, 24, None),
    ("dlp", 40000, 2.0, 1.0, 1e-6, 24, None),
    ("slp", 4_000_000, 1.0, 1.0, 1e-10, 24, None),
    ("dlp", 3, 1.0, 1.0, 1e-12, 24, None),
])
def test_parameter_recipe_matches_oracle(oracle, kind, N, Lx, Ly, tol, P, Nb):
    from fasttools2d_amd import ewald_params
    Q = 2.0 * N + 1
    got = ewald_params(kind, N, N, Q, Lx, Ly, P=P, Nb=Nb, tol=tol)
    npts = 2 * N
    ref = oracle.params(kind, npts, Q, Lx, Ly, P, Nb if Nb else oracle.default_Nb(npts, P), tol)
    for k in ("nside_x", "nside_y", "Mx", "My", "kinf"):
        assert got[k] == ref[k], k
    for k in ("xi", "rc", "w", "eta"):
        assert abs(got[k] - ref[k]) <= 1e-14 * abs(ref[k]), k
    assert got["converged"] == int(ref["converged"])
    assert got["e1_tier"] == 0


@pytest.mark.parametrize("kind", [0, 1])
def test_parameter_recipe_clears_the_other_fields(kind):
    """se2p_params_ref writes every field: values a caller's struct held before
    (here a free-space radius and an FP32 far-pair threshold) do not survive."""
    from fasttools2d_amd import _lib
    L = _lib.lib()
    prm = _lib.Params()
    prm.free_R, prm.far32, prm.symmetric, prm.precision = 3.0, 4.6, 1, 1
    rc = L.se2p_params_ref(kind, 1000, 1000, 2001.0, 2.0, 1.0, None, ctypes.byref(prm))
    assert rc == _lib.SE2P_OK
    assert (prm.free_R, prm.far32, prm.symmetric, prm.precision) == (0.0, 0.0, 0, 0)
    assert prm.Mx > 0 and prm.xi > 0
