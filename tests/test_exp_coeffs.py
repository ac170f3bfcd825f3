"""The reduced-range e^{-x} of the real-space kernel's fast E1 tier
(fasttools2d_amd/csrc/exp_coeffs.h, e1_eval.h exp_neg): the polynomial and the
Cody-Waite reduction restated in numpy, checked against numpy's exp and the
Taylor series of e^r (a dropped or mistyped coefficient fails)."""
import math
import os
import re

import numpy as np

HDR = os.path.join(os.path.dirname(__file__), "..", "fasttools2d_amd", "csrc", "exp_coeffs.h")


def _defs():
    txt = open(HDR).read()
    poly = re.search(r"#define SE2P_EXP_POLY \{([^}]*)\}", txt).group(1)
    c = np.array([float(v) for v in poly.split(",")])
    get = lambda name: float(re.search(rf"#define {name} (\S+)", txt).group(1))  # noqa: E731
    return c, get("SE2P_LOG2E"), get("SE2P_LN2_HI"), get("SE2P_LN2_LO")


def test_polynomial_is_close_to_taylor_and_exp():
    c, _, _, _ = _defs()
    assert len(c) == 11
    # coefficients of e^r: 1/k! (the interpolant moves the high ones slightly)
    for k, ck in enumerate(c):
        assert abs(ck * math.factorial(k) - 1) < (1e-10 if k < 5 else 1e-2)
    r = np.linspace(-math.log(2) / 2, math.log(2) / 2, 20001)
    p = np.polynomial.polynomial.polyval(r, c)
    assert np.max(np.abs(p / np.exp(r) - 1)) < 2e-15


def test_reduced_exp_matches_numpy():
    c, log2e, hi, lo = _defs()
    assert abs(hi + lo - math.log(2)) < 1e-16
    x = np.concatenate([np.linspace(0, 40, 200001), np.linspace(40, 700, 2001)])
    n = np.rint(x * log2e)
    r = (n * hi - x) + n * lo  # n hi is exact: hi has 32 significant bits
    val = np.ldexp(np.polynomial.polynomial.polyval(r, c), -n.astype(int))
    assert np.max(np.abs(val / np.exp(-x) - 1)) < 5e-15
