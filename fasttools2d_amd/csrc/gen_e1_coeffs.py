"""Generate e1_coeffs.h: the coefficient tables of the library's E1 evaluator.

E1(x) = int_1^inf e^{-xt}/t dt is evaluated in two regimes:

* x < 1:  E1(x) = -gamma - ln x + x A(x), with A(x) = sum_{k>=1} (-1)^{k+1}
  x^{k-1}/(k k!) entire.  A is replaced by its Chebyshev interpolant on [0, 1]
  (mpmath.chebyfit), stored as monomial coefficients for Horner evaluation.
* x >= 1: E1(x) = e^{-x} t g(t), t = 1/x, g(t) = x e^x E1(x) smooth on [0, 1]
  with g(0) = 1.  g is replaced by a rational P(t)/Q(t), Q(0) = 1, fitted by
  Loeb's iteratively reweighted linear least squares on Chebyshev points.

Three accuracy tiers are emitted (full FP64, reduced FP64 for tolerances at or
above 1e-10, FP32).  The maximum relative error of each fit, measured against
mpmath at 50 digits on a dense grid, is written into the header.

Run:  python fasttools2d_amd/csrc/gen_e1_coeffs.py   (needs mpmath; a few minutes)
"""
import os
import sys

import mpmath as mp

mp.mp.dps = 50

TIERS = [
    # name, near degree, far (num, den) degree
    ("full", 12, 11),
    ("fast", 8, 7),
    ("f32", 5, 4),
]


def A(x):
    x = mp.mpf(x)
    if x == 0:
        return mp.mpf(1)
    return (mp.e1(x) + mp.euler + mp.log(x)) / x


def g(t):
    t = mp.mpf(t)
    if t == 0:
        return mp.mpf(1)
    x = 1 / t
    return x * mp.exp(x) * mp.e1(x)


def near_fit(deg):
    poly, err = mp.chebyfit(A, [0, 1], deg + 1, error=True)
    coeffs = list(reversed(poly))  # ascending powers
    return coeffs, err


def far_fit(n, npts=160, iters=7):
    ts = [(1 - mp.cos(mp.pi * (i + mp.mpf(1) / 2) / npts)) / 2 for i in range(npts)]
    gs = [g(t) for t in ts]
    w = [mp.mpf(1)] * npts
    p = q = None
    for _ in range(iters):
        rows, rhs = [], []
        for t, gv, wi in zip(ts, gs, w):
            rows.append([wi * t ** k for k in range(n + 1)] +
                        [-wi * gv * t ** k for k in range(1, n + 1)])
            rhs.append(wi * gv)
        sol = mp.qr_solve(mp.matrix(rows), mp.matrix(rhs))[0]
        p = [sol[k] for k in range(n + 1)]
        q = [mp.mpf(1)] + [sol[n + 1 + k] for k in range(n)]
        w = [1 / abs(mp.polyval(list(reversed(q)), t)) for t in ts]
    return p, q


def far_err(p, q, grid=3000):
    e = mp.mpf(0)
    for i in range(grid + 1):
        t = mp.mpf(i) / grid
        v = mp.polyval(list(reversed(p)), t) / mp.polyval(list(reversed(q)), t)
        e = max(e, abs(v - g(t)) / g(t))
    return e


def near_err(c, grid=2000):
    e = mp.mpf(0)
    for i in range(1, grid + 1):
        x = mp.mpf(i) / grid
        v = -mp.euler - mp.log(x) + x * mp.polyval(list(reversed(c)), x)
        e = max(e, abs(v - mp.e1(x)) / mp.e1(x))
    return e


def fmt(v):
    return mp.nstr(v, 25, min_fixed=0, max_fixed=0)


def main(out_path):
    lines = [
        "// Generated by gen_e1_coeffs.py -- do not edit by hand.",
        "// E1(x) = -gamma - ln x + x * NEAR(x)            for x < 1",
        "// E1(x) = e^{-x} * t * FARP(t) / FARQ(t), t = 1/x  for x >= 1",
        "// Coefficients are in ascending powers.",
        "#pragma once",
        "",
        "#define SE2P_EULER_GAMMA 0.5772156649015328606065120900824024",
        "",
    ]
    for name, dn, df in TIERS:
        c, _ = near_fit(dn)
        ne = near_err(c)
        p, q = far_fit(df)
        fe = far_err(p, q)
        print(f"{name}: near deg {dn} rel err {mp.nstr(ne, 3)}; far ({df},{df}) rel err "
              f"{mp.nstr(fe, 3)}", file=sys.stderr)
        lines.append(f"// tier '{name}': max rel error near {mp.nstr(ne, 3)}, far {mp.nstr(fe, 3)}")
        lines.append(f"#define SE2P_E1_{name.upper()}_NEAR_DEG {dn}")
        lines.append(f"#define SE2P_E1_{name.upper()}_FAR_DEG {df}")
        lines.append(f"#define SE2P_E1_{name.upper()}_NEAR {{ " + ", ".join(fmt(v) for v in c) + " }")
        lines.append(f"#define SE2P_E1_{name.upper()}_FARP {{ " + ", ".join(fmt(v) for v in p) + " }")
        lines.append(f"#define SE2P_E1_{name.upper()}_FARQ {{ " + ", ".join(fmt(v) for v in q) + " }")
        lines.append("")
    with open(out_path, "w") as fh:
        fh.write("\n".join(lines))


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(here, "e1_coeffs.h"))
