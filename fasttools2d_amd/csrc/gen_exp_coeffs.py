"""Generate exp_coeffs.h: the polynomial of the library's reduced-range e^{-x}.

e^{-x} (0 <= x < 700) is evaluated by Cody-Waite argument reduction
(W. J. Cody and W. Waite, Software Manual for the Elementary Functions, 1980):
n = rint(x / ln 2), r = n ln 2 - x with ln 2 split into a 32-bit head and a
tail, so |r| <= ln(2)/2 and e^{-x} = 2^{-n} e^r.  e^r on [-ln 2/2, ln 2/2] is
replaced by its Chebyshev interpolant (mpmath.chebyfit), stored as monomial
coefficients for Horner evaluation.  The maximum relative error of the
polynomial, measured against mpmath at 50 digits on a dense grid, is written
into the header.

Run:  python fasttools2d_amd/csrc/gen_exp_coeffs.py   (needs mpmath; seconds)
"""
import os
import sys

import mpmath as mp

mp.mp.dps = 50

DEGREE = 10


def split_ln2():
    ln2 = mp.log(2)
    m, e = mp.frexp(ln2)
    head = mp.ldexp(mp.floor(mp.ldexp(m, 32)), e - 32)  # 32 significant bits
    return head, ln2 - head


def main(out_path):
    half = mp.log(2) / 2
    poly, _ = mp.chebyfit(mp.exp, [-half, half], DEGREE + 1, error=True)
    coeffs = list(reversed(poly))  # ascending powers
    err = mp.mpf(0)
    grid = 4000
    for i in range(grid + 1):
        r = -half + 2 * half * i / grid
        err = max(err, abs(mp.polyval(poly, r) / mp.exp(r) - 1))
    head, tail = split_ln2()
    print(f"exp: degree {DEGREE}, max rel error {mp.nstr(err, 3)}", file=sys.stderr)
    fmt = lambda v: mp.nstr(v, 25, min_fixed=0, max_fixed=0)  # noqa: E731
    lines = [
        "// Generated by gen_exp_coeffs.py -- do not edit by hand.",
        "// e^{-x} = 2^{-n} EXP(r), n = rint(x log2 e), r = n (LN2_HI + LN2_LO) - x.",
        "// Coefficients are in ascending powers.",
        "#pragma once",
        "",
        f"// max rel error of the polynomial on [-ln2/2, ln2/2]: {mp.nstr(err, 3)}",
        f"#define SE2P_EXP_DEG {DEGREE}",
        f"#define SE2P_EXP_POLY {{ " + ", ".join(fmt(v) for v in coeffs) + " }",
        f"#define SE2P_LOG2E {fmt(1 / mp.log(2))}",
        f"#define SE2P_LN2_HI {fmt(head)}",
        f"#define SE2P_LN2_LO {fmt(tail)}",
        "",
    ]
    with open(out_path, "w") as fh:
        fh.write("\n".join(lines))


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(here, "exp_coeffs.h"))
