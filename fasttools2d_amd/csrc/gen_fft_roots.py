"""Generate fft_roots.h: cos and sin of 2 pi m / R for the radices the line
FFT runs as composite in-register butterflies (R = R1 R2 split into an R2-point
and an R1-point DFT with the twiddles W_R^(r1 k2) in between).  Values from
mpmath at 40 digits, printed to 21 significant digits (correctly rounded when
the compiler reads them as doubles).

Run:  python fasttools2d_amd/csrc/gen_fft_roots.py   (needs mpmath; a second)
"""
import os
import sys

import mpmath as mp

mp.mp.dps = 40

RADICES = (6, 9, 10, 12, 15, 16, 20, 25)


def main(out_path):
    fmt = lambda v: mp.nstr(v, 21, min_fixed=-30, max_fixed=30) if v != 0 else "0.0"  # noqa: E731
    lines = [
        "// Generated by gen_fft_roots.py -- do not edit by hand.",
        "// FftRoots<R>::c[m] = cos(2 pi m / R), FftRoots<R>::s[m] = sin(2 pi m / R).",
        "#pragma once",
        "",
        "namespace se2p {",
        "",
        "template <int R>",
        "struct FftRoots;",
    ]
    for R in RADICES:
        c = [mp.cos(2 * mp.pi * m / R) for m in range(R)]
        s = [mp.sin(2 * mp.pi * m / R) for m in range(R)]
        # exact zeros where the angle is a multiple of pi/2
        c = [mp.mpf(0) if abs(v) < mp.mpf(10) ** -35 else v for v in c]
        s = [mp.mpf(0) if abs(v) < mp.mpf(10) ** -35 else v for v in s]
        lines += [
            "template <>",
            f"struct FftRoots<{R}> {{",
            f"    static constexpr double c[{R}] = {{" + ", ".join(fmt(v) for v in c) + "};",
            f"    static constexpr double s[{R}] = {{" + ", ".join(fmt(v) for v in s) + "};",
            "};",
        ]
    lines += ["", "}  // namespace se2p", ""]
    with open(out_path, "w") as fh:
        fh.write("\n".join(lines))


if __name__ == "__main__":
    here = os.path.dirname(os.path.abspath(__file__))
    main(sys.argv[1] if len(sys.argv) > 1 else os.path.join(here, "fft_roots.h"))
