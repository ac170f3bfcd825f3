"""Writes the committed golden vectors under tests/golden/.

e1_mpmath.json: E1(x) at 50 significant digits from mpmath for x spanning
[1e-12, 700] (log-spaced), the regime boundaries x = 1 and x = 16 and their
neighbours.  These pin the oracle's E1 and the GPU evaluator independently of
any code in this repository.

Run: python tests/golden/make_golden.py  (needs mpmath; seconds)
"""
import json
import os

import mpmath as mp
import numpy as np

mp.mp.dps = 50
HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    xs = list(np.logspace(-12, np.log10(700.0), 400))
    for c in (1.0, 16.0, 0.5, 2.0, 4.0, 8.0, 32.0, 64.0):
        xs += [c, float(np.nextafter(c, 0)), float(np.nextafter(c, np.inf))]
    xs = sorted(set(float(x) for x in xs))
    rows = [[repr(x), mp.nstr(mp.e1(mp.mpf(x)), 40)] for x in xs]
    with open(os.path.join(HERE, "e1_mpmath.json"), "w") as fh:
        json.dump({"source": "mpmath.e1 at 50 digits", "rows": rows}, fh, indent=0)


if __name__ == "__main__":
    main()
