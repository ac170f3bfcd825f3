"""The roots of unity the line FFT's composite radices use
(fasttools2d_amd/csrc/fft_roots.h, written by gen_fft_roots.py): every entry
against numpy's cos/sin of 2 pi m / R, and the table sizes against the radices
the FFT kernel instantiates."""
import math
import os
import re

import numpy as np
import pytest

mp = pytest.importorskip("mpmath")
mp.mp.dps = 30

HDR = os.path.join(os.path.dirname(__file__), "..", "fasttools2d_amd", "csrc", "fft_roots.h")


def _tables():
    txt = open(HDR).read()
    out = {}
    pat = (r"struct FftRoots<(\d+)> \{\s*static constexpr double c\[\d+\] = \{([^}]*)\};"
           r"\s*static constexpr double s\[\d+\] = \{([^}]*)\};")
    for R, c, s in re.findall(pat, txt):
        out[int(R)] = (np.array([float(v) for v in c.split(",")]),
                       np.array([float(v) for v in s.split(",")]))
    return out


def test_roots_match_cos_sin():
    tabs = _tables()
    for R in (6, 9, 10, 12, 15, 16):
        assert R in tabs
        c, s = tabs[R]
        m = np.arange(R)
        assert len(c) == R and len(s) == R
        # correctly rounded: within half an ulp of values at most 1 in size
        cr = np.array([float(mp.cos(2 * mp.pi * k / R)) for k in m])
        sr = np.array([float(mp.sin(2 * mp.pi * k / R)) for k in m])
        assert np.max(np.abs(c - cr)) <= 1.2e-16
        assert np.max(np.abs(s - sr)) <= 1.2e-16
        # and the same to a few ulps without mpmath (numpy rounds 2 pi m / R first)
        assert np.max(np.abs(c - np.cos(2 * np.pi * m / R))) < 1e-15
        # exact values where they exist
        assert c[0] == 1.0 and s[0] == 0.0
        if R % 4 == 0:
            assert c[R // 4] == 0.0 and s[R // 4] == 1.0
        if R % 2 == 0:
            assert c[R // 2] == -1.0 and s[R // 2] == 0.0


def test_split_dft_identity():
    """The index split the composite butterflies use, r = r1 + R1 r2 and
    k = k2 + R2 k1: X[k] = sum_r1 W_R1^(r1 k1) W_R^(r1 k2) sum_r2 W_R2^(r2 k2)
    x[r1 + R1 r2], checked against numpy's FFT for every split in use."""
    rng = np.random.default_rng(5)
    for R1, R2 in ((2, 3), (3, 3), (2, 5), (4, 3), (3, 5), (4, 4)):
        R = R1 * R2
        x = rng.standard_normal(R) + 1j * rng.standard_normal(R)
        X = np.zeros(R, complex)
        for k1 in range(R1):
            for k2 in range(R2):
                acc = 0
                for r1 in range(R1):
                    inner = sum(np.exp(-2j * math.pi * r2 * k2 / R2) * x[r1 + R1 * r2]
                                for r2 in range(R2))
                    acc += np.exp(-2j * math.pi * r1 * k1 / R1) * np.exp(
                        -2j * math.pi * r1 * k2 / R) * inner
                X[k2 + R2 * k1] = acc
        assert np.max(np.abs(X - np.fft.fft(x))) < 1e-12
