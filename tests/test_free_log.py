"""The library's natural log (fasttools2d_amd/csrc/e1_eval.h, log_pos; used by
the free-space sums and the fast tier's E1 near 1/xi): the reduction x = m 2^k,
m in [sqrt(1/2), sqrt(2)), and the atanh series, restated in numpy and checked
against numpy's log over the range the kernels see (squared distances from
1e-300 to 1e300, dense near 1, powers of two exact).  The device code itself is
pinned by the free-space and real-space parity tests."""
import numpy as np

LN2_HI = 6.931471803691238164901733e-1
LN2_LO = 1.908214929270587816144266e-10


def log_pos(x):
    m, k = np.frexp(x)
    low = m < 0.70710678118654752440
    m = np.where(low, 2 * m, m)
    k = np.where(low, k - 1, k).astype(float)
    s = (m - 1.0) * (1.0 / (m + 1.0))
    z = s * s
    p = np.full_like(x, 1.0 / 21)
    for d in (19, 17, 15, 13, 11, 9, 7, 5, 3):
        p = p * z + 1.0 / d
    lm = 2.0 * (s * z * p + s)
    return k * LN2_HI + (k * LN2_LO + lm)


def test_log_matches_numpy():
    x = np.concatenate([10.0 ** np.linspace(-300, 300, 200001),
                        1 + np.linspace(-0.3, 0.3, 200001),
                        np.ldexp(1.0, np.arange(-1000, 1000))])
    ref = np.log(x)
    got = log_pos(x)
    big = np.abs(ref) > 1e-3
    ulps = np.abs(got[big] - ref[big]) / np.spacing(np.abs(ref[big]))
    assert ulps.max() <= 2.0
    # near 1 the result is small: absolute error within a few ulps of x - 1
    assert np.max(np.abs(got[~big] - ref[~big])) < 1e-18
    assert np.all(got[x == 1.0] == 0.0)


def test_series_tail_is_negligible():
    # |s| <= (sqrt 2 - 1)/(sqrt 2 + 1): the first dropped term s^22/23 against s
    s = (np.sqrt(2) - 1) / (np.sqrt(2) + 1)
    assert s ** 22 / 23 < 3e-18
