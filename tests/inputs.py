"""Seeded synthetic inputs shared by the tests (the workloads of the benchmark
configurations, fasttools2d_amd.workloads) and the oracle's reference sums."""
import math

import numpy as np

from fasttools2d_amd.workloads import (circles, clusters, normal_density,  # noqa: F401
                                       serving_systems, uniform_points, unit_normals)


def smooth_up(n):
    """Smallest even n' >= n with prime factors in {2, 3, 5} (the oracle's FFT
    is a mixed-radix recursion, slow on large prime factors)."""
    n = max(4, int(math.ceil(n)))
    while True:
        m = n
        for p in (2, 3, 5):
            while m % p == 0:
                m //= p
        if m == 1 and n % 2 == 0:
            return n
        n += 1


def oracle_se(oracle, kind, ps, pt, f, nrm, Lx, Ly, tol=1e-13, P=24):
    """The oracle's full spectral Ewald sum (real + k space, + the stresslet's
    zero mode) of kind 'slp' or 'dlp' at the targets pt, all sources, with the
    reference recipe's parameters at `tol` and the grid raised to the next
    2-3-5-smooth size (a finer grid only lowers the k-space error)."""
    npts = len(ps) + len(pt)
    prm = oracle.params(kind, npts, oracle.density_q(f), Lx, Ly, P, oracle.default_Nb(npts),
                        tol)
    A, B = prm["A"], prm["B"]
    M0 = smooth_up(prm["Mx"] // A)
    Mx, My = A * M0, B * M0
    w = P * Lx / Mx / 2
    eta = (2 * prm["xi"] * w / (0.95 * math.sqrt(math.pi * P))) ** 2
    xi, nsx, nsy = prm["xi"], prm["nside_x"], prm["nside_y"]
    if kind == "slp":
        return (oracle.slp_real(ps, pt, f, xi, nsx, nsy, Lx, Ly) +
                oracle.slp_kspace(ps, pt, xi, eta, f, Mx, My, Lx, Ly, w, P))
    return (oracle.dlp_real(ps, pt, f, nrm, xi, nsx, nsy, Lx, Ly) +
            oracle.dlp_kspace(ps, pt, xi, eta, f, nrm, Mx, My, Lx, Ly, w, P) +
            oracle.dlp_zero_mode(ps, f, nrm, Lx, Ly))


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
