"""Seeded synthetic inputs shared by the tests (the workloads of the benchmark
configurations at test sizes)."""
import numpy as np


def uniform_points(n, Lx, Ly, seed):
    rng = np.random.default_rng(seed)
    p = rng.random((n, 2)) * [Lx, Ly]
    return p - [Lx / 2, Ly / 2]


def normal_density(n, seed):
    return np.random.default_rng(seed).standard_normal((n, 2))


def circles(ncirc, nper, Lx, Ly, seed, rmin=0.02, rmax=0.08):
    """Non-overlapping circles with outward normals, trapezoid weights, N(0,1)
    densities (the stresslet workload)."""
    rng = np.random.default_rng(seed)
    cs, rs = [], []
    while len(cs) < ncirc:
        r = rng.uniform(rmin, rmax)
        c = rng.random(2) * [Lx, Ly] - [Lx / 2, Ly / 2]
        ok = True
        for c2, r2 in zip(cs, rs):
            d = c - c2
            d -= np.round(d / [Lx, Ly]) * [Lx, Ly]
            if np.hypot(*d) < r + r2 + 0.01:
                ok = False
                break
        if ok:
            cs.append(c)
            rs.append(r)
    th = 2 * np.pi * np.arange(nper) / nper
    pts, nrm, wts = [], [], []
    for c, r in zip(cs, rs):
        nn = np.stack([np.cos(th), np.sin(th)], 1)
        pts.append(c + r * nn)
        nrm.append(nn)
        wts.append(np.full(nper, 2 * np.pi * r / nper))
    pts = np.concatenate(pts)
    pts = (pts + [Lx / 2, Ly / 2]) % [Lx, Ly] - [Lx / 2, Ly / 2]
    f = rng.standard_normal((len(pts), 2)) * np.concatenate(wts)[:, None]
    return pts, np.concatenate(nrm), f
