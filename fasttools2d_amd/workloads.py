"""Seeded synthetic point sets and densities of the benchmark workloads.

Used by the benchmark, the tests and the sweeps that train the parameter
selector.  Points are returned as (N, 2) arrays wrapped into the box
[-Lx/2, Lx/2) x [-Ly/2, Ly/2) (the reference's drivers wrap the same way,
StokesSLP_ewald_2p.m:75-78).
"""
import numpy as np


def wrap(p, Lx, Ly):
    """mod(x + L/2, L) - L/2 per coordinate."""
    p = np.asarray(p, dtype=np.float64)
    L = np.array([Lx, Ly])
    return np.ascontiguousarray(np.mod(p + L / 2, L) - L / 2)


def uniform_points(n, Lx, Ly, seed):
    rng = np.random.default_rng(seed)
    p = rng.random((n, 2)) * [Lx, Ly]
    return p - [Lx / 2, Ly / 2]


def normal_density(n, seed):
    return np.random.default_rng(seed).standard_normal((n, 2))


def unit_normals(n, seed):
    th = np.random.default_rng(seed).random(n) * 2 * np.pi
    return np.ascontiguousarray(np.stack([np.cos(th), np.sin(th)], 1))


def circles(ncirc, nper, Lx, Ly, seed, rmin=0.02, rmax=0.08, gap=0.01):
    """Non-overlapping circles (radii uniform in [rmin, rmax], centres uniform,
    periodic distance at least r1 + r2 + gap) with nper equispaced nodes each,
    outward unit normals, trapezoid weights 2 pi r / nper and N(0,1) densities
    times the weights: the stresslet workload.  Returns (points, normals, f)."""
    rng = np.random.default_rng(seed)
    cs, rs = [], []
    while len(cs) < ncirc:
        r = rng.uniform(rmin, rmax)
        c = rng.random(2) * [Lx, Ly] - [Lx / 2, Ly / 2]
        ok = True
        for c2, r2 in zip(cs, rs):
            d = c - c2
            d -= np.round(d / [Lx, Ly]) * [Lx, Ly]
            if np.hypot(*d) < r + r2 + gap:
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
    return np.ascontiguousarray(pts), np.ascontiguousarray(np.concatenate(nrm)), f


def clusters(n, Lx, Ly, seed, ncl=64, sigma=0.01):
    """A mixture of ncl Gaussian clusters (centres uniform in the box, standard
    deviation sigma per coordinate, equal weights), wrapped periodically: the
    strongly nonuniform workload."""
    rng = np.random.default_rng(seed)
    c = rng.random((ncl, 2)) * [Lx, Ly] - [Lx / 2, Ly / 2]
    k = rng.integers(0, ncl, n)
    p = c[k] + sigma * rng.standard_normal((n, 2))
    return wrap(p, Lx, Ly)


def density_moments(p, Lx, Ly, bins=16):
    """Four statistics of the point density: the coefficient of variation,
    skewness and excess kurtosis of the counts on a bins x bins grid, and the
    fraction of empty bins.  Zero for a perfectly even spread."""
    p = np.asarray(p)
    ix = np.clip(((p[:, 0] / Lx + 0.5) * bins).astype(np.int64), 0, bins - 1)
    iy = np.clip(((p[:, 1] / Ly + 0.5) * bins).astype(np.int64), 0, bins - 1)
    c = np.bincount(iy * bins + ix, minlength=bins * bins).astype(np.float64)
    m = c.mean()
    s = c.std()
    if m <= 0 or s <= 0:
        return np.zeros(4)
    z = (c - m) / s
    return np.array([s / m, (z ** 3).mean(), (z ** 4).mean() - 3.0, (c == 0).mean()])


def serving_systems(nsys, seed=11, nmin=500, nmax=200_000, aspects=None):
    """The batched serving workload: box aspect in 1..4 (half steps: the
    parameter recipe needs a rational aspect, and its numerator multiplies the
    grid), N log-uniform in [nmin, nmax], tolerance log-uniform in [1e-10,
    1e-4], uniform points that are both sources and targets.  Returns a list of
    dicts (psrc, ptar the same array, Lx, Ly, tol)."""
    rng = np.random.default_rng(seed)
    if aspects is None:
        aspects = np.arange(2, 9) / 2.0
    out = []
    for _ in range(nsys):
        Lx, Ly = float(rng.choice(aspects)), 1.0
        n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        tol = float(10 ** rng.uniform(-10, -4))
        p = np.ascontiguousarray(rng.random((n, 2)) * [Lx, Ly] - [Lx / 2, Ly / 2])
        out.append(dict(kind="slp", psrc=p, ptar=p, nrm=None, Lx=Lx, Ly=Ly, tol=tol))
    return out
