"""Learned parameter selector: a small network that picks (xi, M, P, rc) for a
periodic system, trained on parameter sweeps of the method itself.

The reference's drivers pick the parameters from analytic error estimates
(StokesSLP_ewald_2p.m:84-104,147-202; StokesDLP_ewald_2p.m:105-125,171-225),
which are conservative.  The autotuner (tune.autotune) measures instead, but a
tuning costs dozens of evaluations.  The selector learns the tuner's choice:

  inputs (8):  log10 N, log10 tol, box aspect Lx/Ly, layer (0 Stokeslet,
               1 stresslet), and four statistics of the point density on a
               16 x 16 histogram (workloads.density_moments: coefficient of
               variation, skewness, excess kurtosis, fraction of empty bins);
  outputs (4): log xi, log (Mx/Lx), P, log rc (rc on a log scale: it spans
               two decades);
  network:     8 -> 64 GELU -> 64 GELU -> 4 in FP32, inputs and outputs
               standardized; forward, loss gradient and Adam steps run on the
               GPU in the library's own kernels (se2p_mlp_*).

Training data (sweep): random uniform and clustered configurations, each
tuned by tune.autotune on the GPU; the label is the fastest configuration
whose error against an FP64 solution at tol 1e-13 (the reference recipe,
itself within 1e-14 of the CPU oracle in the tests) meets the tolerance on
the tuned density and on a held-out one.

Selection (select): the network's prediction is rounded to valid parameters
(integer cells, a 2-3-5-smooth grid, even P in [8, 24]) and checked a
posteriori on the system: its result for a random density is compared with
that of the analytic recipe at the same tolerance (two independent Ewald
splits agree to within their errors, the check of the reference's
consistency tests).  When they differ by more than the tolerance the
analytic recipe is used instead.
"""
import json
import math
import os
import time

import numpy as np

from . import _lib
from ._lib import check, lib
from .plan import DeviceArray, Plan
from .tune import _aspect, far32_for, make_params, precision_for, rel_err, smooth_up
from .workloads import clusters, density_moments, normal_density, uniform_points, unit_normals

N_IN, N_HID, N_OUT = 8, 64, 4
N_PARAMS = 8 * 64 + 64 + 64 * 64 + 64 + 4 * 64 + 4
WEIGHTS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "selector_weights.json")
SWEEPS_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "selector_sweeps.json")


# ---------------------------------------------------------------- features
def features(kind, n, tol, Lx, Ly, points):
    k = _lib.kind_code(kind)
    dlp = 1.0 if _lib.has_normals(k) else 0.0
    m = density_moments(points, Lx, Ly)
    return np.array([math.log10(n), math.log10(tol), Lx / Ly, dlp, *m], dtype=np.float64)


def targets(prm, Lx):
    return np.array([math.log(prm["xi"]), math.log(prm["Mx"] / Lx), float(prm["P"]),
                     math.log(Lx / prm["nside_x"])])


def _smooth(m):
    for q in (2, 3, 5):
        while m % q == 0:
            m //= q
    return m == 1


def params_from_outputs(y, kind, n, tol, Lx, Ly, max_grid=4096, symmetric=0):
    """Valid parameters nearest to the network's outputs y (4 values)."""
    A, B = _aspect(Lx, Ly)
    xi = math.exp(y[0])
    rc = math.exp(y[3])
    a = max(1, int(round(Lx / rc / A)))
    P = int(min(24, max(8, 2 * round(y[2] / 2))))
    cap = max_grid // max(A, B)
    want = min(max(P + 2, math.exp(y[1]) * Lx / A), cap)
    try:
        M0 = smooth_up(want, cap)
    except ValueError:  # no smooth size between want and the cap: the largest below
        M0 = max(m for m in range(P + 2, cap + 1, 2) if _smooth(m))
    tier = 1 if tol >= 1e-10 else 0
    prec = precision_for(tol)
    return make_params(xi, a * A, a * B, A * M0, B * M0, P, Lx, tier, symmetric, prec,
                       far32_for(tol) if tier == 1 and prec == 0 else 0.0)


# ---------------------------------------------------------------- network on the GPU
class Network:
    """Packed FP32 parameters and standardization constants."""

    def __init__(self, w, x_mean, x_std, t_mean, t_std):
        self.w = np.asarray(w, dtype=np.float32)
        assert self.w.size == N_PARAMS
        self.x_mean, self.x_std = np.asarray(x_mean), np.asarray(x_std)
        self.t_mean, self.t_std = np.asarray(t_mean), np.asarray(t_std)

    @staticmethod
    def init(seed, x_mean, x_std, t_mean, t_std):
        """He-normal weights for the GELU layers, zero biases."""
        rng = np.random.default_rng(seed)
        w = np.concatenate([rng.standard_normal(64 * 8) * math.sqrt(2 / 8), np.zeros(64),
                            rng.standard_normal(64 * 64) * math.sqrt(2 / 64), np.zeros(64),
                            rng.standard_normal(4 * 64) * math.sqrt(1 / 64), np.zeros(4)])
        return Network(w, x_mean, x_std, t_mean, t_std)

    def to_json(self):
        return {"w": self.w.tolist(), "x_mean": self.x_mean.tolist(),
                "x_std": self.x_std.tolist(), "t_mean": self.t_mean.tolist(),
                "t_std": self.t_std.tolist()}

    @staticmethod
    def from_json(d):
        return Network(d["w"], d["x_mean"], d["x_std"], d["t_mean"], d["t_std"])

    def predict(self, ctx, X):
        """Outputs (unstandardized) for feature rows X, on the GPU."""
        X = np.atleast_2d(np.asarray(X, dtype=np.float64))
        xs = ((X - self.x_mean) / self.x_std).astype(np.float32)
        y = forward(ctx, self.w, xs)
        return y.astype(np.float64) * self.t_std + self.t_mean


def _dev(ctx, a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    p = ctx.alloc(max(4, a.nbytes))
    if a.nbytes:
        ctx.h2d(p, a)
    return p


def _host(ctx, p, shape):
    out = np.empty(shape, dtype=np.float32)
    if out.nbytes:
        ctx.d2h(out, p)
    return out


def forward(ctx, w, x):
    """Network outputs (n x 4) for standardized inputs x (n x 8), FP32, GPU."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    n = x.shape[0]
    dw, dx, dy = _dev(ctx, w), _dev(ctx, x), ctx.alloc(max(4, n * 16))
    try:
        check(lib().se2p_mlp_forward(ctx.handle, dw, dx, n, dy))
        ctx.sync()
        return _host(ctx, dy, (n, 4))
    finally:
        for p in (dw, dx, dy):
            ctx.free(p)


def loss_grad(ctx, w, x, t):
    """(loss, gradient) of the mean squared error on the GPU."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    n = x.shape[0]
    dw, dx, dt = _dev(ctx, w), _dev(ctx, x), _dev(ctx, t)
    dg, dl = ctx.alloc(4 * N_PARAMS), ctx.alloc(4)
    try:
        check(lib().se2p_mlp_loss_grad(ctx.handle, dw, dx, dt, n, dg, dl))
        ctx.sync()
        return float(_host(ctx, dl, (1,))[0]), _host(ctx, dg, (N_PARAMS,))
    finally:
        for p in (dw, dx, dt, dg, dl):
            ctx.free(p)


def train(ctx, X, T, steps=4000, lr=3e-3, seed=0, log=None):
    """Full-batch Adam on the GPU: the loss gradient and the update are the
    library's kernels, the parameters stay on the device.  Returns the network
    and the loss history (every 100 steps)."""
    X = np.asarray(X, dtype=np.float64)
    T = np.asarray(T, dtype=np.float64)
    x_mean, x_std = X.mean(0), X.std(0)
    x_std[x_std < 1e-6] = 1.0
    t_mean, t_std = T.mean(0), T.std(0)
    t_std[t_std < 1e-6] = 1.0
    net = Network.init(seed, x_mean, x_std, t_mean, t_std)
    xs = ((X - x_mean) / x_std).astype(np.float32)
    ts = ((T - t_mean) / t_std).astype(np.float32)
    n = xs.shape[0]
    dw, dx, dt = _dev(ctx, net.w), _dev(ctx, xs), _dev(ctx, ts)
    dg, dl = ctx.alloc(4 * N_PARAMS), ctx.alloc(4)
    dm, dv = _dev(ctx, np.zeros(N_PARAMS)), _dev(ctx, np.zeros(N_PARAMS))
    hist = []
    try:
        for step in range(1, steps + 1):
            # cosine decay of the step size to a tenth
            a = lr * (0.1 + 0.9 * 0.5 * (1 + math.cos(math.pi * step / steps)))
            check(lib().se2p_mlp_loss_grad(ctx.handle, dw, dx, dt, n, dg, dl))
            check(lib().se2p_mlp_adam(ctx.handle, dw, dg, dm, dv, step, a, 0.9, 0.999, 1e-8))
            if step % 100 == 0 or step == 1:
                ctx.sync()
                hist.append((step, float(_host(ctx, dl, (1,))[0])))
                if log:
                    log(f"selector: step {step} loss {hist[-1][1]:.4e}")
        ctx.sync()
        net.w = _host(ctx, dw, (N_PARAMS,))
    finally:
        for p in (dw, dx, dt, dg, dl, dm, dv):
            ctx.free(p)
    return net, hist


def load(path=WEIGHTS_PATH):
    with open(path) as fh:
        return Network.from_json(json.load(fh))


# ---------------------------------------------------------------- configurations
def random_configs(n, seed, nmin=1e3, nmax=3e5, tmin=1e-10, tmax=1e-4, slp_frac=0.7):
    """Random uniform and clustered periodic systems (sources = targets)."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        kind = "slp" if rng.random() < slp_frac else "dlp"
        Lx, Ly = float(rng.choice([1.0, 1.5, 2.0, 3.0, 4.0])), 1.0
        N = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        tol = float(10 ** rng.uniform(np.log10(tmin), np.log10(tmax)))
        s = int(rng.integers(1 << 30))
        if rng.random() < 0.5:
            p = uniform_points(N, Lx, Ly, s)
            dist = "uniform"
        else:
            ncl = int(rng.integers(4, 65))
            sig = float(rng.uniform(0.01, 0.05))
            p = clusters(N, Lx, Ly, s, ncl=ncl, sigma=sig)
            dist = f"clusters {ncl} sigma {sig:.3f}"
        p = np.ascontiguousarray(p)
        out.append(dict(kind=kind, N=N, tol=tol, Lx=Lx, Ly=Ly, points=p, dist=dist,
                        nrm=unit_normals(N, s + 1) if kind == "dlp" else None,
                        f=normal_density(N, s + 2)))
    return out


def sweep(ctx, configs, log=None, reps=2):
    """Tune every configuration (tune.autotune) and record its features, label
    and the timed sweep."""
    from .tune import autotune
    rows = []
    for i, c in enumerate(configs):
        t0 = time.perf_counter()
        prm, rep = autotune(ctx, c["kind"], c["points"], c["points"], c["f"], c["Lx"], c["Ly"],
                            c["tol"], nrm=c["nrm"], reps=reps, final_rounds=2,
                            nside_scales=(0.75, 1.0, 1.4, 1.9, 2.6, 3.4, 4.7))
        rows.append({"kind": c["kind"], "N": c["N"], "tol": c["tol"], "Lx": c["Lx"],
                     "Ly": c["Ly"], "dist": c["dist"],
                     "features": features(c["kind"], c["N"], c["tol"], c["Lx"], c["Ly"],
                                          c["points"]).tolist(),
                     "label": {k: prm[k] for k in ("xi", "nside_x", "nside_y", "Mx", "My", "P",
                                                    "precision")},
                     "ms": rep["ms"], "err": rep["err"], "err_heldout": rep["err_heldout"],
                     "runs": len(rep["sweep"]),
                     "sweep": [[r[0], r[1], round(r[2], 4), r[3]] for r in rep["sweep"]]})
        if log:
            log(f"sweep {i}: {c['kind']} N {c['N']} tol {c['tol']:.1e} {c['dist']} -> nside "
                f"{prm['nside_x']} M {prm['Mx']} P {prm['P']} {rep['ms']:.3f} ms "
                f"({time.perf_counter() - t0:.1f} s)")
    return rows


# ---------------------------------------------------------------- selection
def analytic_params(kind, n, tol, Lx, Ly, f, max_grid=4096):
    """The reference recipe at tol (its points-per-cell option raised until the
    grid fits max_grid per side), the selector's fallback."""
    from .periodic_ewald import density_q
    from .tune import recipe_params
    prm = recipe_params(kind, n, n, density_q(f[:, 0], f[:, 1]), Lx, Ly, tol, max_grid=max_grid)
    prm["e1_tier"] = 1 if tol >= 1e-10 else 0
    prm["precision"] = precision_for(tol)
    return prm


def safer(prm, Lx, Ly, tol):
    """The same split with more accuracy in both parts: xi (so xi rc) and the grid
    raised by a tenth, the support by two nodes."""
    A, B = _aspect(Lx, Ly)
    P = min(24, prm["P"] + 2)
    cap = 4096 // max(A, B)
    M0 = min(cap, smooth_up(min(cap, 1.1 * prm["Mx"] / A), cap))
    return make_params(1.1 * prm["xi"], prm["nside_x"], prm["nside_y"], A * M0, B * M0, P, Lx,
                       prm["e1_tier"], prm["symmetric"], prm["precision"], prm.get("far32", 0.0))


def select(ctx, net, kind, points, Lx, Ly, tol, nrm=None, check_seed=99, symmetric=1,
           margin=0.5, draws=4):
    """(params, report) for a system whose targets are its sources.  The
    network's choice is checked a posteriori against the analytic recipe (whose
    own error is far below tol) for a random density g.  A relative error
    ||E f|| / ||S f|| (E the parameters' error operator, S the sum) varies
    with the density mostly through ||S f||, which the few lowest Fourier modes
    of f dominate; ||E f|| / ||f|| is nearly the same for every random f.  So
    the check bounds the error of any random density by
    (||E g|| / ||g||) / min_f (||S f|| / ||f||), the minimum over g and `draws`
    more N(0,1) densities applied with the same plan, and accepts it if that is
    at most margin * tol.  If it fails, a safer variant of the same split
    (safer()) is checked the same way, and if that fails too the analytic
    recipe is used."""
    n = len(points)
    k = _lib.kind_code(kind)
    sym = symmetric if k in (_lib.SE2P_SLP, _lib.SE2P_DLP) else 0
    y = net.predict(ctx, features(kind, n, tol, Lx, Ly, points))[0]
    prm = params_from_outputs(y, kind, n, tol, Lx, Ly, symmetric=sym)
    rng = np.random.default_rng(check_seed)
    dens = [rng.standard_normal((n, 2)) for _ in range(1 + draws)]
    g = dens[0]
    ref = analytic_params(kind, n, tol, Lx, Ly, g)
    ref["symmetric"] = sym

    def run(q, ds):
        pl = Plan(ctx, kind, points, points, Lx, Ly, q, nrm=nrm)
        try:
            return [pl.apply(d) for d in ds]
        finally:
            pl.close()

    u_ref = run(ref, [g])[0]

    def bound(q):
        us = run(q, dens)
        e_abs = np.linalg.norm(us[0] - u_ref) / np.linalg.norm(g)
        s_min = min([np.linalg.norm(u_ref) / np.linalg.norm(g)] +
                    [np.linalg.norm(u) / np.linalg.norm(d) for u, d in zip(us[1:], dens[1:])])
        return float(e_abs / max(s_min, 1e-300)), rel_err(us[0], u_ref)

    e, e_g = bound(prm)
    rep = {"check_err": e, "check_err_g": e_g, "predicted": prm, "fallback": False,
           "safer": False}
    if e <= margin * tol:
        return prm, rep
    try:
        prm2 = safer(prm, Lx, Ly, tol)
    except ValueError:
        prm2 = None
    if prm2 is not None:
        e2, _ = bound(prm2)
        rep["check_err_safer"] = e2
        if e2 <= margin * tol:
            rep["safer"] = True
            return prm2, rep
    rep["fallback"] = True
    return ref, rep


def time_params(ctx, kind, points, Lx, Ly, prm, f, nrm=None, reps=10):
    """(seconds per apply, output) of a plan at prm on device-resident data."""
    pl = Plan(ctx, kind, points, points, Lx, Ly, prm, nrm=nrm)
    df = DeviceArray.from_host(ctx, f)
    du = DeviceArray(ctx, (len(points), 2))
    try:
        t0 = time.perf_counter()
        pl.apply_device(df.ptr, du.ptr)
        ctx.sync()
        if time.perf_counter() - t0 > 0.05:  # slow configurations: one timed apply
            reps, rounds = 1, 1
        else:
            rounds = 3
        ts = []
        for _ in range(rounds):
            t0 = time.perf_counter()
            for _ in range(reps):
                pl.apply_device(df.ptr, du.ptr)
            ctx.sync()
            ts.append((time.perf_counter() - t0) / reps)
        return float(np.median(ts)), du.download()
    finally:
        pl.close()
        df.free()
        du.free()


def evaluate(ctx, net, configs, log=None):
    """Per configuration: time and error (against an FP64 solution at tol 1e-13)
    of the selector's parameters (after its check), tune.rule_params and the
    analytic recipe."""
    from .tune import rule_params
    rows = []
    for i, c in enumerate(configs):
        kind, p, n, tol, Lx, Ly = c["kind"], c["points"], c["N"], c["tol"], c["Lx"], c["Ly"]
        sym = 1 if kind in ("slp", "dlp") else 0
        f = c["f"]
        acc = analytic_params(kind, n, 1e-13, Lx, Ly, f)
        acc["precision"] = 0
        _, u_ref = time_params(ctx, kind, p, Lx, Ly, acc, f, c["nrm"], reps=1)
        sel, srep = select(ctx, net, kind, p, Lx, Ly, tol, nrm=c["nrm"], symmetric=sym)
        cand = {"selector": sel,
                "rule": rule_params(kind, n, n, Lx, Ly, tol, symmetric=sym, precision=None),
                "analytic": dict(analytic_params(kind, n, tol, Lx, Ly, f), symmetric=sym)}
        row = {"kind": kind, "N": n, "tol": tol, "Lx": Lx, "dist": c["dist"],
               "fallback": srep["fallback"], "safer": srep["safer"],
               "check_err": srep["check_err"]}
        for name, q in cand.items():
            t, u = time_params(ctx, kind, p, Lx, Ly, q, f, c["nrm"])
            row[name] = {"ms": t * 1e3, "err": rel_err(u, u_ref),
                         "nside": q["nside_x"], "M": q["Mx"], "P": q["P"]}
        rows.append(row)
        if log:
            log(f"eval {i}: {kind} N {n} tol {tol:.1e} " + " ".join(
                f"{k} {row[k]['ms']:.3f} ms err/tol {row[k]['err'] / tol:.2f}"
                for k in cand) + (" (fallback)" if row["fallback"] else ""))
    return rows


def summarize(rows):
    """Geometric-mean speedups of the selector over the rule and the analytic
    recipe, and the worst error / tolerance of each."""
    def gm(a):
        return float(np.exp(np.mean(np.log(a))))
    out = {"configs": len(rows), "fallbacks": int(sum(r["fallback"] for r in rows)),
           "safer_variants": int(sum(r.get("safer", False) for r in rows))}
    for k in ("selector", "rule", "analytic"):
        out[f"worst_err_over_tol_{k}"] = max(r[k]["err"] / r["tol"] for r in rows)
        out[f"ms_total_{k}"] = sum(r[k]["ms"] for r in rows)
    out["speedup_vs_rule_gmean"] = gm([r["rule"]["ms"] / r["selector"]["ms"] for r in rows])
    out["speedup_vs_analytic_gmean"] = gm([r["analytic"]["ms"] / r["selector"]["ms"]
                                           for r in rows])
    return out


# ---------------------------------------------------------------- command line
def _main():
    import argparse
    import sys
    ap = argparse.ArgumentParser(description="selector sweeps, training and evaluation")
    ap.add_argument("what", choices=["sweep", "train", "evaluate"])
    ap.add_argument("--n", type=int, default=400)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--nmax", type=float, default=3e5)
    ap.add_argument("--sweeps", default=SWEEPS_PATH)
    ap.add_argument("--weights", default=WEIGHTS_PATH)
    ap.add_argument("--out", default=None)
    ap.add_argument("--steps", type=int, default=4000)
    args = ap.parse_args()
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    ctx = _lib.Context(0)
    if args.what == "sweep":
        rows = sweep(ctx, random_configs(args.n, args.seed, nmax=args.nmax), log=log)
        with open(args.out or args.sweeps, "w") as fh:
            json.dump({"seed": args.seed, "rows": rows}, fh)
    elif args.what == "train":
        rows = []
        for path in args.sweeps.split(","):
            with open(path) as fh:
                rows += json.load(fh)["rows"]
        # labels whose choice did not hold on the held-out density are left out
        rows = [r for r in rows if r["err_heldout"] is not None and r["err_heldout"] <= r["tol"]]
        X = np.array([r["features"] for r in rows])
        T = np.array([targets(r["label"], r["Lx"]) for r in rows])
        net, hist = train(ctx, X, T, steps=args.steps, log=log)
        with open(args.out or args.weights, "w") as fh:
            json.dump(dict(net.to_json(), loss=hist, samples=len(rows)), fh)
    else:
        net = load(args.weights)
        rows = evaluate(ctx, net, random_configs(args.n, args.seed, nmin=1e3, nmax=args.nmax),
                        log=log)
        res = {"summary": summarize(rows), "rows": rows}
        print(json.dumps(res["summary"]))
        if args.out:
            with open(args.out, "w") as fh:
                json.dump(res, fh)


if __name__ == "__main__":
    _main()
