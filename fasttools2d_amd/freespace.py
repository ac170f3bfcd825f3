"""Free-space (non-periodic) Stokes layer potentials: the reference's FMM mex
functions stokesSLPfmm and stokesDLPfmm (FMM/src/StokesSLP/stokesSLPfmm.F,
FMM/src/StokesDLP/stokesDLPfmm.F), with the same names, arguments and outputs.
libse2p.so evaluates them on the GPU: by direct summation in FP64 (exact to
rounding, O(Ns Nt) work), or for the Stokeslet by a free-space spectral Ewald
sum at the FMM tolerance of iprec (O(N log N); the context's free-space method
chooses, see Context.set_free_method).  free_slp_plan sets up the same sum for
repeated evaluations on a fixed geometry."""
import ctypes

import numpy as np

from ._lib import Params, check, default_context, lib


def _col(a, name):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def stokesSLPfmm(f1, f2, xsrc, ysrc, xtar, ytar, iequal=0, iprec=5, ctx=None):
    """[u1, u2] = stokesSLPfmm(f1, f2, xsrc, ysrc, xtar, ytar, iequal, iprec)
    (stokesSLPfmm.F:10-149): u = sum_s (-1/2 log rho^2 f + (r.f) r/rho^2)/(4 pi);
    iequal = 1 when the targets are the sources (own terms left out).  iprec
    (-2..5) sets the tolerance, 0.5 10^-k for k = 0, 1, 2, 3, 6, 9, 12, 15
    (cfmm2dpart.f:351-361), which the spectral Ewald sum meets when the
    context's method takes it (iprec <= 4)."""
    f1, f2, xs, ys = (_col(v, n) for v, n in ((f1, "f1"), (f2, "f2"), (xsrc, "xsrc"),
                                              (ysrc, "ysrc")))
    xt, yt = _col(xtar, "xtar"), _col(ytar, "ytar")
    Ns, Nt = xs.size, xt.size
    if f1.size != Ns or f2.size != Ns or ys.size != Ns or yt.size != Nt:
        raise ValueError("inconsistent input lengths")
    u1, u2 = np.zeros(Nt), np.zeros(Nt)
    check(lib().se2p_free_slp((ctx or default_context()).handle, _p(f1), _p(f2), _p(xs), _p(ys),
                              Ns, _p(xt), _p(yt), Nt, int(iequal), int(iprec), _p(u1), _p(u2)))
    return u1, u2


def stokesDLPfmm(f1, f2, xsrc, ysrc, n1, n2, ctx=None):
    """[u1, u2] = stokesDLPfmm(f1, f2, xsrc, ysrc, n1, n2) (stokesDLPfmm.F,
    stokesDLP.f): at the sources, u = sum_{s != t} (r.n)(r.f) r/rho^4 / pi."""
    f1, f2, xs, ys, a, b = (_col(v, n) for v, n in ((f1, "f1"), (f2, "f2"), (xsrc, "xsrc"),
                                                    (ysrc, "ysrc"), (n1, "n1"), (n2, "n2")))
    N = xs.size
    if any(v.size != N for v in (f1, f2, ys, a, b)):
        raise ValueError("inconsistent input lengths")
    u1, u2 = np.zeros(N), np.zeros(N)
    check(lib().se2p_free_dlp((ctx or default_context()).handle, _p(f1), _p(f2), _p(xs), _p(ys),
                              _p(a), _p(b), N, _p(u1), _p(u2)))
    return u1, u2


def ewald_params(Nsrc, Ntar, wx, wy, tol):
    """(params dict, Lx, Ly) of a free-space spectral Ewald Stokeslet sum over
    points inside a wx x wy box (se2p_free_params): the points go in centred
    on the box's centre, in a periodic box Lx x Ly, with params["free_R"] set."""
    prm = Params()
    Lx, Ly = ctypes.c_double(), ctypes.c_double()
    check(lib().se2p_free_params(int(Nsrc), int(Ntar), float(wx), float(wy), float(tol),
                                 ctypes.byref(prm), ctypes.byref(Lx), ctypes.byref(Ly)))
    return prm.as_dict(), Lx.value, Ly.value


def free_slp_plan(ctx, psrc, ptar, tol, symmetric=1):
    """A plan (fasttools2d_amd.plan.Plan) for repeated free-space Stokeslet sums
    u(x_t) = sum_s G(x_t - x_s) f_s on fixed sources psrc and targets ptar (N x 2
    arrays; pass the same array for both to leave each source's own term out)
    at relative tolerance tol: the points are centred in the box of
    ewald_params.  Returns (plan, centre); plan.apply(f) gives u at ptar."""
    from .plan import Plan
    ps = np.ascontiguousarray(psrc, dtype=np.float64)
    same = ptar is psrc
    pt = ps if same else np.ascontiguousarray(ptar, dtype=np.float64)
    allp = ps if same else np.concatenate([ps, pt])
    lo, hi = allp.min(0), allp.max(0)
    c = 0.5 * (lo + hi)
    prm, Lx, Ly = ewald_params(len(ps), len(pt), hi[0] - lo[0], hi[1] - lo[1], tol)
    prm["symmetric"] = int(symmetric) if same else 0
    ps_c = np.ascontiguousarray(ps - c)
    pt_c = ps_c if same else np.ascontiguousarray(pt - c)
    return Plan(ctx, "slp", ps_c, pt_c, Lx, Ly, prm), c
