"""Free-space (non-periodic) Stokes layer potentials: the reference's FMM mex
functions stokesSLPfmm and stokesDLPfmm (FMM/src/StokesSLP/stokesSLPfmm.F,
FMM/src/StokesDLP/stokesDLPfmm.F), with the same names, arguments and outputs.
libse2p.so evaluates the sums directly on the GPU in FP64, so the results are
exact to rounding instead of to the FMM's tolerance (and the work is O(Ns Nt))."""
import ctypes

import numpy as np

from ._lib import check, default_context, lib


def _col(a, name):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def stokesSLPfmm(f1, f2, xsrc, ysrc, xtar, ytar, iequal=0, iprec=5, ctx=None):
    """[u1, u2] = stokesSLPfmm(f1, f2, xsrc, ysrc, xtar, ytar, iequal, iprec)
    (stokesSLPfmm.F:10-149): u = sum_s (-1/2 log rho^2 f + (r.f) r/rho^2)/(4 pi);
    iequal = 1 when the targets are the sources (own terms left out)."""
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
