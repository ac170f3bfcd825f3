"""Python mirror of the reference's PeriodicEwald interface, running on the GPU.

Same names, argument order and meaning as the MATLAB/mex functions of
annabroms/FastTools2D (paths relative to PeriodicEwald/src/):

* ``mex_stokes_slp_real``   - mex/StokesSLP/mex_stokes_slp_real.cpp:58
* ``mex_stokes_slp_kspace`` - mex/StokesSLP/mex_stokes_slp_kspace.cpp:7
* ``mex_stokes_dlp_real``   - mex/StokesDLP/mex_stokes_dlp_real.cpp:15
* ``mex_stokes_dlp_kspace`` - mex/StokesDLP/mex_stokes_dlp_kspace.cpp:10
* ``StokesSLP_ewald_2p``    - matlab/StokesSLP_ewald_2p.m:1
* ``StokesDLP_ewald_2p``    - matlab/StokesDLP_ewald_2p.m:1

Point, density and normal arrays are 2 x N like the MATLAB matrices; outputs
are 2 x N.  Every call goes through libse2p.so; shape errors raise
``ValueError`` where the mex raised mexErrMsgTxt, and library failures raise
``Se2pError``.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import Options, Params, check, default_context, lib


def _as_2xn(a, name):
    a = np.asarray(a, dtype=np.float64)
    if a.ndim != 2 or a.shape[0] != 2:
        raise ValueError(f"{name} must be a 2xn matrix.")
    # MATLAB's 2 x N column-major memory order is interleaved [x0, y0, x1, y1, ...]
    return np.ascontiguousarray(a.T)


def _vec(a, name):
    a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))
    return a


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _ctx(ctx):
    return (ctx or default_context()).handle


def mex_stokes_slp_real(psrc, ptar, f, xi, nside_x, nside_y, Lx, Ly, ctx=None):
    ps, pt, ff = _as_2xn(psrc, "psrc"), _as_2xn(ptar, "ptar"), _as_2xn(f, "f")
    if ff.shape[0] != ps.shape[0]:
        raise ValueError("psrc and f must be the same size.")
    out = np.zeros((pt.shape[0], 2))
    check(lib().se2p_slp_real(_ctx(ctx), _p(ps), _p(pt), _p(ff), ps.shape[0], pt.shape[0],
                              float(xi), int(nside_x), int(nside_y), float(Lx), float(Ly),
                              _p(out)))
    return out.T.copy()


def mex_stokes_slp_kspace(psrc, ptar, xi, eta, f, Mx, My, Lx, Ly, w, P, ctx=None):
    ps, pt, ff = _as_2xn(psrc, "psrc"), _as_2xn(ptar, "ptar"), _as_2xn(f, "f")
    if ff.shape[0] != ps.shape[0]:
        raise ValueError("psrc and f must be the same size.")
    out = np.zeros((pt.shape[0], 2))
    check(lib().se2p_slp_kspace(_ctx(ctx), _p(ps), _p(pt), float(xi), float(eta), _p(ff),
                                ps.shape[0], pt.shape[0], int(Mx), int(My), float(Lx), float(Ly),
                                float(w), int(P), _p(out)))
    return out.T.copy()


def mex_stokes_dlp_real(psrc, ptar, f, n, xi, nside_x, nside_y, Lx, Ly, ctx=None):
    ps, pt = _as_2xn(psrc, "psrc"), _as_2xn(ptar, "ptar")
    ff, nn = _as_2xn(f, "f"), _as_2xn(n, "n")
    if ff.shape[0] != ps.shape[0]:
        raise ValueError("psrc and f must be the same size.")
    if nn.shape[0] != ps.shape[0]:
        raise ValueError("psrc and n must be the same size.")
    out = np.zeros((pt.shape[0], 2))
    check(lib().se2p_dlp_real(_ctx(ctx), _p(ps), _p(pt), _p(ff), _p(nn), ps.shape[0],
                              pt.shape[0], float(xi), int(nside_x), int(nside_y), float(Lx),
                              float(Ly), _p(out)))
    return out.T.copy()


def mex_stokes_dlp_kspace(psrc, ptar, xi, eta, f, n, Mx, My, Lx, Ly, w, P, ctx=None):
    ps, pt = _as_2xn(psrc, "psrc"), _as_2xn(ptar, "ptar")
    ff, nn = _as_2xn(f, "f"), _as_2xn(n, "n")
    if ff.shape[0] != ps.shape[0]:
        raise ValueError("psrc and f must be the same size.")
    if nn.shape[0] != ps.shape[0]:
        raise ValueError("psrc and n must be the same size.")
    out = np.zeros((pt.shape[0], 2))
    check(lib().se2p_dlp_kspace(_ctx(ctx), _p(ps), _p(pt), float(xi), float(eta), _p(ff),
                                _p(nn), ps.shape[0], pt.shape[0], int(Mx), int(My), float(Lx),
                                float(Ly), float(w), int(P), _p(out)))
    return out.T.copy()


def mex_stokes_slp_pressure_real(psrc, ptar, f, xi, nside_x, nside_y, Lx, Ly, ctx=None):
    """pr = mex_stokes_slp_pressure_real(psrc, ptar, f, xi, nside_x, nside_y, Lx, Ly)
    (mex_stokes_slp_pressure_real.cpp:17-160); 1 x Ntar."""
    ps, pt, ff = _as_2xn(psrc, "psrc"), _as_2xn(ptar, "ptar"), _as_2xn(f, "f")
    if ff.shape[0] != ps.shape[0]:
        raise ValueError("psrc and f must be the same size.")
    out = np.zeros(pt.shape[0])
    check(lib().se2p_slp_pressure_real(_ctx(ctx), _p(ps), _p(pt), _p(ff), ps.shape[0],
                                       pt.shape[0], float(xi), int(nside_x), int(nside_y),
                                       float(Lx), float(Ly), _p(out)))
    return out[None, :].copy()


def mex_stokes_slp_pressure_kspace(psrc, ptar, f, xi, eta, Mx, My, Lx, Ly, w, P, ctx=None):
    """pk = mex_stokes_slp_pressure_kspace(psrc, ptar, f, xi, eta, Mx, My, Lx, Ly, w, P)
    (mex_stokes_slp_pressure_kspace.cpp:7-190, f before xi); 1 x Ntar, no zero mode."""
    ps, pt, ff = _as_2xn(psrc, "psrc"), _as_2xn(ptar, "ptar"), _as_2xn(f, "f")
    if ff.shape[0] != ps.shape[0]:
        raise ValueError("psrc and f must be the same size.")
    out = np.zeros(pt.shape[0])
    check(lib().se2p_slp_pressure_kspace(_ctx(ctx), _p(ps), _p(pt), _p(ff), float(xi),
                                         float(eta), ps.shape[0], pt.shape[0], int(Mx), int(My),
                                         float(Lx), float(Ly), float(w), int(P), _p(out)))
    return out[None, :].copy()


def StokesSLP_pressure_ewald_2p(xsrc, ysrc, xtar, ytar, f1, f2, Lx, Ly, P=24, Nb=None,
                                tol=1e-16, verbose=0, ctx=None):
    """[p, pr, pk, xi] = StokesSLP_pressure_ewald_2p(...) (StokesSLP_pressure_ewald_2p.m:1-138)."""
    xs, ys, xt, yt = (_vec(v, n) for v, n in ((xsrc, "xsrc"), (ysrc, "ysrc"), (xtar, "xtar"),
                                                 (ytar, "ytar")))
    a1, a2 = _vec(f1, "f1"), _vec(f2, "f2")
    Ns, Nt = xs.size, xt.size
    if ys.size != Ns or a1.size != Ns or a2.size != Ns or yt.size != Nt:
        raise ValueError("inconsistent input lengths")
    p, pr, pk = np.zeros(Nt), np.zeros(Nt), np.zeros(Nt)
    xi = ctypes.c_double()
    o = _options(P, Nb, tol, verbose)
    check(lib().se2p_slp_pressure_ewald(_ctx(ctx), _p(xs), _p(ys), _p(xt), _p(yt), _p(a1),
                                        _p(a2), Ns, Nt, float(Lx), float(Ly), ctypes.byref(o),
                                        _p(p), _p(pr), _p(pk), ctypes.byref(xi)))
    return p, pr, pk, xi.value


def mex_stokes_dlp_pressure_real(psrc, ptar, f, n, xi, nside_x, nside_y, Lx, Ly, ctx=None):
    """pr = mex_stokes_dlp_pressure_real(psrc, ptar, f, n, xi, nside_x, nside_y, Lx, Ly)
    (mex_stokes_dlp_pressure_real.cpp:19-63); 1 x Ntar."""
    ps, pt = _as_2xn(psrc, "psrc"), _as_2xn(ptar, "ptar")
    ff, nn = _as_2xn(f, "f"), _as_2xn(n, "n")
    if ff.shape[0] != ps.shape[0]:
        raise ValueError("psrc and f must be the same size.")
    if nn.shape[0] != ps.shape[0]:
        raise ValueError("psrc and n must be the same size.")
    out = np.zeros(pt.shape[0])
    check(lib().se2p_dlp_pressure_real(_ctx(ctx), _p(ps), _p(pt), _p(ff), _p(nn), ps.shape[0],
                                       pt.shape[0], float(xi), int(nside_x), int(nside_y),
                                       float(Lx), float(Ly), _p(out)))
    return out[None, :].copy()


def mex_stokes_dlp_pressure_kspace(psrc, ptar, f, n, xi, eta, Mx, My, Lx, Ly, w, P, ctx=None):
    """pk = mex_stokes_dlp_pressure_kspace(psrc, ptar, f, n, xi, eta, Mx, My, Lx, Ly, w, P)
    (mex_stokes_dlp_pressure_kspace.cpp:10-55); 1 x Ntar, no zero mode."""
    ps, pt = _as_2xn(psrc, "psrc"), _as_2xn(ptar, "ptar")
    ff, nn = _as_2xn(f, "f"), _as_2xn(n, "n")
    if ff.shape[0] != ps.shape[0]:
        raise ValueError("psrc and f must be the same size.")
    if nn.shape[0] != ps.shape[0]:
        raise ValueError("psrc and n must be the same size.")
    out = np.zeros(pt.shape[0])
    check(lib().se2p_dlp_pressure_kspace(_ctx(ctx), _p(ps), _p(pt), _p(ff), _p(nn), float(xi),
                                         float(eta), ps.shape[0], pt.shape[0], int(Mx), int(My),
                                         float(Lx), float(Ly), float(w), int(P), _p(out)))
    return out[None, :].copy()


def StokesDLP_pressure_ewald_2p(xsrc, ysrc, xtar, ytar, n1, n2, f1, f2, Lx, Ly, P=24, Nb=None,
                                tol=1e-16, verbose=0, ctx=None):
    """[p, pr, pk, xi] = StokesDLP_pressure_ewald_2p(...) (StokesDLP_pressure_ewald_2p.m:1-160)."""
    xs, ys, xt, yt = (_vec(v, n) for v, n in ((xsrc, "xsrc"), (ysrc, "ysrc"), (xtar, "xtar"),
                                                 (ytar, "ytar")))
    b1, b2 = _vec(n1, "n1"), _vec(n2, "n2")
    a1, a2 = _vec(f1, "f1"), _vec(f2, "f2")
    Ns, Nt = xs.size, xt.size
    if (ys.size != Ns or a1.size != Ns or a2.size != Ns or b1.size != Ns or b2.size != Ns
            or yt.size != Nt):
        raise ValueError("inconsistent input lengths")
    p, pr, pk = np.zeros(Nt), np.zeros(Nt), np.zeros(Nt)
    xi = ctypes.c_double()
    o = _options(P, Nb, tol, verbose)
    check(lib().se2p_dlp_pressure_ewald(_ctx(ctx), _p(xs), _p(ys), _p(xt), _p(yt), _p(b1),
                                        _p(b2), _p(a1), _p(a2), Ns, Nt, float(Lx), float(Ly),
                                        ctypes.byref(o), _p(p), _p(pr), _p(pk), ctypes.byref(xi)))
    return p, pr, pk, xi.value


def _options(P=24, Nb=None, tol=1e-16, verbose=0):
    o = Options()
    lib().se2p_options_default(ctypes.byref(o))
    o.P = int(P)
    o.Nb = float(Nb) if Nb is not None else 0.0
    o.tol = float(tol)
    o.verbose = int(verbose)
    return o


def ewald_params(kind, Nsrc, Ntar, Q, Lx, Ly, P=24, Nb=None, tol=1e-16):
    """Reference parameter recipe (StokesSLP_ewald_2p.m:84-104 / StokesDLP_ewald_2p.m:104-125;
    each pressure uses its family's recipe, StokesSLP_pressure_ewald_2p.m:83-104 and
    StokesDLP_pressure_ewald_2p.m:111-134)."""
    o = _options(P, Nb, tol)
    prm = Params()
    k = _lib.kind_code(kind)
    check(lib().se2p_params_ref(k, int(Nsrc), int(Ntar), float(Q), float(Lx), float(Ly),
                                ctypes.byref(o), ctypes.byref(prm)))
    return prm.as_dict()


def StokesSLP_ewald_2p(xsrc, ysrc, xtar, ytar, f1, f2, Lx, Ly, P=24, Nb=None, tol=1e-16,
                       verbose=0, ctx=None):
    """[u1, u2, ur, uk, xi] = StokesSLP_ewald_2p(...) (StokesSLP_ewald_2p.m:1-138)."""
    xs, ys, xt, yt = (_vec(v, n) for v, n in ((xsrc, "xsrc"), (ysrc, "ysrc"), (xtar, "xtar"),
                                                 (ytar, "ytar")))
    a1, a2 = _vec(f1, "f1"), _vec(f2, "f2")
    Ns, Nt = xs.size, xt.size
    if ys.size != Ns or a1.size != Ns or a2.size != Ns or yt.size != Nt:
        raise ValueError("inconsistent input lengths")
    u1, u2 = np.zeros(Nt), np.zeros(Nt)
    ur, uk = np.zeros((Nt, 2)), np.zeros((Nt, 2))
    xi = ctypes.c_double()
    o = _options(P, Nb, tol, verbose)
    check(lib().se2p_slp_ewald(_ctx(ctx), _p(xs), _p(ys), _p(xt), _p(yt), _p(a1), _p(a2), Ns, Nt,
                               float(Lx), float(Ly), ctypes.byref(o), _p(u1), _p(u2), _p(ur),
                               _p(uk), ctypes.byref(xi)))
    return u1, u2, ur.T.copy(), uk.T.copy(), xi.value


def StokesDLP_ewald_2p(xsrc, ysrc, xtar, ytar, n1, n2, f1, f2, Lx, Ly, P=24, Nb=None, tol=1e-16,
                       verbose=0, ctx=None):
    """[u1, u2, ur, uk, xi] = StokesDLP_ewald_2p(...) (StokesDLP_ewald_2p.m:1-163).

    The zero mode sum(f.n x_src)/(Lx Ly) is included in uk.  The driver's
    MATLAB-release workaround (1e-60 offsets, StokesDLP_ewald_2p.m:65-80) is
    not reproduced.
    """
    xs, ys, xt, yt = (_vec(v, n) for v, n in ((xsrc, "xsrc"), (ysrc, "ysrc"), (xtar, "xtar"),
                                                 (ytar, "ytar")))
    b1, b2, a1, a2 = _vec(n1, "n1"), _vec(n2, "n2"), _vec(f1, "f1"), _vec(f2, "f2")
    Ns, Nt = xs.size, xt.size
    if any(v.size != Ns for v in (ys, b1, b2, a1, a2)) or yt.size != Nt:
        raise ValueError("inconsistent input lengths")
    u1, u2 = np.zeros(Nt), np.zeros(Nt)
    ur, uk = np.zeros((Nt, 2)), np.zeros((Nt, 2))
    xi = ctypes.c_double()
    o = _options(P, Nb, tol, verbose)
    check(lib().se2p_dlp_ewald(_ctx(ctx), _p(xs), _p(ys), _p(xt), _p(yt), _p(b1), _p(b2), _p(a1),
                               _p(a2), Ns, Nt, float(Lx), float(Ly), ctypes.byref(o), _p(u1),
                               _p(u2), _p(ur), _p(uk), ctypes.byref(xi)))
    return u1, u2, ur.T.copy(), uk.T.copy(), xi.value


# ---------------------------------------------------------------------------
# Stage-level functions of PeriodicEwald/src/mex/common/ewald_tools.h
# ---------------------------------------------------------------------------
def Assign(p, Lx, Ly, nside_x, nside_y, ctx=None):
    """Assign (ewald_tools.cpp:7-80) for one 2 x N point set.

    Returns (particle_offsets, box_offsets, n_in_box) as int32 arrays.
    """
    pp = _as_2xn(p, "p")
    n = pp.shape[0]
    nbox = int(nside_x) * int(nside_y)
    po = np.zeros(max(n, 1), np.int32)
    bo = np.zeros(nbox + 1, np.int32)
    nb = np.zeros(nbox, np.int32)
    check(lib().se2p_assign(_ctx(ctx), _p(pp), n, float(Lx), float(Ly), int(nside_x),
                            int(nside_y), _p(po), _p(bo), _p(nb)))
    return po[:n], bo, nb


def FindClosestNode(p, Lx, Ly, h, P, ctx=None):
    """FindClosestNode (ewald_tools.cpp:86-114) for every column of a 2 x N array.

    Returns (mx, my, px, py).
    """
    pp = _as_2xn(p, "p")
    n = pp.shape[0]
    mx, my = np.zeros(n, np.int32), np.zeros(n, np.int32)
    px, py = np.zeros(n), np.zeros(n)
    check(lib().se2p_find_closest_node(_ctx(ctx), _p(pp), n, float(Lx), float(Ly), float(h),
                                       int(P), _p(mx), _p(my), _p(px), _p(py)))
    return mx, my, px, py


def Spread(psrc, f, Lx, Ly, xi, w, eta, P, Mx, My, ctx=None):
    """Spread (ewald_tools.cpp:120-199): returns (H1, H2) as My x Mx arrays."""
    ps, ff = _as_2xn(psrc, "psrc"), _as_2xn(f, "f")
    H1 = np.zeros((int(Mx), int(My)))
    H2 = np.zeros((int(Mx), int(My)))
    check(lib().se2p_spread(_ctx(ctx), _p(ps), _p(ff), ps.shape[0], float(Lx), float(Ly),
                            float(xi), float(w), float(eta), int(P), int(Mx), int(My), _p(H1),
                            _p(H2)))
    return H1.T.copy(), H2.T.copy()


def Gather(H, ptar, Lx, Ly, xi, w, eta, P, ctx=None):
    """Gather (ewald_tools.cpp:206-271) of one My x Mx real grid at 2 x N targets."""
    Hm = np.asarray(H, dtype=np.float64)
    My, Mx = Hm.shape
    Hc = np.ascontiguousarray(Hm.T)  # column-major My x Mx
    pt = _as_2xn(ptar, "ptar")
    out = np.zeros(pt.shape[0])
    check(lib().se2p_gather(_ctx(ctx), _p(Hc), _p(pt), pt.shape[0], float(Lx), float(Ly),
                            float(xi), float(w), float(eta), int(P), int(Mx), int(My), _p(out)))
    return out


def fft2(Z, inverse=False, ctx=None):
    """MATLAB fft2 / ifft2 of a My x Mx complex array on the GPU."""
    Zm = np.asarray(Z, dtype=np.complex128)
    My, Mx = Zm.shape
    re = np.ascontiguousarray(Zm.real.T)
    im = np.ascontiguousarray(Zm.imag.T)
    check(lib().se2p_fft2(_ctx(ctx), _p(re), _p(im), int(My), int(Mx), 1 if inverse else 0))
    return (re + 1j * im).T.copy()


def expint_e1(x, tier=0, ctx=None):
    """The real-space E1 evaluator on the GPU (tier 0 full FP64, 1 reduced)."""
    xx = np.ascontiguousarray(np.asarray(x, dtype=np.float64).reshape(-1))
    out = np.zeros_like(xx)
    check(lib().se2p_e1(_ctx(ctx), _p(xx), xx.size, int(tier), _p(out)))
    return out
