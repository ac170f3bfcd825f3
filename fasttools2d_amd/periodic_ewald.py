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


def density_q(f1, f2):
    """Q = sum(sum(f.^2)) + 1 of the MATLAB drivers, f = [f1'; f2'] (2 x N;
    StokesSLP_ewald_2p.m:82,87): each column's f1(i)^2 + f2(i)^2, then those
    sums added in index order (a sequential sum, as the library's C++ drivers
    add them)."""
    a1 = np.asarray(f1, dtype=np.float64).ravel()
    a2 = np.asarray(f2, dtype=np.float64).ravel()
    if a1.size == 0:
        return 1.0
    return float(np.cumsum(a1 * a1 + a2 * a2)[-1] + 1.0)


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


# ---------------------------------------------------------------------------
# Derived quantities: velocity gradient, pressure gradient, vorticity, stress.
# Tensors come back as 4 x Ntar in the mex order [d/dx u1, d/dx u2, d/dy u1,
# d/dy u2] (mex_stokes_slp_gradient_real.cpp:199-206).
# ---------------------------------------------------------------------------
def _derived(kind, part, ps, pt, ff, nn, a, ctx):
    """One derived mex call on (N, 2) arrays; returns (Ntar, nc)."""
    k = _lib.kind_code(kind)
    nc = _lib.out_comps(k)
    out = np.zeros((pt.shape[0], nc))
    layer, q = kind.split("_", 1)
    fn = getattr(lib(), f"se2p_{layer}_{q}_{part}")
    Ns, Nt = ps.shape[0], pt.shape[0]
    pn = [_p(nn)] if nn is not None else []
    if part == "real":
        xi, nsx, nsy, Lx, Ly = a
        check(fn(_ctx(ctx), _p(ps), _p(pt), _p(ff), *pn, Ns, Nt, float(xi), int(nsx), int(nsy),
                 float(Lx), float(Ly), _p(out)))
    else:
        xi, eta, Mx, My, Lx, Ly, w, P = a
        tail = [Ns, Nt, int(Mx), int(My), float(Lx), float(Ly), float(w), int(P), _p(out)]
        if q in ("pressure_grad",) or (layer == "dlp" and q == "vorticity"):
            # f (and n) before xi, eta: mex_stokes_*_pressure_grad_kspace.cpp,
            # mex_stokes_dlp_vorticity_kspace.cpp
            check(fn(_ctx(ctx), _p(ps), _p(pt), _p(ff), *pn, float(xi), float(eta), *tail))
        else:
            check(fn(_ctx(ctx), _p(ps), _p(pt), float(xi), float(eta), _p(ff), *pn, *tail))
    return out


def _derived_real_mex(kind, psrc, ptar, f, n, xi, nside_x, nside_y, Lx, Ly, ctx):
    ps, pt, ff = _as_2xn(psrc, "psrc"), _as_2xn(ptar, "ptar"), _as_2xn(f, "f")
    nn = _as_2xn(n, "n") if n is not None else None
    if ff.shape[0] != ps.shape[0]:
        raise ValueError("psrc and f must be the same size.")
    if nn is not None and nn.shape[0] != ps.shape[0]:
        raise ValueError("psrc and n must be the same size.")
    return _derived(kind, "real", ps, pt, ff, nn, (xi, nside_x, nside_y, Lx, Ly), ctx).T.copy()


def _derived_kspace_mex(kind, psrc, ptar, f, n, xi, eta, Mx, My, Lx, Ly, w, P, ctx):
    ps, pt, ff = _as_2xn(psrc, "psrc"), _as_2xn(ptar, "ptar"), _as_2xn(f, "f")
    nn = _as_2xn(n, "n") if n is not None else None
    if ff.shape[0] != ps.shape[0]:
        raise ValueError("psrc and f must be the same size.")
    if nn is not None and nn.shape[0] != ps.shape[0]:
        raise ValueError("psrc and n must be the same size.")
    return _derived(kind, "kspace", ps, pt, ff, nn, (xi, eta, Mx, My, Lx, Ly, w, P),
                    ctx).T.copy()


def mex_stokes_slp_gradient_real(psrc, ptar, f, xi, nside_x, nside_y, Lx, Ly, ctx=None):
    """mex_stokes_slp_gradient_real.cpp:9-217; 4 x Ntar."""
    return _derived_real_mex("slp_gradient", psrc, ptar, f, None, xi, nside_x, nside_y, Lx, Ly,
                             ctx)


def mex_stokes_slp_gradient_kspace(psrc, ptar, xi, eta, f, Mx, My, Lx, Ly, w, P, ctx=None):
    """mex_stokes_slp_gradient_kspace.cpp:3-288; 4 x Ntar."""
    return _derived_kspace_mex("slp_gradient", psrc, ptar, f, None, xi, eta, Mx, My, Lx, Ly, w,
                               P, ctx)


def mex_stokes_slp_pressure_grad_real(psrc, ptar, f, xi, nside_x, nside_y, Lx, Ly, ctx=None):
    """mex_stokes_slp_pressure_grad_real.cpp:17-186; 2 x Ntar."""
    return _derived_real_mex("slp_pressure_grad", psrc, ptar, f, None, xi, nside_x, nside_y, Lx,
                             Ly, ctx)


def mex_stokes_slp_pressure_grad_kspace(psrc, ptar, f, xi, eta, Mx, My, Lx, Ly, w, P,
                                        ctx=None):
    """mex_stokes_slp_pressure_grad_kspace.cpp:7-195 (f before xi); 2 x Ntar."""
    return _derived_kspace_mex("slp_pressure_grad", psrc, ptar, f, None, xi, eta, Mx, My, Lx,
                               Ly, w, P, ctx)


def mex_stokes_slp_vorticity_real(psrc, ptar, f, xi, nside_x, nside_y, Lx, Ly, ctx=None):
    """mex_stokes_slp_vorticity_real.cpp:9-187; 1 x Ntar."""
    return _derived_real_mex("slp_vorticity", psrc, ptar, f, None, xi, nside_x, nside_y, Lx, Ly,
                             ctx)


def mex_stokes_slp_vorticity_kspace(psrc, ptar, xi, eta, f, Mx, My, Lx, Ly, w, P, ctx=None):
    """mex_stokes_slp_vorticity_kspace.cpp:3-203; 1 x Ntar."""
    return _derived_kspace_mex("slp_vorticity", psrc, ptar, f, None, xi, eta, Mx, My, Lx, Ly, w,
                               P, ctx)


def mex_stokes_slp_stress_real(psrc, ptar, f, xi, nside_x, nside_y, Lx, Ly, ctx=None):
    """mex_stokes_slp_stress_real.cpp:9-217; 4 x Ntar."""
    return _derived_real_mex("slp_stress", psrc, ptar, f, None, xi, nside_x, nside_y, Lx, Ly,
                             ctx)


def mex_stokes_slp_stress_kspace(psrc, ptar, xi, eta, f, Mx, My, Lx, Ly, w, P, ctx=None):
    """mex_stokes_slp_stress_kspace.cpp:3-293; 4 x Ntar."""
    return _derived_kspace_mex("slp_stress", psrc, ptar, f, None, xi, eta, Mx, My, Lx, Ly, w, P,
                               ctx)


def mex_stokes_dlp_gradient_real(psrc, ptar, f, n, xi, nside_x, nside_y, Lx, Ly, ctx=None):
    """mex_stokes_dlp_gradient_real.cpp:9-240; 4 x Ntar."""
    return _derived_real_mex("dlp_gradient", psrc, ptar, f, n, xi, nside_x, nside_y, Lx, Ly, ctx)


def mex_stokes_dlp_gradient_kspace(psrc, ptar, xi, eta, f, n, Mx, My, Lx, Ly, w, P, ctx=None):
    """mex_stokes_dlp_gradient_kspace.cpp:3-345; 4 x Ntar."""
    return _derived_kspace_mex("dlp_gradient", psrc, ptar, f, n, xi, eta, Mx, My, Lx, Ly, w, P,
                               ctx)


def mex_stokes_dlp_pressure_grad_real(psrc, ptar, f, n, xi, nside_x, nside_y, Lx, Ly, ctx=None):
    """mex_stokes_dlp_pressure_grad_real.cpp:17-207; 2 x Ntar."""
    return _derived_real_mex("dlp_pressure_grad", psrc, ptar, f, n, xi, nside_x, nside_y, Lx, Ly,
                             ctx)


def mex_stokes_dlp_pressure_grad_kspace(psrc, ptar, f, n, xi, eta, Mx, My, Lx, Ly, w, P,
                                        ctx=None):
    """mex_stokes_dlp_pressure_grad_kspace.cpp:10-269; 2 x Ntar."""
    return _derived_kspace_mex("dlp_pressure_grad", psrc, ptar, f, n, xi, eta, Mx, My, Lx, Ly,
                               w, P, ctx)


def mex_stokes_dlp_vorticity_real(psrc, ptar, f, n, xi, nside_x, nside_y, Lx, Ly, ctx=None):
    """mex_stokes_dlp_vorticity_real.cpp:19-221; 1 x Ntar."""
    return _derived_real_mex("dlp_vorticity", psrc, ptar, f, n, xi, nside_x, nside_y, Lx, Ly,
                             ctx)


def mex_stokes_dlp_vorticity_kspace(psrc, ptar, f, n, xi, eta, Mx, My, Lx, Ly, w, P, ctx=None):
    """mex_stokes_dlp_vorticity_kspace.cpp:3-252 (f, n before xi); 1 x Ntar."""
    return _derived_kspace_mex("dlp_vorticity", psrc, ptar, f, n, xi, eta, Mx, My, Lx, Ly, w, P,
                               ctx)


def mex_stokes_dlp_stress_real(psrc, ptar, f, n, xi, nside_x, nside_y, Lx, Ly, ctx=None):
    """mex_stokes_dlp_stress_real.cpp:9-248; 4 x Ntar."""
    return _derived_real_mex("dlp_stress", psrc, ptar, f, n, xi, nside_x, nside_y, Lx, Ly, ctx)


def mex_stokes_dlp_stress_kspace(psrc, ptar, xi, eta, f, n, Mx, My, Lx, Ly, w, P, ctx=None):
    """mex_stokes_dlp_stress_kspace.cpp:3-340; 4 x Ntar."""
    return _derived_kspace_mex("dlp_stress", psrc, ptar, f, n, xi, eta, Mx, My, Lx, Ly, w, P,
                               ctx)


# ---------------------------------------------------------------------------
# Drivers of the derived quantities (Stokes{SLP,DLP}_{gradient,pressure_grad,
# vorticity,stress}_ewald_2p.m): wrap, the recipe of each driver, the two mex
# parts, then each driver's projection, zero mode, sign and self terms.
# ---------------------------------------------------------------------------
def _mwrap(v, L):
    """mod(v + L/2, L) - L/2 with MATLAB's mod (StokesSLP_ewald_2p.m:75-78)."""
    a = v + L / 2
    return (a - np.floor(a / L) * L) - L / 2


def _coincident(xs, ys, xt, yt):
    """(source, target) index pairs with identical raw coordinates: the drivers'
    ismember(zsrc, ztar) (e.g. StokesDLP_gradient_ewald_2p.m:66-72).  The
    reference adds the self term of the k-th coincident source to the k-th
    target, which is the coincident target when the targets are the sources;
    this pairs each source with the targets at its position."""
    where = {}
    for j, key in enumerate(zip(xt.tolist(), yt.tolist())):
        where.setdefault(key, []).append(j)
    return [(i, j) for i, key in enumerate(zip(xs.tolist(), ys.tolist()))
            for j in where.get(key, ())]


def _derived_driver(kind, xsrc, ysrc, xtar, ytar, f1, f2, n1, n2, Lx, Ly, P, Nb, tol, ctx):
    xs, ys, xt, yt = (_vec(v, nm) for v, nm in ((xsrc, "xsrc"), (ysrc, "ysrc"), (xtar, "xtar"),
                                                 (ytar, "ytar")))
    a1, a2 = _vec(f1, "f1"), _vec(f2, "f2")
    Ns, Nt = xs.size, xt.size
    if ys.size != Ns or a1.size != Ns or a2.size != Ns or yt.size != Nt:
        raise ValueError("inconsistent input lengths")
    nn = None
    if n1 is not None:
        b1, b2 = _vec(n1, "n1"), _vec(n2, "n2")
        if b1.size != Ns or b2.size != Ns:
            raise ValueError("inconsistent input lengths")
        nn = np.ascontiguousarray(np.stack([b1, b2], 1))
    pairs = _coincident(xs, ys, xt, yt)
    ps = np.ascontiguousarray(np.stack([_mwrap(xs, Lx), _mwrap(ys, Ly)], 1))
    pt = np.ascontiguousarray(np.stack([_mwrap(xt, Lx), _mwrap(yt, Ly)], 1))
    ff = np.ascontiguousarray(np.stack([a1, a2], 1))
    prm = ewald_params(kind, Ns, Nt, density_q(a1, a2), Lx, Ly, P=P, Nb=Nb, tol=tol)
    R = _derived(kind, "real", ps, pt, ff, nn,
                 (prm["xi"], prm["nside_x"], prm["nside_y"], Lx, Ly), ctx)
    K = _derived(kind, "kspace", ps, pt, ff, nn,
                 (prm["xi"], prm["eta"], prm["Mx"], prm["My"], Lx, Ly, prm["w"], prm["P"]), ctx)
    return ps, ff, nn, pairs, prm, R, K


def _project(T, b1, b2, order=(0, 2, 1, 3)):
    """[T[a] b1 + T[b] b2; T[c] b1 + T[d] b2] as 2 x Ntar."""
    a, b, c, d = order
    return np.stack([T[:, a] * b1 + T[:, b] * b2, T[:, c] * b1 + T[:, d] * b2])


def StokesSLP_gradient_ewald_2p(xsrc, ysrc, xtar, ytar, f1, f2, b1, b2, Lx, Ly, P=24, Nb=None,
                                tol=1e-16, verbose=0, ctx=None):
    """[u1, u2, ur, uk, xi] (StokesSLP_gradient_ewald_2p.m:1-160, projection :133-135,148-150): the velocity
    gradient applied to the per-target vector b."""
    _, _, _, _, prm, R, K = _derived_driver("slp_gradient", xsrc, ysrc, xtar, ytar, f1, f2,
                                            None, None, Lx, Ly, P, Nb, tol, ctx)
    bb1, bb2 = _vec(b1, "b1"), _vec(b2, "b2")
    ur, uk = _project(R, bb1, bb2), _project(K, bb1, bb2)
    u = ur + uk
    return u[0].copy(), u[1].copy(), ur, uk, prm["xi"]


def StokesSLP_pressure_grad_ewald_2p(xsrc, ysrc, xtar, ytar, f1, f2, Lx, Ly, P=24, Nb=None,
                                     tol=1e-16, verbose=0, ctx=None):
    """[p, pr, pk, xi] (StokesSLP_pressure_grad_ewald_2p.m:1-158): p = -(pr + pk +
    pself), pself = xi^2 f/(2 pi) at targets that are sources (:152-156)."""
    _, ff, _, pairs, prm, R, K = _derived_driver("slp_pressure_grad", xsrc, ysrc, xtar, ytar, f1,
                                                 f2, None, None, Lx, Ly, P, Nb, tol, ctx)
    pr, pk = R.T.copy(), K.T.copy()
    pself = np.zeros_like(pr)
    xi = prm["xi"]
    for i, j in pairs:
        pself[:, j] += xi * xi * ff[i] / (2 * np.pi)
    return -(pr + pk + pself), pr, pk, xi


def StokesSLP_vorticity_ewald_2p(xsrc, ysrc, xtar, ytar, f1, f2, Lx, Ly, P=24, Nb=None,
                                 tol=1e-16, verbose=0, ctx=None):
    """[omega, omegar, omegak, xi] (StokesSLP_vorticity_ewald_2p.m:1-147): omega =
    omegar + omegak - sum(f1 y - f2 x)/(2 Lx Ly) (:142-145)."""
    ps, ff, _, _, prm, R, K = _derived_driver("slp_vorticity", xsrc, ysrc, xtar, ytar, f1, f2,
                                              None, None, Lx, Ly, P, Nb, tol, ctx)
    omr, omk = R[:, 0].copy(), K[:, 0].copy()
    zm = (ff[:, 0] * ps[:, 1] - ff[:, 1] * ps[:, 0]).sum() / (2 * Lx * Ly)
    return omr + omk - zm, omr, omk, prm["xi"]


def StokesSLP_stress_ewald_2p(xsrc, ysrc, xtar, ytar, f1, f2, b1, b2, Lx, Ly, P=24, Nb=None,
                              tol=1e-16, verbose=0, ctx=None):
    """[sigma1, sigma2, sigmar, sigmak, xi] (StokesSLP_stress_ewald_2p.m:1-173): the
    stress applied to b, plus sum(f1 x + f2 y)/(2 Lx Ly) on both values (:168-169)."""
    ps, ff, _, _, prm, R, K = _derived_driver("slp_stress", xsrc, ysrc, xtar, ytar, f1, f2, None,
                                              None, Lx, Ly, P, Nb, tol, ctx)
    bb1, bb2 = _vec(b1, "b1"), _vec(b2, "b2")
    sr, sk = _project(R, bb1, bb2), _project(K, bb1, bb2)
    zm = (ff[:, 0] * ps[:, 0] + ff[:, 1] * ps[:, 1]).sum() / (2 * Lx * Ly)
    s = sr + sk + zm
    return s[0].copy(), s[1].copy(), sr, sk, prm["xi"]


def _dlp_ugrad_self(xi, f, n, b):
    """xi^2 (f (b.n) + n (f.b) + b (f.n))/(2 pi) (StokesDLP_gradient_ewald_2p.m:175-184)."""
    return xi * xi * (f * (b @ n) + n * (f @ b) + b * (f @ n)) / (2 * np.pi)


def StokesDLP_gradient_ewald_2p(xsrc, ysrc, xtar, ytar, n1, n2, f1, f2, b1, b2, Lx, Ly, P=24,
                                Nb=None, tol=1e-16, verbose=0, ctx=None):
    """[u1, u2, ur, uk, xi] (StokesDLP_gradient_ewald_2p.m:1-190): the velocity
    gradient applied to b, plus the self term at targets that are sources."""
    _, ff, nn, pairs, prm, R, K = _derived_driver("dlp_gradient", xsrc, ysrc, xtar, ytar, f1, f2,
                                                  n1, n2, Lx, Ly, P, Nb, tol, ctx)
    bb1, bb2 = _vec(b1, "b1"), _vec(b2, "b2")
    ur, uk = _project(R, bb1, bb2), _project(K, bb1, bb2)
    u = ur + uk
    for i, j in pairs:
        u[:, j] += _dlp_ugrad_self(prm["xi"], ff[i], nn[i], np.array([bb1[j], bb2[j]]))
    return u[0].copy(), u[1].copy(), ur, uk, prm["xi"]


def StokesDLP_pressure_grad_ewald_2p(xsrc, ysrc, xtar, ytar, n1, n2, f1, f2, Lx, Ly, P=24,
                                     Nb=None, tol=1e-16, verbose=0, ctx=None):
    """[p, pr, pk, xi] (StokesDLP_pressure_grad_ewald_2p.m:1-156): p = -2 (pr + pk)
    (:154)."""
    _, _, _, _, prm, R, K = _derived_driver("dlp_pressure_grad", xsrc, ysrc, xtar, ytar, f1, f2,
                                            n1, n2, Lx, Ly, P, Nb, tol, ctx)
    pr, pk = R.T.copy(), K.T.copy()
    return -2 * (pr + pk), pr, pk, prm["xi"]


def StokesDLP_vorticity_ewald_2p(xsrc, ysrc, xtar, ytar, n1, n2, f1, f2, Lx, Ly, P=24, Nb=None,
                                 tol=1e-16, verbose=0, ctx=None):
    """[omega, omegar, omegak, xi] (StokesDLP_vorticity_ewald_2p.m:1-161), with the
    SLP recipe (:113-131)."""
    _, _, _, _, prm, R, K = _derived_driver("dlp_vorticity", xsrc, ysrc, xtar, ytar, f1, f2, n1,
                                            n2, Lx, Ly, P, Nb, tol, ctx)
    omr, omk = R[:, 0].copy(), K[:, 0].copy()
    return omr + omk, omr, omk, prm["xi"]


def StokesDLP_stress_ewald_2p(xsrc, ysrc, xtar, ytar, n1, n2, f1, f2, b1, b2, Lx, Ly, P=24,
                              Nb=None, tol=1e-16, verbose=0, ctx=None):
    """[sigma1, sigma2, sigmar, sigmak, xi] (StokesDLP_stress_ewald_2p.m:1-200): the
    real part projected as [T1 b1 + T2 b2; T3 b1 + T4 b2] (:153-155), the k-space
    part as [T1 b1 + T3 b2; T2 b1 + T4 b2] plus sum(n.f)/(2 Lx Ly) (:164-170),
    and at targets that are sources the self term of :180-195.  The reference
    builds that self term as a complex matrix, -[pself'; 0] + 2 [Re ugradself;
    Im ugradself] with pself = xi^2 b (f.n)/(2 pi) complex; its real parts are
    kept here (sigma1 gets -xi^2 b1 (f.n)/(2 pi), sigma2 none), and the imaginary
    part the reference leaves on sigma1 when b2 != 0 is dropped."""
    _, ff, nn, pairs, prm, R, K = _derived_driver("dlp_stress", xsrc, ysrc, xtar, ytar, f1, f2,
                                                  n1, n2, Lx, Ly, P, Nb, tol, ctx)
    bb1, bb2 = _vec(b1, "b1"), _vec(b2, "b2")
    sr = _project(R, bb1, bb2, order=(0, 1, 2, 3))
    sk = _project(K, bb1, bb2) + (nn * ff).sum() / (2 * Lx * Ly)
    s = sr + sk
    xi = prm["xi"]
    for i, j in pairs:
        b = np.array([bb1[j], bb2[j]])
        ug = _dlp_ugrad_self(xi, ff[i], nn[i], b)
        s[0, j] += -xi * xi * b[0] * (ff[i] @ nn[i]) / (2 * np.pi) + 2 * ug[0]
        s[1, j] += 2 * ug[1]
    return s[0].copy(), s[1].copy(), sr, sk, xi
