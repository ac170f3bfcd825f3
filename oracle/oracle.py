"""ctypes front end of the CPU oracle (oracle/se2p_oracle.c).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()``
as the checker and by ``bench.py``'s cpu_baseline leg; never by the product
package.  Arrays follow the reference layout: points, densities and normals
are (N, 2) C-contiguous float64 arrays (the reference's 2 x N column-major
layout has the same memory order), grids are (Mx, My) C arrays (the
reference's My x Mx column-major grid has the same memory order).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libse2p_oracle.so")
_lib = None

_D = ctypes.c_double
_I = ctypes.c_int
_P = ctypes.c_void_p


def build(force=False):
    """Compile the oracle with its Makefile (gcc)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", _HERE],
                       check=True)
    return _LIB_PATH


_REF_PATH = os.path.join(_HERE, "libse2p_refcpu.so")
_ref = None


def refcpu():
    """The CPU baseline build of the same stages (Makefile target
    libse2p_refcpu.so: -O3, AVX2/FMA, rational-fit E1); bench.py times it."""
    global _ref
    if _ref is None:
        if not os.path.exists(_REF_PATH):
            build(force=True)
        _ref = _load(_REF_PATH)
    return _ref


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _load(_LIB_PATH)
    return _lib


def _load(path):
    """ctypes handle of an oracle build with its argument types."""
    L = ctypes.CDLL(path)
    sig = {
        "orc_assign": [_P, _I, _D, _D, _I, _I, _P, _P, _P],
        "orc_find_closest_node": [_D, _D, _D, _D, _D, _I, _P, _P, _P, _P],
        "orc_slp_real": [_P, _P, _P, _I, _I, _D, _I, _I, _D, _D, _P],
        "orc_dlp_real": [_P, _P, _P, _P, _I, _I, _D, _I, _I, _D, _D, _P],
        "orc_spread": [_P, _P, _I, _D, _D, _D, _D, _D, _I, _I, _I, _P, _P],
        "orc_fft2": [_P, _P, _I, _I, _I],
        "orc_slp_filter": [_P, _P, _P, _P, _I, _I, _D, _D, _D, _D],
        "orc_dlp_filter": [_P] * 8 + [_I, _I, _D, _D, _D, _D],
        "orc_gather": [_P, _I, _I, _P, _I, _D, _D, _D, _D, _D, _I, _I, _I, _P],
        "orc_slp_kspace": [_P, _P, _D, _D, _P, _I, _I, _I, _I, _D, _D, _D, _I, _P],
        "orc_dlp_kspace": [_P, _P, _D, _D, _P, _P, _I, _I, _I, _I, _D, _D, _D, _I, _P],
        "orc_slp_real_ds": [_P, _P, _P, _I, _I, _D, _D, _D, _P],
        "orc_slp_kspace_ds": [_P, _P, _P, _I, _I, _D, _D, _D, _I, _P],
        "orc_dlp_real_ds": [_P, _P, _P, _P, _I, _I, _D, _D, _D, _P],
        "orc_dlp_kspace_ds": [_P, _P, _P, _P, _I, _I, _D, _D, _D, _I, _P],
        "orc_dlp_zero_mode": [_P, _P, _P, _I, _D, _D, _P],
        "orc_params": [_I, _I, _D, _D, _D, _I, _D, _D, _P, _P],
        "orc_rat": [_D, _P, _P],
        "orc_slp_pressure_real": [_P, _P, _P, _I, _I, _D, _I, _I, _D, _D, _P],
        "orc_slp_pressure_filter": [_P, _P, _P, _P, _I, _I, _D, _D, _D, _D],
        "orc_slp_pressure_kspace": [_P, _P, _P, _D, _D, _I, _I, _I, _I, _D, _D, _D, _I, _P],
        "orc_slp_pressure_real_ds": [_P, _P, _P, _I, _I, _D, _D, _D, _P],
        "orc_slp_pressure_kspace_ds": [_P, _P, _P, _I, _I, _D, _D, _D, _I, _P],
        "orc_dlp_pressure_real": [_P, _P, _P, _P, _I, _I, _D, _I, _I, _D, _D, _P],
        "orc_dlp_pressure_filter": [_P] * 8 + [_I, _I, _D, _D, _D, _D],
        "orc_dlp_pressure_kspace": [_P, _P, _P, _P, _D, _D, _I, _I, _I, _I, _D, _D, _D, _I,
                                    _P],
        "orc_dlp_pressure_real_ds": [_P, _P, _P, _P, _I, _I, _D, _D, _D, _P],
        "orc_dlp_pressure_kspace_ds": [_P, _P, _P, _P, _I, _I, _D, _D, _D, _I, _P],
        "orc_derived_real": [_I, _P, _P, _P, _P, _I, _I, _D, _I, _I, _D, _D, _P],
        "orc_derived_kspace": [_I, _P, _P, _P, _P, _D, _D, _I, _I, _I, _I, _D, _D, _D, _I,
                               _P],
        "orc_derived_real_ds": [_I, _P, _P, _P, _P, _P, _I, _I, _D, _D, _D, _P],
        "orc_derived_kspace_ds": [_I, _P, _P, _P, _P, _P, _I, _I, _D, _D, _D, _I, _P],
        "orc_free_slp": [_P, _P, _I, _P, _I, _I, _P],
        "orc_free_dlp": [_P, _P, _P, _I, _P],
    }
    for name, args in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = None
    L.orc_params.restype = _I
    L.orc_e1.argtypes = [_D]
    L.orc_e1.restype = _D
    L.orc_slp_pressure_zero_mode.argtypes = [_P, _P, _I, _D, _D]
    L.orc_slp_pressure_zero_mode.restype = _D
    L.orc_dlp_pressure_zero_mode.argtypes = [_P, _P, _I, _D, _D]
    L.orc_dlp_pressure_zero_mode.restype = _D
    return L


def _c(a, dtype=np.float64):
    return np.ascontiguousarray(a, dtype=dtype)


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def assign(p, Lx, Ly, nside_x, nside_y):
    p = _c(p)
    n = p.shape[0]
    nbox = nside_x * nside_y
    po = np.zeros(max(n, 1), np.int32)
    bo = np.zeros(nbox + 1, np.int32)
    nb = np.zeros(nbox, np.int32)
    lib().orc_assign(_ptr(p), n, Lx, Ly, nside_x, nside_y, _ptr(po), _ptr(bo), _ptr(nb))
    return po[:n], bo, nb


def find_closest_node(x, y, Lx, Ly, h, P):
    mx, my = ctypes.c_int(), ctypes.c_int()
    px, py = ctypes.c_double(), ctypes.c_double()
    lib().orc_find_closest_node(x, y, Lx, Ly, h, P, ctypes.byref(mx), ctypes.byref(my),
                                ctypes.byref(px), ctypes.byref(py))
    return mx.value, my.value, px.value, py.value


def e1(x):
    return lib().orc_e1(float(x))


def slp_real(psrc, ptar, f, xi, nside_x, nside_y, Lx, Ly, L=None):
    psrc, ptar, f = _c(psrc), _c(ptar), _c(f)
    u = np.zeros((ptar.shape[0], 2))
    (L or lib()).orc_slp_real(_ptr(psrc), _ptr(ptar), _ptr(f), psrc.shape[0], ptar.shape[0], xi,
                       nside_x, nside_y, Lx, Ly, _ptr(u))
    return u


def dlp_real(psrc, ptar, f, n, xi, nside_x, nside_y, Lx, Ly):
    psrc, ptar, f, n = _c(psrc), _c(ptar), _c(f), _c(n)
    u = np.zeros((ptar.shape[0], 2))
    lib().orc_dlp_real(_ptr(psrc), _ptr(ptar), _ptr(f), _ptr(n), psrc.shape[0],
                       ptar.shape[0], xi, nside_x, nside_y, Lx, Ly, _ptr(u))
    return u


def spread(psrc, f, Lx, Ly, xi, w, eta, P, Mx, My):
    psrc, f = _c(psrc), _c(f)
    H1 = np.zeros((Mx, My))
    H2 = np.zeros((Mx, My))
    lib().orc_spread(_ptr(psrc), _ptr(f), psrc.shape[0], Lx, Ly, xi, w, eta, P, Mx, My,
                     _ptr(H1), _ptr(H2))
    return H1, H2


def fft2(Z, inverse=False):
    """2D DFT of a complex (Mx, My) grid (axis 1 = y is contiguous)."""
    re = _c(np.real(Z))
    im = _c(np.imag(Z))
    Mx, My = re.shape
    lib().orc_fft2(_ptr(re), _ptr(im), My, Mx, 1 if inverse else 0)
    return re + 1j * im


def slp_filter(H1hat, H2hat, Lx, Ly, xi, eta):
    a, b = _c(H1hat.real), _c(H1hat.imag)
    c, d = _c(H2hat.real), _c(H2hat.imag)
    Mx, My = a.shape
    lib().orc_slp_filter(_ptr(a), _ptr(b), _ptr(c), _ptr(d), Mx, My, Lx, Ly, xi, eta)
    return a + 1j * b, c + 1j * d


def dlp_filter(H1hat, H2hat, H3hat, H4hat, Lx, Ly, xi, eta):
    arrs = [_c(x) for h in (H1hat, H2hat, H3hat, H4hat) for x in (h.real, h.imag)]
    Mx, My = arrs[0].shape
    lib().orc_dlp_filter(*[_ptr(a) for a in arrs], Mx, My, Lx, Ly, xi, eta)
    return arrs[0] + 1j * arrs[1], arrs[2] + 1j * arrs[3]


def gather(H, ptar, Lx, Ly, xi, w, eta, P):
    H, ptar = _c(H), _c(ptar)
    Mx, My = H.shape
    out = np.zeros((ptar.shape[0], 1))
    lib().orc_gather(_ptr(H), 1, 0, _ptr(ptar), ptar.shape[0], Lx, Ly, xi, w, eta, P, Mx,
                     My, _ptr(out))
    return out[:, 0]


def slp_kspace(psrc, ptar, xi, eta, f, Mx, My, Lx, Ly, w, P, L=None):
    psrc, ptar, f = _c(psrc), _c(ptar), _c(f)
    u = np.zeros((ptar.shape[0], 2))
    (L or lib()).orc_slp_kspace(_ptr(psrc), _ptr(ptar), xi, eta, _ptr(f), psrc.shape[0],
                         ptar.shape[0], Mx, My, Lx, Ly, w, P, _ptr(u))
    return u


def dlp_kspace(psrc, ptar, xi, eta, f, n, Mx, My, Lx, Ly, w, P):
    psrc, ptar, f, n = _c(psrc), _c(ptar), _c(f), _c(n)
    u = np.zeros((ptar.shape[0], 2))
    lib().orc_dlp_kspace(_ptr(psrc), _ptr(ptar), xi, eta, _ptr(f), _ptr(n), psrc.shape[0],
                         ptar.shape[0], Mx, My, Lx, Ly, w, P, _ptr(u))
    return u


def slp_real_ds(psrc, ptar, f, Lx, Ly, xi):
    psrc, ptar, f = _c(psrc), _c(ptar), _c(f)
    u = np.zeros((ptar.shape[0], 2))
    lib().orc_slp_real_ds(_ptr(psrc), _ptr(ptar), _ptr(f), psrc.shape[0], ptar.shape[0],
                          Lx, Ly, xi, _ptr(u))
    return u


def slp_kspace_ds(psrc, ptar, f, Lx, Ly, xi, kinf):
    psrc, ptar, f = _c(psrc), _c(ptar), _c(f)
    u = np.zeros((ptar.shape[0], 2))
    lib().orc_slp_kspace_ds(_ptr(psrc), _ptr(ptar), _ptr(f), psrc.shape[0], ptar.shape[0],
                            Lx, Ly, xi, kinf, _ptr(u))
    return u


def dlp_real_ds(psrc, ptar, f, n, Lx, Ly, xi):
    psrc, ptar, f, n = _c(psrc), _c(ptar), _c(f), _c(n)
    u = np.zeros((ptar.shape[0], 2))
    lib().orc_dlp_real_ds(_ptr(psrc), _ptr(ptar), _ptr(f), _ptr(n), psrc.shape[0],
                          ptar.shape[0], Lx, Ly, xi, _ptr(u))
    return u


def dlp_kspace_ds(psrc, ptar, f, n, Lx, Ly, xi, kinf):
    psrc, ptar, f, n = _c(psrc), _c(ptar), _c(f), _c(n)
    u = np.zeros((ptar.shape[0], 2))
    lib().orc_dlp_kspace_ds(_ptr(psrc), _ptr(ptar), _ptr(f), _ptr(n), psrc.shape[0],
                            ptar.shape[0], Lx, Ly, xi, kinf, _ptr(u))
    return u


def dlp_zero_mode(psrc, f, n, Lx, Ly):
    psrc, f, n = _c(psrc), _c(f), _c(n)
    out = np.zeros(2)
    lib().orc_dlp_zero_mode(_ptr(psrc), _ptr(f), _ptr(n), psrc.shape[0], Lx, Ly, _ptr(out))
    return out


def slp_pressure_real(psrc, ptar, f, xi, nside_x, nside_y, Lx, Ly):
    psrc, ptar, f = _c(psrc), _c(ptar), _c(f)
    p = np.zeros(ptar.shape[0])
    lib().orc_slp_pressure_real(_ptr(psrc), _ptr(ptar), _ptr(f), psrc.shape[0], ptar.shape[0],
                                xi, nside_x, nside_y, Lx, Ly, _ptr(p))
    return p


def slp_pressure_filter(H1hat, H2hat, Lx, Ly, xi, eta):
    """Pressure filter of the two spread spectra (Mx x My arrays, [ix, iy]);
    returns the filtered spectrum."""
    a, b = _c(H1hat.real), _c(H1hat.imag)
    c, d = _c(H2hat.real), _c(H2hat.imag)
    Mx, My = a.shape
    lib().orc_slp_pressure_filter(_ptr(a), _ptr(b), _ptr(c), _ptr(d), Mx, My, Lx, Ly, xi, eta)
    return a + 1j * b


def slp_pressure_kspace(psrc, ptar, f, xi, eta, Mx, My, Lx, Ly, w, P):
    psrc, ptar, f = _c(psrc), _c(ptar), _c(f)
    p = np.zeros(ptar.shape[0])
    lib().orc_slp_pressure_kspace(_ptr(psrc), _ptr(ptar), _ptr(f), xi, eta, psrc.shape[0],
                                  ptar.shape[0], Mx, My, Lx, Ly, w, P, _ptr(p))
    return p


def slp_pressure_zero_mode(psrc, f, Lx, Ly):
    psrc, f = _c(psrc), _c(f)
    return lib().orc_slp_pressure_zero_mode(_ptr(psrc), _ptr(f), psrc.shape[0], Lx, Ly)


def slp_pressure_real_ds(psrc, ptar, f, Lx, Ly, xi):
    psrc, ptar, f = _c(psrc), _c(ptar), _c(f)
    p = np.zeros(ptar.shape[0])
    lib().orc_slp_pressure_real_ds(_ptr(psrc), _ptr(ptar), _ptr(f), psrc.shape[0],
                                   ptar.shape[0], Lx, Ly, xi, _ptr(p))
    return p


def slp_pressure_kspace_ds(psrc, ptar, f, Lx, Ly, xi, kinf):
    """Direct k-space sum including the zero mode (stokes_slp_pressure_kspace_ds.m)."""
    psrc, ptar, f = _c(psrc), _c(ptar), _c(f)
    p = np.zeros(ptar.shape[0])
    lib().orc_slp_pressure_kspace_ds(_ptr(psrc), _ptr(ptar), _ptr(f), psrc.shape[0],
                                     ptar.shape[0], Lx, Ly, xi, kinf, _ptr(p))
    return p


def dlp_pressure_real(psrc, ptar, f, n, xi, nside_x, nside_y, Lx, Ly):
    psrc, ptar, f, n = _c(psrc), _c(ptar), _c(f), _c(n)
    p = np.zeros(ptar.shape[0])
    lib().orc_dlp_pressure_real(_ptr(psrc), _ptr(ptar), _ptr(f), _ptr(n), psrc.shape[0],
                                ptar.shape[0], xi, nside_x, nside_y, Lx, Ly, _ptr(p))
    return p


def dlp_pressure_filter(H1, H2, H3, H4, Lx, Ly, xi, eta):
    """Pressure filter of the four spread spectra (f1 n1, f2 n1, f1 n2, f2 n2),
    Mx x My arrays [ix, iy]; returns the filtered spectrum."""
    arrs = [_c(x) for h in (H1, H2, H3, H4) for x in (h.real, h.imag)]
    Mx, My = arrs[0].shape
    lib().orc_dlp_pressure_filter(*[_ptr(a) for a in arrs], Mx, My, Lx, Ly, xi, eta)
    return arrs[0] + 1j * arrs[1]


def dlp_pressure_kspace(psrc, ptar, f, n, xi, eta, Mx, My, Lx, Ly, w, P):
    psrc, ptar, f, n = _c(psrc), _c(ptar), _c(f), _c(n)
    p = np.zeros(ptar.shape[0])
    lib().orc_dlp_pressure_kspace(_ptr(psrc), _ptr(ptar), _ptr(f), _ptr(n), xi, eta,
                                  psrc.shape[0], ptar.shape[0], Mx, My, Lx, Ly, w, P, _ptr(p))
    return p


def dlp_pressure_zero_mode(f, n, Lx, Ly):
    f, n = _c(f), _c(n)
    return lib().orc_dlp_pressure_zero_mode(_ptr(f), _ptr(n), f.shape[0], Lx, Ly)


def dlp_pressure_real_ds(psrc, ptar, f, n, Lx, Ly, xi):
    psrc, ptar, f, n = _c(psrc), _c(ptar), _c(f), _c(n)
    p = np.zeros(ptar.shape[0])
    lib().orc_dlp_pressure_real_ds(_ptr(psrc), _ptr(ptar), _ptr(f), _ptr(n), psrc.shape[0],
                                   ptar.shape[0], Lx, Ly, xi, _ptr(p))
    return p


def dlp_pressure_kspace_ds(psrc, ptar, f, n, Lx, Ly, xi, kinf):
    """Direct k-space sum including the zero mode (stokes_dlp_pressure_kspace_ds.m)."""
    psrc, ptar, f, n = _c(psrc), _c(ptar), _c(f), _c(n)
    p = np.zeros(ptar.shape[0])
    lib().orc_dlp_pressure_kspace_ds(_ptr(psrc), _ptr(ptar), _ptr(f), _ptr(n), psrc.shape[0],
                                     ptar.shape[0], Lx, Ly, xi, kinf, _ptr(p))
    return p


# Derived quantities (kind codes of the C oracle; the same numbers as the
# product's SE2P_* kinds).
DERIVED = {"slp_gradient": 4, "slp_pressure_grad": 5, "slp_vorticity": 6, "slp_stress": 7,
           "dlp_gradient": 8, "dlp_pressure_grad": 9, "dlp_vorticity": 10, "dlp_stress": 11}
NCOMP = {4: 4, 5: 2, 6: 1, 7: 4, 8: 4, 9: 2, 10: 1, 11: 4}


def _kind(kind):
    return DERIVED[kind] if isinstance(kind, str) else int(kind)


def _nullable(a):
    return None if a is None else _c(a)


def derived_real(kind, psrc, ptar, f, n, xi, nside_x, nside_y, Lx, Ly):
    """Real-space mex restatement; returns (Nt, nc) (rows in the mex order)."""
    k = _kind(kind)
    psrc, ptar, f, n = _c(psrc), _c(ptar), _c(f), _nullable(n)
    out = np.zeros((ptar.shape[0], NCOMP[k]))
    lib().orc_derived_real(k, _ptr(psrc), _ptr(ptar), _ptr(f), None if n is None else _ptr(n),
                           psrc.shape[0], ptar.shape[0], xi, nside_x, nside_y, Lx, Ly,
                           _ptr(out))
    return out


def derived_kspace(kind, psrc, ptar, f, n, xi, eta, Mx, My, Lx, Ly, w, P):
    """k-space mex restatement (no zero mode); returns (Nt, nc)."""
    k = _kind(kind)
    psrc, ptar, f, n = _c(psrc), _c(ptar), _c(f), _nullable(n)
    out = np.zeros((ptar.shape[0], NCOMP[k]))
    lib().orc_derived_kspace(k, _ptr(psrc), _ptr(ptar), _ptr(f),
                             None if n is None else _ptr(n), xi, eta, psrc.shape[0],
                             ptar.shape[0], Mx, My, Lx, Ly, w, P, _ptr(out))
    return out


def derived_real_ds(kind, psrc, ptar, f, n, b, Lx, Ly, xi):
    """stokes_*_real_ds.m: (Nt, 2) (vorticity (Nt,)), gradient/stress projected on b."""
    k = _kind(kind)
    psrc, ptar, f, n, b = _c(psrc), _c(ptar), _c(f), _nullable(n), _nullable(b)
    no = 1 if k in (6, 10) else 2
    out = np.zeros((ptar.shape[0], no))
    lib().orc_derived_real_ds(k, _ptr(psrc), _ptr(ptar), _ptr(f),
                              None if n is None else _ptr(n), None if b is None else _ptr(b),
                              psrc.shape[0], ptar.shape[0], Lx, Ly, xi, _ptr(out))
    return out[:, 0] if no == 1 else out


def derived_kspace_ds(kind, psrc, ptar, f, n, b, Lx, Ly, xi, kinf):
    """stokes_*_kspace_ds.m (pair by pair: small cases only)."""
    k = _kind(kind)
    psrc, ptar, f, n, b = _c(psrc), _c(ptar), _c(f), _nullable(n), _nullable(b)
    no = 1 if k in (6, 10) else 2
    out = np.zeros((ptar.shape[0], no))
    lib().orc_derived_kspace_ds(k, _ptr(psrc), _ptr(ptar), _ptr(f),
                                None if n is None else _ptr(n), None if b is None else _ptr(b),
                                psrc.shape[0], ptar.shape[0], Lx, Ly, xi, kinf, _ptr(out))
    return out[:, 0] if no == 1 else out


def free_slp(f, psrc, ptar, self=False):
    """Free-space SLP direct sum (stokes_SLP_FMM_test.m:29-51); (Nt, 2)."""
    f, psrc, ptar = _c(f), _c(psrc), _c(ptar)
    u = np.zeros((ptar.shape[0], 2))
    lib().orc_free_slp(_ptr(f), _ptr(psrc), psrc.shape[0], _ptr(ptar), ptar.shape[0],
                       int(bool(self)), _ptr(u))
    return u


def free_dlp(f, n, p):
    """Free-space DLP direct sum at the sources (stokes_DLP_FMM_test.m:17-35); (N, 2)."""
    f, n, p = _c(f), _c(n), _c(p)
    u = np.zeros((p.shape[0], 2))
    lib().orc_free_dlp(_ptr(f), _ptr(n), _ptr(p), p.shape[0], _ptr(u))
    return u


def density_q(f):
    """Q = sum(sum(f.^2)) + 1 for f = [f1'; f2'] (StokesSLP_ewald_2p.m:82,87):
    the column sums f1(i)^2 + f2(i)^2 added sequentially in index order."""
    f = np.asarray(f, dtype=np.float64).reshape(-1, 2)
    q = 0.0
    for v in (f[:, 0] * f[:, 0] + f[:, 1] * f[:, 1]).tolist():
        q += v
    return q + 1.0


def params(kind, npts, Q, Lx, Ly, P, Nb, tol):
    """Reference parameter recipe; kind 'slp' or 'dlp'.  Returns a dict."""
    out = np.zeros(4)
    iout = np.zeros(7, np.int32)
    rc = lib().orc_params(0 if kind == "slp" else 1, int(npts), float(Q), float(Lx),
                          float(Ly), int(P), float(Nb), float(tol), _ptr(out), _ptr(iout))
    keys = ("nside_x", "nside_y", "Mx", "My", "kinf", "A", "B")
    d = dict(zip(keys, (int(v) for v in iout)))
    d.update(xi=out[0], rc=out[1], w=out[2], eta=out[3], converged=(rc == 0), P=int(P))
    return d


def rat(X):
    N, D = ctypes.c_int(), ctypes.c_int()
    lib().orc_rat(float(X), ctypes.byref(N), ctypes.byref(D))
    return N.value, D.value


def default_Nb(npts, P=24):
    """StokesSLP_ewald_2p.m:33: min(P*round(log2(npts)), npts/4), MATLAB round."""
    l2 = np.log2(npts)
    r = np.floor(l2 + 0.5)
    return min(P * r, npts / 4)


def wrap(p, Lx, Ly):
    """StokesSLP_ewald_2p.m:75-78: mod(x + L/2, L) - L/2 with MATLAB mod."""
    p = np.array(p, dtype=np.float64, copy=True)
    for c, L in ((0, Lx), (1, Ly)):
        v = p[:, c] + L / 2
        p[:, c] = (v - np.floor(v / L) * L) - L / 2
    return p
