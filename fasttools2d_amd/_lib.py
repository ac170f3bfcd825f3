"""ctypes binding of libse2p.so (the C ABI declared in include/se2p.h).

The library is built in-tree (fasttools2d_amd/libse2p.so) by
``__graft_entry__.build()`` or ``make -C fasttools2d_amd/csrc``.  Importing
this module fails with an explicit error when the library is missing: there
is no fallback implementation.
"""
import ctypes
import os
import threading

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libse2p.so")

SE2P_OK = 0
SE2P_ERR_ARG = -1
SE2P_ERR_HIP = -2
SE2P_ERR_NOMEM = -3
SE2P_ERR_UNSUPPORTED = -4
SE2P_ERR_NODEV = -5
SE2P_SLP = 0
SE2P_DLP = 1
SE2P_SLP_PRESSURE = 2
SE2P_DLP_PRESSURE = 3
SE2P_SLP_GRADIENT = 4
SE2P_SLP_PRESSURE_GRAD = 5
SE2P_SLP_VORTICITY = 6
SE2P_SLP_STRESS = 7
SE2P_DLP_GRADIENT = 8
SE2P_DLP_PRESSURE_GRAD = 9
SE2P_DLP_VORTICITY = 10
SE2P_DLP_STRESS = 11
SE2P_FP64 = 0
SE2P_FP32C = 1
SE2P_FREE_AUTO = 0
SE2P_FREE_DIRECT = 1
SE2P_FREE_EWALD = 2

_KINDS = {"slp": SE2P_SLP, "dlp": SE2P_DLP, "slp_pressure": SE2P_SLP_PRESSURE,
          "dlp_pressure": SE2P_DLP_PRESSURE, "slp_gradient": SE2P_SLP_GRADIENT,
          "slp_pressure_grad": SE2P_SLP_PRESSURE_GRAD, "slp_vorticity": SE2P_SLP_VORTICITY,
          "slp_stress": SE2P_SLP_STRESS, "dlp_gradient": SE2P_DLP_GRADIENT,
          "dlp_pressure_grad": SE2P_DLP_PRESSURE_GRAD, "dlp_vorticity": SE2P_DLP_VORTICITY,
          "dlp_stress": SE2P_DLP_STRESS}
_NCOMP = {SE2P_SLP: 2, SE2P_DLP: 2, SE2P_SLP_PRESSURE: 1, SE2P_DLP_PRESSURE: 1,
          SE2P_SLP_GRADIENT: 4, SE2P_SLP_PRESSURE_GRAD: 2, SE2P_SLP_VORTICITY: 1,
          SE2P_SLP_STRESS: 4, SE2P_DLP_GRADIENT: 4, SE2P_DLP_PRESSURE_GRAD: 2,
          SE2P_DLP_VORTICITY: 1, SE2P_DLP_STRESS: 4}


def kind_code(kind):
    """Kind name ('slp', 'dlp', 'slp_pressure', ..., 'dlp_stress') or integer code -> code."""
    if isinstance(kind, str):
        return _KINDS[kind]
    if kind not in _KINDS.values():
        raise ValueError(f"unknown kind {kind}")
    return int(kind)


def out_comps(kind):
    """Values per target: 2 (velocity, pressure gradient), 1 (pressure, vorticity)
    or 4 (velocity gradient, stress)."""
    return _NCOMP[kind_code(kind)]


def has_normals(kind):
    """Stresslet kinds carry source normals."""
    k = kind_code(kind)
    return k in (SE2P_DLP, SE2P_DLP_PRESSURE) or k >= SE2P_DLP_GRADIENT

EXPORTS = (
    "se2p_ctx_create", "se2p_ctx_destroy", "se2p_last_error", "se2p_version",
    "se2p_device_count", "se2p_ctx_set_precision", "se2p_slp_real", "se2p_slp_kspace", "se2p_dlp_real",
    "se2p_dlp_kspace", "se2p_options_default", "se2p_params_ref", "se2p_slp_ewald",
    "se2p_dlp_ewald", "se2p_dev_alloc", "se2p_dev_free", "se2p_memcpy_h2d",
    "se2p_memcpy_d2h", "se2p_memcpy_d2d", "se2p_ctx_sync", "se2p_eval_device",
    "se2p_eval_device_noalloc", "se2p_prof_enable",
    "se2p_prof_reset", "se2p_prof_get", "se2p_assign", "se2p_find_closest_node",
    "se2p_spread", "se2p_gather", "se2p_fft2", "se2p_e1", "se2p_plan_create",
    "se2p_plan_apply", "se2p_plan_apply_graph", "se2p_plan_destroy", "se2p_plan_params",
    "se2p_batch_create", "se2p_batch_sizes", "se2p_batch_apply", "se2p_batch_apply_graph",
    "se2p_batch_destroy", "se2p_batch_params",
    "se2p_slp_pressure_real", "se2p_slp_pressure_kspace", "se2p_slp_pressure_ewald",
    "se2p_dlp_pressure_real", "se2p_dlp_pressure_kspace", "se2p_dlp_pressure_ewald",
) + tuple(f"se2p_{layer}_{q}_{part}" for layer in ("slp", "dlp")
          for q in ("gradient", "pressure_grad", "vorticity", "stress")
          for part in ("real", "kspace")) + ("se2p_free_slp", "se2p_free_dlp",
                                                  "se2p_free_params", "se2p_ctx_set_free_method",
                                                  "se2p_mlp_forward", "se2p_mlp_loss_grad",
                                                  "se2p_mlp_adam")


class Se2pError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"se2p error {code}: {msg}")
        self.code = code


class Options(ctypes.Structure):
    _fields_ = [("P", ctypes.c_int), ("Nb", ctypes.c_double), ("tol", ctypes.c_double),
                ("verbose", ctypes.c_int)]


class Params(ctypes.Structure):
    _fields_ = [("xi", ctypes.c_double), ("rc", ctypes.c_double), ("w", ctypes.c_double),
                ("eta", ctypes.c_double), ("nside_x", ctypes.c_int), ("nside_y", ctypes.c_int),
                ("Mx", ctypes.c_int), ("My", ctypes.c_int), ("P", ctypes.c_int),
                ("kinf", ctypes.c_int), ("converged", ctypes.c_int), ("e1_tier", ctypes.c_int),
                ("symmetric", ctypes.c_int), ("precision", ctypes.c_int),
                ("far32", ctypes.c_double), ("free_R", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class System(ctypes.Structure):
    """se2p_system: one system of a batch (device pointers)."""
    _fields_ = [("kind", ctypes.c_int), ("Nsrc", ctypes.c_int), ("Ntar", ctypes.c_int),
                ("Lx", ctypes.c_double), ("Ly", ctypes.c_double), ("d_psrc", ctypes.c_void_p),
                ("d_ptar", ctypes.c_void_p), ("d_n", ctypes.c_void_p), ("prm", Params)]


_lib = None
_lock = threading.Lock()


def lib():
    """Load libse2p.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; "
                f"g.build()'` or `make -C fasttools2d_amd/csrc`")
        L = ctypes.CDLL(LIB_PATH)
        D, I, P, Z = ctypes.c_double, ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t
        sig = {
            "se2p_ctx_create": [I, ctypes.POINTER(P)],
            "se2p_ctx_destroy": [P],
            "se2p_device_count": [ctypes.POINTER(I)],
            "se2p_ctx_set_precision": [P, I],
            "se2p_slp_real": [P, P, P, P, I, I, D, I, I, D, D, P],
            "se2p_slp_kspace": [P, P, P, D, D, P, I, I, I, I, D, D, D, I, P],
            "se2p_dlp_real": [P, P, P, P, P, I, I, D, I, I, D, D, P],
            "se2p_dlp_kspace": [P, P, P, D, D, P, P, I, I, I, I, D, D, D, I, P],
            "se2p_params_ref": [I, I, I, D, D, D, ctypes.POINTER(Options), ctypes.POINTER(Params)],
            "se2p_slp_ewald": [P, P, P, P, P, P, P, I, I, D, D, ctypes.POINTER(Options), P, P, P,
                               P, P],
            "se2p_dlp_ewald": [P, P, P, P, P, P, P, P, P, I, I, D, D, ctypes.POINTER(Options), P,
                               P, P, P, P],
            "se2p_dev_alloc": [P, Z, ctypes.POINTER(P)],
            "se2p_dev_free": [P, P],
            "se2p_memcpy_h2d": [P, P, P, Z],
            "se2p_memcpy_d2h": [P, P, P, Z],
            "se2p_memcpy_d2d": [P, P, P, Z],
            "se2p_ctx_sync": [P],
            "se2p_eval_device": [P, I, P, P, P, P, I, I, D, D, ctypes.POINTER(Params), P],
            "se2p_eval_device_noalloc": [P, I, P, P, P, P, I, I, D, D, ctypes.POINTER(Params), P],
            "se2p_prof_enable": [P, I],
            "se2p_prof_reset": [P],
            "se2p_prof_get": [P, ctypes.c_char_p, ctypes.POINTER(D), ctypes.POINTER(I)],
            "se2p_assign": [P, P, I, D, D, I, I, P, P, P],
            "se2p_find_closest_node": [P, P, I, D, D, D, I, P, P, P, P],
            "se2p_spread": [P, P, P, I, D, D, D, D, D, I, I, I, P, P],
            "se2p_gather": [P, P, P, I, D, D, D, D, D, I, I, I, P],
            "se2p_fft2": [P, P, P, I, I, I],
            "se2p_e1": [P, P, I, I, P],
            "se2p_plan_create": [P, I, P, P, P, I, I, D, D, ctypes.POINTER(Params),
                                 ctypes.POINTER(P)],
            "se2p_plan_apply": [P, P, P],
            "se2p_plan_apply_graph": [P, P, P],
            "se2p_plan_destroy": [P],
            "se2p_plan_params": [P, ctypes.POINTER(Params)],
            "se2p_batch_create": [P, I, ctypes.POINTER(System), ctypes.POINTER(P)],
            "se2p_batch_sizes": [P, ctypes.POINTER(ctypes.c_longlong),
                                 ctypes.POINTER(ctypes.c_longlong)],
            "se2p_batch_apply": [P, P, P],
            "se2p_batch_apply_graph": [P, P, P],
            "se2p_batch_destroy": [P],
            "se2p_batch_params": [P, I, ctypes.POINTER(Params)],
            "se2p_slp_pressure_real": [P, P, P, P, I, I, D, I, I, D, D, P],
            "se2p_slp_pressure_kspace": [P, P, P, P, D, D, I, I, I, I, D, D, D, I, P],
            "se2p_slp_pressure_ewald": [P, P, P, P, P, P, P, I, I, D, D, ctypes.POINTER(Options),
                                        P, P, P, P],
            "se2p_dlp_pressure_real": [P, P, P, P, P, I, I, D, I, I, D, D, P],
            "se2p_dlp_pressure_kspace": [P, P, P, P, P, D, D, I, I, I, I, D, D, D, I, P],
            "se2p_dlp_pressure_ewald": [P, P, P, P, P, P, P, P, P, I, I, D, D,
                                        ctypes.POINTER(Options), P, P, P, P],
            # derived quantities: real parts (psrc, ptar, f[, n], Ns, Nt, xi, nsx, nsy,
            # Lx, Ly, out); k-space parts in each mex function's argument order
            **{f"se2p_slp_{q}_real": [P, P, P, P, I, I, D, I, I, D, D, P]
               for q in ("gradient", "pressure_grad", "vorticity", "stress")},
            **{f"se2p_dlp_{q}_real": [P, P, P, P, P, I, I, D, I, I, D, D, P]
               for q in ("gradient", "pressure_grad", "vorticity", "stress")},
            **{f"se2p_slp_{q}_kspace": [P, P, P, D, D, P, I, I, I, I, D, D, D, I, P]
               for q in ("gradient", "vorticity", "stress")},
            "se2p_slp_pressure_grad_kspace": [P, P, P, P, D, D, I, I, I, I, D, D, D, I, P],
            **{f"se2p_dlp_{q}_kspace": [P, P, P, D, D, P, P, I, I, I, I, D, D, D, I, P]
               for q in ("gradient", "stress")},
            **{f"se2p_dlp_{q}_kspace": [P, P, P, P, P, D, D, I, I, I, I, D, D, D, I, P]
               for q in ("pressure_grad", "vorticity")},
            "se2p_free_slp": [P, P, P, P, P, I, P, P, I, I, I, P, P],
            "se2p_free_dlp": [P, P, P, P, P, P, P, I, P, P],
            "se2p_free_params": [I, I, D, D, D, P, P, P],
            "se2p_ctx_set_free_method": [P, I],
            "se2p_mlp_forward": [P, P, P, I, P],
            "se2p_mlp_loss_grad": [P, P, P, P, I, P, P],
            "se2p_mlp_adam": [P, P, P, P, P, I, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                              ctypes.c_float],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = I
        L.se2p_last_error.argtypes = []
        L.se2p_last_error.restype = ctypes.c_char_p
        L.se2p_version.argtypes = []
        L.se2p_version.restype = I
        L.se2p_options_default.argtypes = [ctypes.POINTER(Options)]
        L.se2p_options_default.restype = None
        _lib = L
    return _lib


def check(rc):
    if rc != SE2P_OK:
        raise Se2pError(rc, lib().se2p_last_error().decode(errors="replace"))
    return rc


def device_count():
    n = ctypes.c_int(0)
    check(lib().se2p_device_count(ctypes.byref(n)))
    return n.value


class Context:
    """One GPU and one HIP stream (se2p_ctx)."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        check(lib().se2p_ctx_create(int(device), ctypes.byref(self._h)))
        self.device = device

    @property
    def handle(self):
        if not self._h:
            raise Se2pError(SE2P_ERR_ARG, "context is closed")
        return self._h

    def close(self):
        if self._h:
            lib().se2p_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(lib().se2p_ctx_sync(self.handle))

    def set_free_method(self, method):
        """SE2P_FREE_AUTO (0), SE2P_FREE_DIRECT (1) or SE2P_FREE_EWALD (2) for
        this context's free-space Stokeslet sums (stokesSLPfmm)."""
        check(lib().se2p_ctx_set_free_method(self.handle, int(method)))

    def set_precision(self, precision):
        """SE2P_FP64 (0) or SE2P_FP32C (1) for this context's one-shot calls (the
        mex equivalents and drivers)."""
        check(lib().se2p_ctx_set_precision(self.handle, int(precision)))

    # device memory owned by the caller
    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        check(lib().se2p_dev_alloc(self.handle, int(nbytes), ctypes.byref(p)))
        return p

    def free(self, p):
        check(lib().se2p_dev_free(self.handle, p))

    def h2d(self, dst, arr):
        check(lib().se2p_memcpy_h2d(self.handle, dst, arr.ctypes.data_as(ctypes.c_void_p),
                                    arr.nbytes))

    def d2h(self, arr, src):
        check(lib().se2p_memcpy_d2h(self.handle, arr.ctypes.data_as(ctypes.c_void_p), src,
                                    arr.nbytes))

    def d2d(self, dst, src, nbytes):
        check(lib().se2p_memcpy_d2d(self.handle, dst, src, int(nbytes)))

    def prof_enable(self, on=True, serial=False):
        """Per-stage HIP-event timing.  serial: every stage of an apply on one
        stream in order, so each is timed alone; otherwise the real-space and
        k-space passes keep running concurrently and each stage is timed on
        its own stream."""
        check(lib().se2p_prof_enable(self.handle, (2 if serial else 1) if on else 0))

    def prof_reset(self):
        check(lib().se2p_prof_reset(self.handle))

    def prof_get(self, stage):
        ms, cnt = ctypes.c_double(), ctypes.c_int()
        check(lib().se2p_prof_get(self.handle, stage.encode(), ctypes.byref(ms),
                                  ctypes.byref(cnt)))
        return ms.value, cnt.value


_tls = threading.local()


def default_context():
    """Per-thread context on device 0, created on first use."""
    ctx = getattr(_tls, "ctx", None)
    if ctx is None:
        ctx = Context(0)
        _tls.ctx = ctx
    return ctx
