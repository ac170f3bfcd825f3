"""Plans: fixed geometry, repeated matvecs on the GPU (se2p_plan_* in include/se2p.h)."""
import ctypes

import numpy as np

from . import _lib
from ._lib import Params, check, lib


def params_struct(d):
    """se2p_params from a dict with the field names of the C struct."""
    s = Params()
    for k, _ in Params._fields_:
        setattr(s, k, d.get(k, 0))
    return s


class DeviceArray:
    """A float64 device buffer owned by a Context."""

    def __init__(self, ctx, shape):
        self.ctx = ctx
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        self.ptr = ctx.alloc(max(self.nbytes, 8))

    @classmethod
    def from_host(cls, ctx, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        d = cls(ctx, a.shape)
        if a.nbytes:
            ctx.h2d(d.ptr, a)
        return d

    def upload(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == self.shape
        if a.nbytes:
            self.ctx.h2d(self.ptr, a)

    def download(self):
        out = np.empty(self.shape)
        if out.nbytes:
            self.ctx.d2h(out, self.ptr)
        return out

    def free(self):
        if self.ptr:
            self.ctx.free(self.ptr)
            self.ptr = None


class Plan:
    """Geometry binned once; apply() evaluates u = ur + uk (+ DLP zero mode).

    psrc, ptar, nrm: (N, 2) host arrays (interleaved like the mex 2 x N input).
    """

    def __init__(self, ctx, kind, psrc, ptar, Lx, Ly, prm, nrm=None):
        self.ctx = ctx
        self.kind = _lib.kind_code(kind)
        self.Ns, self.Nt = len(psrc), len(ptar)
        self.prm = params_struct(prm) if isinstance(prm, dict) else prm
        # the same array for sources and targets: one upload, so the library sees
        # one pointer (pair-symmetric sums when prm asks for them)
        same = ptar is psrc
        dps = DeviceArray.from_host(ctx, psrc)
        dpt = dps if same else DeviceArray.from_host(ctx, ptar)
        dn = DeviceArray.from_host(ctx, nrm) if nrm is not None else None
        self._h = ctypes.c_void_p()
        try:
            check(lib().se2p_plan_create(ctx.handle, self.kind, dps.ptr, dpt.ptr,
                                         dn.ptr if dn else None, self.Ns, self.Nt, float(Lx),
                                         float(Ly), ctypes.byref(self.prm),
                                         ctypes.byref(self._h)))
        finally:
            dps.free()
            if not same:
                dpt.free()
            if dn:
                dn.free()
        self.f = DeviceArray(ctx, (self.Ns, 2))
        nc = _lib.out_comps(self.kind)
        self.u = DeviceArray(ctx, (self.Nt, nc) if nc > 1 else (self.Nt,))

    def apply_device(self, d_f, d_u, graph=False):
        fn = lib().se2p_plan_apply_graph if graph else lib().se2p_plan_apply
        check(fn(self._h, d_f, d_u))

    def apply(self, f, graph=False):
        self.f.upload(f)
        self.apply_device(self.f.ptr, self.u.ptr, graph)
        return self.u.download()

    def params_in_use(self):
        """The parameters evaluations run with (se2p_plan_params): `symmetric`
        is 1 only when pair-symmetric real-space sums are in use."""
        out = Params()
        check(lib().se2p_plan_params(self._h, ctypes.byref(out)))
        return out.as_dict()

    def close(self):
        if self._h:
            lib().se2p_plan_destroy(self._h)
            self._h = ctypes.c_void_p()
        for a in (getattr(self, "f", None), getattr(self, "u", None)):
            if a is not None:
                a.free()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
