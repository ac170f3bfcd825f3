"""Batches of independent periodic systems evaluated together (se2p_batch_* in
include/se2p.h): each system has its own points, box, kind (SLP or DLP) and
Ewald parameters; every stage of an evaluation runs as one launch over the
blocks of all systems, and repeated evaluations on the fixed geometry can be
replayed from a HIP graph.  The reference evaluates one system per call
(PeriodicEwald/src/matlab/StokesSLP_ewald_2p.m:119,126)."""
import ctypes

import numpy as np

from . import _lib
from ._lib import Params, System, check, lib
from .plan import DeviceArray, params_struct


class Batch:
    """systems: list of dicts with keys kind ("slp", "dlp", "slp_pressure",
    "dlp_pressure"), psrc, ptar (N x 2 host arrays), nrm (stresslet kinds), Lx, Ly,
    prm (parameter dict or se2p_params)."""

    def __init__(self, ctx, systems):
        self.ctx = ctx
        self.ns = [len(y["psrc"]) for y in systems]
        self.nt = [len(y["ptar"]) for y in systems]
        self.nc = [_lib.out_comps(y.get("kind", "slp")) for y in systems]
        tmp = []
        arr = (System * max(1, len(systems)))()
        try:
            for i, y in enumerate(systems):
                kind = _lib.kind_code(y.get("kind", "slp"))
                dps = DeviceArray.from_host(ctx, y["psrc"])
                # the same array for sources and targets: one upload, so the
                # library sees one pointer (pair-symmetric sums when asked for)
                same = y["ptar"] is y["psrc"]
                dpt = dps if same else DeviceArray.from_host(ctx, y["ptar"])
                dn = DeviceArray.from_host(ctx, y["nrm"]) if _lib.has_normals(kind) else None
                tmp += [a for a in (dps, None if same else dpt, dn) if a is not None]
                prm = y["prm"]
                arr[i] = System(kind, self.ns[i], self.nt[i], float(y["Lx"]), float(y["Ly"]),
                                dps.ptr, dpt.ptr, dn.ptr if dn else None,
                                params_struct(prm) if isinstance(prm, dict) else prm)
            self._h = ctypes.c_void_p()
            check(lib().se2p_batch_create(ctx.handle, len(systems), arr, ctypes.byref(self._h)))
        finally:
            for a in tmp:
                a.free()
        nf, nu = ctypes.c_longlong(), ctypes.c_longlong()
        check(lib().se2p_batch_sizes(self._h, ctypes.byref(nf), ctypes.byref(nu)))
        self.nf, self.nu = nf.value, nu.value
        self.f = DeviceArray(ctx, (max(1, self.nf),))
        self.u = DeviceArray(ctx, (max(1, self.nu),))
        self.f_off = np.concatenate([[0], np.cumsum(2 * np.array(self.ns, dtype=np.int64))])
        self.u_off = np.concatenate([[0], np.cumsum(np.array(self.nc, dtype=np.int64) *
                                                    np.array(self.nt, dtype=np.int64))])

    def apply_device(self, d_f, d_u, graph=False):
        fn = lib().se2p_batch_apply_graph if graph else lib().se2p_batch_apply
        check(fn(self._h, d_f, d_u))

    def apply(self, fs, graph=False):
        """fs: list of (Nsrc_i x 2) densities; returns the list of (Ntar_i x 2) velocities."""
        if self.nf:
            self.f.upload(np.concatenate([np.ascontiguousarray(f, dtype=np.float64).ravel()
                                          for f in fs]))
        self.apply_device(self.f.ptr, self.u.ptr, graph)
        u = self.u.download() if self.nu else np.zeros(0)
        return [u[self.u_off[i]:self.u_off[i + 1]].reshape(-1, self.nc[i]) if self.nc[i] > 1
                else u[self.u_off[i]:self.u_off[i + 1]] for i in range(len(self.nt))]

    def params_in_use(self, i):
        """Parameters system i runs with (se2p_batch_params); `symmetric` is 1
        only when its pair-symmetric real-space sums are in use."""
        out = Params()
        check(lib().se2p_batch_params(self._h, int(i), ctypes.byref(out)))
        return out.as_dict()

    def close(self):
        if getattr(self, "_h", None):
            lib().se2p_batch_destroy(self._h)
            self._h = ctypes.c_void_p()
        for a in (getattr(self, "f", None), getattr(self, "u", None)):
            if a is not None:
                a.free()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
