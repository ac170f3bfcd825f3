"""fasttools2d_amd: MI355X-native spectral Ewald sums for doubly-periodic 2D Stokes flow.

The compute path is libse2p.so (hand-written HIP for gfx950 behind a C ABI,
include/se2p.h); this package is its Python front end, mirroring the
reference's PeriodicEwald functions.
"""
from ._lib import (Context, Se2pError, default_context, device_count, LIB_PATH,  # noqa: F401
                   SE2P_SLP, SE2P_DLP, SE2P_SLP_PRESSURE, SE2P_DLP_PRESSURE, SE2P_SLP_GRADIENT,
                   SE2P_SLP_PRESSURE_GRAD, SE2P_SLP_VORTICITY, SE2P_SLP_STRESS,
                   SE2P_DLP_GRADIENT, SE2P_DLP_PRESSURE_GRAD, SE2P_DLP_VORTICITY,
                   SE2P_DLP_STRESS)
from . import periodic_ewald as _pe
from .freespace import stokesDLPfmm, stokesSLPfmm  # noqa: F401
from .periodic_ewald import (StokesDLP_ewald_2p, StokesDLP_pressure_ewald_2p,  # noqa: F401
                             StokesSLP_ewald_2p, StokesSLP_pressure_ewald_2p, density_q,
                             ewald_params,
                             mex_stokes_dlp_kspace, mex_stokes_dlp_pressure_kspace,
                             mex_stokes_dlp_pressure_real, mex_stokes_dlp_real,
                             mex_stokes_slp_kspace, mex_stokes_slp_pressure_kspace,
                             mex_stokes_slp_pressure_real, mex_stokes_slp_real)

__all__ = [
    "Context", "Se2pError", "default_context", "device_count", "LIB_PATH", "SE2P_SLP",
    "SE2P_DLP", "SE2P_SLP_PRESSURE", "SE2P_DLP_PRESSURE", "StokesSLP_ewald_2p",
    "StokesDLP_ewald_2p", "StokesSLP_pressure_ewald_2p", "StokesDLP_pressure_ewald_2p",
    "density_q", "ewald_params", "mex_stokes_slp_real", "mex_stokes_slp_kspace", "mex_stokes_dlp_real",
    "mex_stokes_dlp_kspace", "mex_stokes_slp_pressure_real", "mex_stokes_slp_pressure_kspace",
    "mex_stokes_dlp_pressure_real", "mex_stokes_dlp_pressure_kspace",
]
# derived quantities: mex_stokes_{slp,dlp}_{gradient,pressure_grad,vorticity,stress}_
# {real,kspace} and Stokes{SLP,DLP}_{gradient,pressure_grad,vorticity,stress}_ewald_2p
for _layer in ("slp", "dlp"):
    for _q in ("gradient", "pressure_grad", "vorticity", "stress"):
        for _name in (f"mex_stokes_{_layer}_{_q}_real", f"mex_stokes_{_layer}_{_q}_kspace",
                      f"Stokes{_layer.upper()}_{_q}_ewald_2p"):
            globals()[_name] = getattr(_pe, _name)
            __all__.append(_name)
for _k in ("SLP_GRADIENT", "SLP_PRESSURE_GRAD", "SLP_VORTICITY", "SLP_STRESS", "DLP_GRADIENT",
           "DLP_PRESSURE_GRAD", "DLP_VORTICITY", "DLP_STRESS"):
    __all__.append("SE2P_" + _k)
__all__ += ["stokesSLPfmm", "stokesDLPfmm"]
del _layer, _q, _name, _k
