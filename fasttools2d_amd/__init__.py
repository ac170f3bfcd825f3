"""fasttools2d_amd: MI355X-native spectral Ewald sums for doubly-periodic 2D Stokes flow.

The compute path is libse2p.so (hand-written HIP for gfx950 behind a C ABI,
include/se2p.h); this package is its Python front end, mirroring the
reference's PeriodicEwald functions.
"""
from ._lib import (Context, Se2pError, default_context, device_count, LIB_PATH,  # noqa: F401
                   SE2P_SLP, SE2P_DLP, SE2P_SLP_PRESSURE, SE2P_DLP_PRESSURE)
from .periodic_ewald import (StokesDLP_ewald_2p, StokesDLP_pressure_ewald_2p,  # noqa: F401
                             StokesSLP_ewald_2p, StokesSLP_pressure_ewald_2p, ewald_params,
                             mex_stokes_dlp_kspace, mex_stokes_dlp_pressure_kspace,
                             mex_stokes_dlp_pressure_real, mex_stokes_dlp_real,
                             mex_stokes_slp_kspace, mex_stokes_slp_pressure_kspace,
                             mex_stokes_slp_pressure_real, mex_stokes_slp_real)

__all__ = [
    "Context", "Se2pError", "default_context", "device_count", "LIB_PATH", "SE2P_SLP",
    "SE2P_DLP", "SE2P_SLP_PRESSURE", "SE2P_DLP_PRESSURE", "StokesSLP_ewald_2p",
    "StokesDLP_ewald_2p", "StokesSLP_pressure_ewald_2p", "StokesDLP_pressure_ewald_2p",
    "ewald_params", "mex_stokes_slp_real", "mex_stokes_slp_kspace", "mex_stokes_dlp_real",
    "mex_stokes_dlp_kspace", "mex_stokes_slp_pressure_real", "mex_stokes_slp_pressure_kspace",
    "mex_stokes_dlp_pressure_real", "mex_stokes_dlp_pressure_kspace",
]
