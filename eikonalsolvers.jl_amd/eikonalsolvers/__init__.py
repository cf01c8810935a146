"""EikonalSolvers.jl's public API on a native MI355X (gfx950) engine.

Host-side mirror of the reference's exports (src/EikonalSolvers.jl:23-37) for the traveltime and
gradient path; the computation runs in libeik.so (hand-written HIP, include/eik.h).
"""
from .grids import (ExtraParams, Grid2DCart, Grid2DSphere, Grid3DCart, Grid3DSphere, GridRefinementPars)
from .api import (Context, EikModel, default_context, eikgradient, eikttimemisfit, eiktraveltime)
from .serving import Request, Server, traveltime_batch
from . import hmc, autodiff
from .vtk import savemodelvtk
from .rays import tracerays
from .autodiff import misfit_torch, rrule_eikttimemisfit

HMCtraveltimes = hmc  # the reference's submodule name (src/HMCtraveltimes.jl)

__all__ = [
    "ExtraParams", "GridRefinementPars", "Grid2DCart", "Grid3DCart", "Grid2DSphere", "Grid3DSphere",
    "eiktraveltime", "eikgradient", "eikttimemisfit", "EikModel", "Context", "default_context",
    "Server", "Request", "traveltime_batch", "HMCtraveltimes", "rrule_eikttimemisfit", "misfit_torch", "savemodelvtk", "tracerays",
]
