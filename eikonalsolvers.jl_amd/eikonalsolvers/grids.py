"""Grid geometry and configuration types of EikonalSolvers.jl, restated for the Python host.

Mirrors ``src/EikSolv/eikstructs.jl``:

* ``Grid2DCart``   -- eikstructs.jl:33-55
* ``Grid3DCart``   -- eikstructs.jl:78-103
* ``Grid2DSphere`` -- eikstructs.jl:126-158  (r, theta in degrees)
* ``Grid3DSphere`` -- eikstructs.jl:182-219  (r, theta, phi in degrees)
* ``GridRefinementPars`` -- eikstructs.jl:230-235
* ``ExtraParams``  -- eikstructs.jl:246-282, plus ``precision`` (FP64 or FP32 stencil with FP64
  near-source, adjoint and reductions) which has no counterpart in the reference.

Arrays follow the reference's layout: ``vel[i, j(, k)]`` with index 1 (x or r) fastest in memory,
i.e. numpy arrays of shape ``grsize`` in Fortran order.  Constructor argument names and
validation messages follow the reference so that callers see the same ``AssertionError`` /
``ValueError`` (Julia ``ErrorException``) behaviour.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# grid kind codes shared with the C-ABI (include/eik.h)
KIND_CART2D = 0
KIND_CART3D = 1
KIND_SPH2D = 2
KIND_SPH3D = 3


def _range(start: float, step: float, n: int) -> np.ndarray:
    """Node coordinates start + (i-1)*step (Julia ``range(start=..., step=..., length=n)``).

    Exact for the spacings used in practice (binary fractions and integers); for decimal steps
    such as 0.1 deg the last bit can differ from Julia's twice-precision range.
    """
    return start + np.arange(n, dtype=np.float64) * step


class AbstractGridEik:
    kind: int
    ndim: int
    grsize: tuple
    cooinit: tuple

    @property
    def spacing3(self) -> tuple:
        raise NotImplementedError

    @property
    def npts(self) -> int:
        return int(np.prod(self.grsize))

    def axes(self):
        raise NotImplementedError


class Grid2DCart(AbstractGridEik):
    """2D Cartesian grid (eikstructs.jl:33-55)."""

    kind = KIND_CART2D
    ndim = 2

    def __init__(self, *, hgrid: float, cooinit, grsize):
        hgrid = float(hgrid)
        cooinit = tuple(float(c) for c in cooinit)
        grsize = tuple(int(n) for n in grsize)
        assert hgrid > 0.0, "hgrid>0.0"
        assert len(cooinit) == 2 and len(grsize) == 2
        assert all(n > 0 for n in grsize), "all(grsize.>0)"
        self.hgrid = hgrid
        self.cooinit = cooinit
        self.grsize = grsize
        self.x = _range(cooinit[0], hgrid, grsize[0])
        self.y = _range(cooinit[1], hgrid, grsize[1])
        self.cooend = (float(self.x[-1]), float(self.y[-1]))

    @property
    def spacing3(self):
        return (self.hgrid, self.hgrid, self.hgrid)

    def axes(self):
        return (self.x, self.y)

    def __repr__(self):
        return f"Grid2DCart(hgrid={self.hgrid}, cooinit={self.cooinit}, grsize={self.grsize})"


class Grid3DCart(AbstractGridEik):
    """3D Cartesian grid (eikstructs.jl:78-103)."""

    kind = KIND_CART3D
    ndim = 3

    def __init__(self, *, hgrid: float, cooinit, grsize):
        hgrid = float(hgrid)
        cooinit = tuple(float(c) for c in cooinit)
        grsize = tuple(int(n) for n in grsize)
        assert hgrid > 0.0, "hgrid>0.0"
        assert len(cooinit) == 3 and len(grsize) == 3
        assert all(n > 0 for n in grsize), "all(grsize.>0)"
        self.hgrid = hgrid
        self.cooinit = cooinit
        self.grsize = grsize
        self.x = _range(cooinit[0], hgrid, grsize[0])
        self.y = _range(cooinit[1], hgrid, grsize[1])
        self.z = _range(cooinit[2], hgrid, grsize[2])
        self.cooend = (float(self.x[-1]), float(self.y[-1]), float(self.z[-1]))

    @property
    def spacing3(self):
        return (self.hgrid, self.hgrid, self.hgrid)

    def axes(self):
        return (self.x, self.y, self.z)

    def __repr__(self):
        return f"Grid3DCart(hgrid={self.hgrid}, cooinit={self.cooinit}, grsize={self.grsize})"


class Grid2DSphere(AbstractGridEik):
    """2D spherical (polar) grid in (r, theta[deg]) (eikstructs.jl:126-158)."""

    kind = KIND_SPH2D
    ndim = 2

    def __init__(self, *, cooinit, grsize, Δr: float | None = None, Δθ: float | None = None,
                 dr: float | None = None, dtheta: float | None = None):
        dr = float(Δr if Δr is not None else dr)
        dth = float(Δθ if Δθ is not None else dtheta)
        cooinit = tuple(float(c) for c in cooinit)
        grsize = tuple(int(n) for n in grsize)
        assert cooinit[0] >= 0.0, "cooinit[1]>=0.0"
        assert dr > 0.0, "Δr>0.0"
        assert dth > 0.0, "Δθ>0.0"
        assert all(n > 0 for n in grsize), "all(grsize.>0)"
        self.Δr, self.Δθ = dr, dth
        self.cooinit = cooinit
        self.grsize = grsize
        self.r = _range(cooinit[0], dr, grsize[0])
        self.θ = _range(cooinit[1], dth, grsize[1])
        assert np.all((0.0 <= self.θ) & (self.θ <= 180.0)), "all(0.0.<=θ.<=180.0)"
        self.cooend = (float(self.r[-1]), float(self.θ[-1]))

    @property
    def spacing3(self):
        return (self.Δr, self.Δθ, 0.0)

    def axes(self):
        return (self.r, self.θ)

    def __repr__(self):
        return f"Grid2DSphere(Δr={self.Δr}, Δθ={self.Δθ}, cooinit={self.cooinit}, grsize={self.grsize})"


class Grid3DSphere(AbstractGridEik):
    """3D spherical grid in (r, theta[deg], phi[deg]) (eikstructs.jl:182-219).

    The reference constructs this type but its solvers do not run on it; this package gives it its
    own semantics: the phi spacing of the stencil is r*sin(theta)*dphi and distances are chords.
    """

    kind = KIND_SPH3D
    ndim = 3

    def __init__(self, *, cooinit, grsize, Δr: float | None = None, Δθ: float | None = None,
                 Δφ: float | None = None, dr=None, dtheta=None, dphi=None):
        dr = float(Δr if Δr is not None else dr)
        dth = float(Δθ if Δθ is not None else dtheta)
        dph = float(Δφ if Δφ is not None else dphi)
        cooinit = tuple(float(c) for c in cooinit)
        grsize = tuple(int(n) for n in grsize)
        assert cooinit[0] > 0.0, "cooinit[1]>0.0"
        assert all(n > 0 for n in grsize), "all(grsize.>0)"
        self.Δr, self.Δθ, self.Δφ = dr, dth, dph
        self.cooinit = cooinit
        self.grsize = grsize
        self.r = _range(cooinit[0], dr, grsize[0])
        self.θ = _range(cooinit[1], dth, grsize[1])
        self.φ = _range(cooinit[2], dph, grsize[2])
        assert np.all((5.0 <= self.θ) & (self.θ <= 175.0)), "all(5.0.<=θ.<=175.0)"
        assert np.all((0.0 <= self.φ) & (self.φ <= 180.0)), "all(0.0.<=φ.<=180.0)"
        self.cooend = (float(self.r[-1]), float(self.θ[-1]), float(self.φ[-1]))

    @property
    def spacing3(self):
        return (self.Δr, self.Δθ, self.Δφ)

    def axes(self):
        return (self.r, self.θ, self.φ)

    def __repr__(self):
        return (f"Grid3DSphere(Δr={self.Δr}, Δθ={self.Δθ}, Δφ={self.Δφ}, cooinit={self.cooinit}, "
                f"grsize={self.grsize})")


@dataclass(frozen=True)
class GridRefinementPars:
    """Refined grid around the source (eikstructs.jl:230-235)."""

    downscalefactor: int = 5
    noderadius: int = 3


PRECISION_FP64 = 0
PRECISION_MIXED = 1   # FP32 stencil increments on FP64 traveltimes; FP64 near-source init, adjoint, reductions
PRECISION_FP32 = 2    # FP32 traveltimes and stencil; FP64 near-source init, adjoint and reductions
PRECISIONS = {"fp64": PRECISION_FP64, "mixed": PRECISION_MIXED, "fp32": PRECISION_FP32}


@dataclass(frozen=True)
class ExtraParams:
    """Extra parameters (eikstructs.jl:246-282) plus the device-side ``precision`` knob.

    ``parallelkind`` is accepted and validated like the reference but every source runs batched on
    the one GPU.  ``precision`` is ``"fp64"`` (default, the reference's arithmetic), ``"mixed"``
    (FP64 traveltimes, the stencil's quadratic increment in FP32; the heap's order is decided in
    FP64) or ``"fp32"`` (FP32 traveltimes and stencil: a throughput mode whose results can leave the
    stated tolerances downstream of heap near-ties within FP32 resolution).
    ``fast`` (FP64): False (default) reproduces the heap FMM bit for bit; True lets changes below
    1e-12 relative stop propagating (~1/3 faster; a suppressed change can flip a decision at an
    exact heap tie, measured up to 5e-5 relative on a grid, inside the stated tolerances).
    """

    refinearoundsrc: bool = True
    manualGCtrigger: bool = False
    parallelkind: str = "serial"
    radiussmoothgradsrc: int = 3
    smoothgradkern: int = 0
    grdrefpars: GridRefinementPars = field(default_factory=GridRefinementPars)
    precision: str = "fp64"
    fast: bool = False

    def __post_init__(self):
        pk = self.parallelkind.lstrip(":")
        object.__setattr__(self, "parallelkind", pk)
        assert self.radiussmoothgradsrc >= 0, "radiussmoothgradsrc>= 0"
        assert self.smoothgradkern >= 0, "smoothgradkern>=0"
        if self.smoothgradkern > 0 and self.smoothgradkern % 2 == 0:
            raise ValueError("ExtraParams: smoothgradkern must be an odd number")
        if pk not in ("serial", "sharedmem", "distribmem"):
            raise ValueError("ExtraParams(): 'parallelkind' must be one of :serial, :sharedmem or :distribmem")
        if self.precision not in PRECISIONS:
            raise ValueError("ExtraParams(): 'precision' must be 'fp64', 'mixed' or 'fp32'")
