"""Problem callables for Hamiltonian Monte Carlo inversion (HMCtraveltimes.jl:38-283).

``EikonalProbVel``, ``EikonalProbSrcLoc`` and ``EikonalProbVelAndSrcLoc`` hold the fixed part of a
traveltime inversion (grid, observations, geometry) and are called as ``prob(x, kind)``:

* ``kind = "nlogpdf"``     -> the misfit ``eikttimemisfit`` at ``x`` (a float);
* ``kind = "gradnlogpdf"`` -> its gradient, flattened like Julia's ``vec`` (column-major).

``x`` is the flattened velocity (``vec(vel)``, column-major), the flattened source coordinates
(``vec(coordsrc)`` of the ``nsrc x ndim`` array: all first coordinates, then all second ...), or
both concatenated, as in the reference.  Any other ``kind`` raises ``ValueError`` (the
reference's ``error(...)``); a problem built with another ``whichgrad`` raises
``AssertionError`` (its ``@assert``).  ``logVel=True`` means ``x`` holds ``log(vel)``; the
gradient is then multiplied by ``1 / vel`` exactly as the reference does
(HMCtraveltimes.jl:101-104 and 269-272; the chain rule of ``vel = exp(x)`` would multiply by
``vel``), so results stay interchangeable with it.

Unlike the reference, which re-creates every solver buffer on each call, a problem keeps one
``EikModel`` resident in HBM across calls (refreshed with ``update_vel`` when the velocity is a
parameter) and ``nlogpdf_and_grad(x)`` returns both values from a single adjoint call, the pair a
leapfrog step needs.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any, Optional

import numpy as np

from .api import EikModel
from .grids import ExtraParams, Grid2DCart, Grid3DCart

_KINDS = ("nlogpdf", "gradnlogpdf")


def _kind(kind, name):
    k = str(kind).lstrip(":")
    if k not in _KINDS:
        raise ValueError(f"eikprob::{name}(): Wrong argument 'kind': {kind}...")
    return k


def _which(w):
    return str(w).lstrip(":")


class _Resident:
    """One resident model per problem, created on first use."""

    _model: Optional[EikModel] = None

    def _model_for(self, vel, fixed=False):
        if self._model is None:
            self._model = EikModel(vel, self.grd)
        elif not fixed:
            self._model.update_vel(vel)
        return self._model

    def close(self):
        if self._model is not None:
            self._model.close()
            self._model = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _vel_from(x, grd, logvel):
    v = np.exp(x) if logvel else np.asarray(x, dtype=np.float64)
    return v.reshape(grd.grsize, order="F")


@dataclass(eq=False)
class EikonalProbVel(_Resident):
    """Velocity as the HMC parameter (HMCtraveltimes.jl:38-112); whichgrad must be :gradvel."""

    grd: Any
    dobs: list
    stdobs: list
    coordsrc: np.ndarray
    coordrec: list
    whichgrad: str
    logVel: bool = False
    extraparams: Optional[ExtraParams] = None
    _model: Optional[EikModel] = field(default=None, repr=False)

    def __post_init__(self):
        assert isinstance(self.grd, (Grid2DCart, Grid3DCart)), "grd::Union{Grid2DCart,Grid3DCart}"

    def __call__(self, inpvecvel, kind):
        assert _which(self.whichgrad) == "gradvel"
        velnd = _vel_from(inpvecvel, self.grd, self.logVel)
        k = _kind(kind, "EikonalProbVel")
        m = self._model_for(velnd)
        if k == "nlogpdf":
            return m.misfit(self.coordsrc, self.coordrec, self.dobs, self.stdobs, extraparams=self.extraparams)
        grad, _ = m.gradient(self.coordsrc, self.coordrec, self.dobs, self.stdobs, "gradvel",
                             extraparams=self.extraparams)
        if self.logVel:
            grad = (1.0 / velnd) * grad
        return grad.ravel(order="F")

    def nlogpdf_and_grad(self, inpvecvel):
        """(nlogpdf, gradnlogpdf) from one adjoint call."""
        assert _which(self.whichgrad) == "gradvel"
        velnd = _vel_from(inpvecvel, self.grd, self.logVel)
        grad, misf = self._model_for(velnd).gradient(self.coordsrc, self.coordrec, self.dobs, self.stdobs, "gradvel",
                                                     extraparams=self.extraparams)
        if self.logVel:
            grad = (1.0 / velnd) * grad
        return misf, grad.ravel(order="F")


@dataclass(eq=False)
class EikonalProbSrcLoc(_Resident):
    """Source coordinates as the HMC parameter (HMCtraveltimes.jl:117-184); the velocity model
    is fixed and stays resident.  whichgrad must be :gradsrcloc."""

    grd: Any
    dobs: list
    stdobs: list
    coordrec: list
    velmod: np.ndarray
    whichgrad: str
    extraparams: Optional[ExtraParams] = None
    _model: Optional[EikModel] = field(default=None, repr=False)

    def __post_init__(self):
        assert isinstance(self.grd, (Grid2DCart, Grid3DCart)), "grd::Union{Grid2DCart,Grid3DCart}"

    def _coosrc(self, inpcoosrc):
        nd = np.ndim(self.velmod)
        return np.asarray(inpcoosrc, dtype=np.float64).reshape((-1, nd), order="F")

    def __call__(self, inpcoosrc, kind):
        assert _which(self.whichgrad) == "gradsrcloc"
        coosrcnd = self._coosrc(inpcoosrc)
        k = _kind(kind, "EikonalProbSrcLoc")
        m = self._model_for(self.velmod, fixed=True)
        if k == "nlogpdf":
            return m.misfit(coosrcnd, self.coordrec, self.dobs, self.stdobs, extraparams=self.extraparams)
        gradsrcloc, _ = m.gradient(coosrcnd, self.coordrec, self.dobs, self.stdobs, "gradsrcloc",
                                   extraparams=self.extraparams)
        return gradsrcloc.ravel(order="F")

    def nlogpdf_and_grad(self, inpcoosrc):
        assert _which(self.whichgrad) == "gradsrcloc"
        gradsrcloc, misf = self._model_for(self.velmod, fixed=True).gradient(
            self._coosrc(inpcoosrc), self.coordrec, self.dobs, self.stdobs, "gradsrcloc", extraparams=self.extraparams)
        return misf, gradsrcloc.ravel(order="F")


@dataclass(eq=False)
class EikonalProbVelAndSrcLoc(_Resident):
    """Velocity and source coordinates together (HMCtraveltimes.jl:189-283): x = [vec(vel);
    vec(coordsrc)].  whichgrad must be :gradvelandsrcloc."""

    grd: Any
    dobs: list
    stdobs: list
    coordrec: list
    whichgrad: str
    logVel: bool = False
    extraparams: Optional[ExtraParams] = None
    _model: Optional[EikModel] = field(default=None, repr=False)

    def __post_init__(self):
        assert isinstance(self.grd, (Grid2DCart, Grid3DCart)), "grd::Union{Grid2DCart,Grid3DCart}"

    def _split(self, inppars):
        x = np.asarray(inppars, dtype=np.float64)
        n = int(np.prod(self.grd.grsize))
        velnd = _vel_from(x[:n], self.grd, self.logVel)
        coosrcnd = x[n:].reshape((-1, velnd.ndim), order="F")
        return velnd, coosrcnd

    def __call__(self, inppars, kind):
        assert _which(self.whichgrad) == "gradvelandsrcloc"
        velnd, coosrcnd = self._split(inppars)
        k = _kind(kind, "EikonalVelAndSrcLocProb")
        m = self._model_for(velnd)
        if k == "nlogpdf":
            return m.misfit(coosrcnd, self.coordrec, self.dobs, self.stdobs, extraparams=self.extraparams)
        return self.nlogpdf_and_grad(inppars, _split=(velnd, coosrcnd))[1]

    def nlogpdf_and_grad(self, inppars, _split=None):
        assert _which(self.whichgrad) == "gradvelandsrcloc"
        velnd, coosrcnd = _split if _split is not None else self._split(inppars)
        gradvel, gradsrcloc, misf = self._model_for(velnd, fixed=_split is not None).gradient(
            coosrcnd, self.coordrec, self.dobs, self.stdobs, "gradvelandsrcloc", extraparams=self.extraparams)
        if self.logVel:
            gradvel = (1.0 / velnd) * gradvel
        return misf, np.concatenate([gradvel.ravel(order="F"), gradsrcloc.ravel(order="F")])
