"""HMC problem callables: argument handling that needs no device (HMCtraveltimes.jl:56-283).

Errors are raised before any solve: a problem built for another ``whichgrad`` fails its
``@assert`` (AssertionError), an unknown ``kind`` its ``error(...)`` (ValueError), and only
Cartesian grids are accepted (the fields' ``Union{Grid2DCart,Grid3DCart}``).  The parameter
vectors are split and reshaped column-major, like Julia's ``reshape``/``vec``.
"""
import numpy as np
import pytest

from eikonalsolvers.grids import Grid2DCart, Grid2DSphere, Grid3DCart
from eikonalsolvers.hmc import EikonalProbSrcLoc, EikonalProbVel, EikonalProbVelAndSrcLoc


def _geom(grd):
    nd = grd.ndim
    return dict(dobs=[np.zeros(2)], stdobs=[np.ones(2)], coordrec=[np.zeros((2, nd))])


def test_errors_before_any_solve():
    grd = Grid2DCart(hgrid=1.0, cooinit=(0.0, 0.0), grsize=(10, 8))
    x = np.ones(80)
    P = EikonalProbVel(grd=grd, coordsrc=np.array([[1.0, 1.0]]), whichgrad="gradvel", **_geom(grd))
    with pytest.raises(ValueError):
        P(x, "logpdf")
    Q = EikonalProbVel(grd=grd, coordsrc=np.array([[1.0, 1.0]]), whichgrad=":gradsrcloc", **_geom(grd))
    with pytest.raises(AssertionError):
        Q(x, "nlogpdf")
    S = EikonalProbSrcLoc(grd=grd, velmod=np.ones((10, 8)), whichgrad="gradvel", **_geom(grd))
    with pytest.raises(AssertionError):
        S(np.ones(2), "nlogpdf")
    S2 = EikonalProbSrcLoc(grd=grd, velmod=np.ones((10, 8)), whichgrad="gradsrcloc", **_geom(grd))
    with pytest.raises(ValueError):
        S2(np.ones(2), "grad")
    J = EikonalProbVelAndSrcLoc(grd=grd, whichgrad="gradvelandsrcloc", **_geom(grd))
    with pytest.raises(ValueError):
        J(np.ones(82), ":gradlogpdf")
    with pytest.raises(AssertionError):
        EikonalProbVel(grd=Grid2DSphere(Δr=1.0, Δθ=1.0, grsize=(10, 8), cooinit=(10.0, 10.0)),
                       coordsrc=np.array([[11.0, 11.0]]), whichgrad="gradvel", **_geom(grd))


def test_parameter_vectors_are_column_major():
    grd = Grid3DCart(hgrid=1.0, cooinit=(0.0, 0.0, 0.0), grsize=(4, 3, 2))
    J = EikonalProbVelAndSrcLoc(grd=grd, whichgrad="gradvelandsrcloc", logVel=True, **_geom(grd))
    vel = np.arange(1.0, 25.0).reshape((4, 3, 2), order="F")
    src = np.array([[0.5, 1.0, 0.25], [2.0, 1.5, 0.75]])
    x = np.concatenate([np.log(vel).ravel(order="F"), src.ravel(order="F")])
    v, s = J._split(x)
    np.testing.assert_allclose(v, vel, rtol=1e-14)
    np.testing.assert_array_equal(s, src)
    S = EikonalProbSrcLoc(grd=grd, velmod=vel, whichgrad="gradsrcloc", **_geom(grd))
    np.testing.assert_array_equal(S._coosrc(src.ravel(order="F")), src)
