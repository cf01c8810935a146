"""tracerays on the GPU (eik_tracerays) against the CPU restatement (oracle/rays.py).

Traveltime grids come from the GPU forward solve (FP64) of smooth random models; the same grids
go to both tracers.  The two build the B-spline coefficients with different solvers (device
Thomas sweeps vs a banded LU) and sum the gradient in different orders, so points agree to
rounding accumulated over the steps (1e-8 here) and the point counts are equal.
"""
import numpy as np
import pytest

from cases import smooth_field
from eikonalsolvers.grids import ExtraParams, Grid2DCart, Grid2DSphere, Grid3DCart, Grid3DSphere
from oracle import rays as R

pytestmark = pytest.mark.gpu


def _compare(gpu, ref):
    assert len(gpu) == len(ref)
    for gr, rr in zip(gpu, ref):
        assert gr.shape == rr.shape, (gr.shape, rr.shape)
        np.testing.assert_allclose(gr, rr, rtol=0, atol=1e-8)


@pytest.mark.parametrize("case", ["2d", "3d", "sph2d"])
def test_rays_match_oracle(eik, case):
    if case == "2d":
        grd = Grid2DCart(hgrid=0.5, cooinit=(0.0, 0.0), grsize=(64, 48))
        vel = smooth_field(grd.grsize, 41)
        src = np.array([[20.2, 17.9], [5.3, 3.1]])
        recs = [np.array([[1.0, 22.0], [30.0, 0.5], [31.5, 23.5]]), np.array([[25.0, 20.0], [5.3, 3.3]])]
        steplen = 0.05
    elif case == "3d":
        grd = Grid3DCart(hgrid=1.0, cooinit=(0.0, 0.0, 0.0), grsize=(24, 20, 16))
        vel = smooth_field(grd.grsize, 42, passes=4)
        src = np.array([[11.3, 9.8, 7.1]])
        recs = [np.array([[0.5, 0.5, 0.0], [22.0, 18.5, 0.0], [3.0, 17.0, 15.0]])]
        steplen = 0.1
    else:
        grd = Grid2DSphere(Δr=2.0, Δθ=0.02, grsize=(40, 60), cooinit=(6000.0, 30.0))
        vel = smooth_field(grd.grsize, 43, lo=4.0, hi=8.0, base=6.0, amp=0.5)
        src = np.array([[6041.0, 30.61]])
        recs = [np.array([[6078.0, 30.05], [6010.0, 31.1]])]
        steplen = 0.5
    _, tt = eik.eiktraveltime(vel, grd, src, recs, returntt=True, extraparams=ExtraParams(refinearoundsrc=False))
    gpu = eik.tracerays(grd, tt, src, recs, steplen=steplen)
    npts = []
    for s in range(len(src)):
        ref = R.trace(grd, tt[s], src[s], recs[s], steplen=steplen, max_steps=200000)
        _compare(gpu[s], ref)
        npts += [len(r) for r in ref]
    print(f"{case}: rays of {npts} points match the oracle", flush=True)


def test_rays_errors(eik):
    g3s = Grid3DSphere(Δr=1.0, Δθ=1.0, Δφ=1.0, cooinit=(100.0, 40.0, 10.0), grsize=(6, 6, 6))
    with pytest.raises(ValueError):
        eik.tracerays(g3s, [np.ones((6, 6, 6))], np.array([[101.0, 41.0, 11.0]]), [np.array([[102.0, 42.0, 12.0]])])
    g = Grid2DCart(hgrid=1.0, cooinit=(0.0, 0.0), grsize=(8, 8))
    with pytest.raises(AssertionError):
        eik.tracerays(g, [np.ones((8, 8))], np.array([[3.0, 3.0]]), [np.array([[9.5, 1.0]])])
