"""The oracle's :gradvel on Grid3DSphere, pinned by central finite differences of its own misfit.

The reference raises for Grid3DSphere (createsparsederivativematrices!, eikmainfwd.jl:885-887); the
oracle and the GPU differentiate the engine's own 3D spherical forward (r sin(theta) dphi metric,
chord distances), with the reference's discrete adjoint otherwise unchanged (eikmaingrad.jl).
That adjoint is not the exact derivative of the discrete misfit (the heap's acceptance order
and the derivative codes are held fixed), so finite differences pin it only to ~1e-3: on a
Cartesian 3D grid of the same size the reference's own adjoint differs from them by 8e-3 over
the 20 most sensitive cells and 3e-2 along a smooth direction; here it measures 1.1e-3 and
4.7e-3.  The case lies at theta ~ 30 deg, where dropping the sin(theta) of the phi metric (a
plausible mistake) gives 11 and 0.16.
"""
import numpy as np
import pytest

from cases import smooth_field, sphere3d
from eikonalsolvers.grids import ExtraParams


@pytest.fixture(scope="module")
def case(oracle):
    grd, vel, src, recs = sphere3d()
    ep = ExtraParams(refinearoundsrc=False, radiussmoothgradsrc=0)
    obs = oracle.traveltime(vel, grd, src, recs, extraparams=ep)
    obs = [o * 1.01 + 0.002 * np.sin(np.arange(len(o))) for o in obs]
    std = [np.full(len(o), 0.01) for o in obs]
    return grd, vel, src, recs, obs, std, ep


def test_sphere3d_gradvel_cells(oracle, case):
    grd, vel, src, recs, obs, std, ep = case
    gv, misf = oracle.gradient(vel, grd, src, recs, obs, std, "gradvel", extraparams=ep)
    gv = np.asarray(gv)
    assert misf == pytest.approx(oracle.misfit(vel, grd, src, recs, obs, std, extraparams=ep), rel=1e-12)
    cells = [np.unravel_index(i, gv.shape) for i in np.argsort(-np.abs(gv).ravel())[:20]]
    dh = 1e-5
    fd = []
    for c in cells:
        vp = vel.copy(); vp[c] += dh
        vm = vel.copy(); vm[c] -= dh
        fd.append((oracle.misfit(vp, grd, src, recs, obs, std, extraparams=ep)
                   - oracle.misfit(vm, grd, src, recs, obs, std, extraparams=ep)) / (2 * dh))
    fd = np.array(fd)
    ad = np.array([gv[c] for c in cells])
    assert np.linalg.norm(ad - fd) <= 3e-3 * np.linalg.norm(fd)


def test_sphere3d_gradvel_direction(oracle, case):
    grd, vel, src, recs, obs, std, ep = case
    gv, _ = oracle.gradient(vel, grd, src, recs, obs, std, "gradvel", extraparams=ep)
    d = smooth_field(grd.grsize, 5, lo=-1.0, hi=1.0, base=0.0, amp=0.5)
    eps = 1e-5
    fd = (oracle.misfit(vel + eps * d, grd, src, recs, obs, std, extraparams=ep)
          - oracle.misfit(vel - eps * d, grd, src, recs, obs, std, extraparams=ep)) / (2 * eps)
    assert abs(float(np.sum(np.asarray(gv) * d)) - fd) <= 1e-2 * abs(fd)


def test_sphere3d_gradsrcloc_raises(oracle, case):
    grd, vel, src, recs, obs, std, ep = case
    with pytest.raises(Exception):
        oracle.gradient(vel, grd, src, recs, obs, std, "gradsrcloc", extraparams=ep)
