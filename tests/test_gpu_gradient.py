"""GPU discrete-adjoint gradients against the CPU oracle on the same inputs.

FP64: the adjoint is the same linear system solved in a different summation order, so the
gradients agree to ~1e-9 relative L2 (checked at 1e-7).  Mixed precision (FP32 stencil increments
on FP64 traveltimes, FP64 adjoint): within 1e-3 relative L2 of FP64 -- the stated gradient
tolerance -- and within 1e-7 of the oracle's mixed restatement.
"""
import numpy as np
import pytest

from cases import random2d, random3d, smooth_field, sphere2d
from eikonalsolvers.grids import ExtraParams, GridRefinementPars

pytestmark = pytest.mark.gpu


def _obs(oracle, vel, grd, src, recs, ep, seed):
    velobs = vel * (1.0 + 0.03 * (smooth_field(grd.grsize, seed, lo=-3, hi=3, base=0.0, amp=1.0)))
    picks = oracle.traveltime(velobs, grd, src, recs, extraparams=ep)
    std = [np.full(len(p), 0.01) for p in picks]
    return picks, std


def _rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


GCASES = {
    "rand2d": lambda: random2d(seed=5, n=(64, 48), nsrc=3, nrec=10),
    "rand3d": lambda: random3d(seed=6, n=(28, 24, 20), nsrc=2, nrec=9),
    "sph2d": sphere2d,
}


@pytest.mark.parametrize("name", list(GCASES))
@pytest.mark.parametrize("refine", [False, True])
@pytest.mark.parametrize("radius", [0, 3])
def test_gradvel_fp64_matches_oracle(eik, oracle, name, refine, radius):
    grd, vel, src, recs = GCASES[name]()
    ep = ExtraParams(refinearoundsrc=refine, radiussmoothgradsrc=radius,
                     grdrefpars=GridRefinementPars(downscalefactor=5 if grd.ndim == 2 else 3, noderadius=3))
    print(f"case={name} refine={refine} radius={radius} grid={grd.grsize} nsrc={len(src)}", flush=True)
    obs, std = _obs(oracle, vel, grd, src, recs, ep, 77)
    gg, mg = eik.eikgradient(vel, grd, src, recs, obs, std, "gradvel", extraparams=ep)
    go, mo = oracle.gradient(vel, grd, src, recs, obs, std, "gradvel", extraparams=ep)
    e = _rel_l2(gg, go)
    print(f"  misfit gpu {mg:.12e} oracle {mo:.12e}; gradvel rel L2 {e:.3e}")
    assert mg == pytest.approx(mo, rel=1e-10)
    assert e <= 1e-7


@pytest.mark.parametrize("name", ["rand2d", "rand3d"])
@pytest.mark.parametrize("refine", [False, True])
def test_gradsrcloc_fp64_matches_oracle(eik, oracle, name, refine):
    grd, vel, src, recs = GCASES[name]()
    ep = ExtraParams(refinearoundsrc=refine,
                     grdrefpars=GridRefinementPars(downscalefactor=5 if grd.ndim == 2 else 3, noderadius=3))
    print(f"case={name} refine={refine} grid={grd.grsize} nsrc={len(src)}", flush=True)
    obs, std = _obs(oracle, vel, grd, src, recs, ep, 78)
    gv, gs, mg = eik.eikgradient(vel, grd, src, recs, obs, std, "gradvelandsrcloc", extraparams=ep)
    ov, os_, mo = oracle.gradient(vel, grd, src, recs, obs, std, "gradvelandsrcloc", extraparams=ep)
    print(f"  gradsrc gpu {gs.ravel()[:4]} oracle {os_.ravel()[:4]}; rel L2 {_rel_l2(gs, os_):.3e}")
    assert _rel_l2(gs, os_) <= 1e-7
    assert _rel_l2(gv, ov) <= 1e-7


@pytest.mark.parametrize("name,refine", [("rand2d", False), ("rand2d", True), ("rand3d", False), ("rand3d", True),
                                         ("sph2d", False), ("sph2d", True)])
def test_gradient_mixed_within_tolerance(eik, oracle, name, refine):
    """Mixed precision against the FP64 oracle within the stated 1e-3 relative L2, and against the
    oracle's restatement of the mixed arithmetic (the same discrete problem) to rounding of the
    adjoint's summation order.  Grid2DSphere runs the FP64 stencil under the mixed precision."""
    grd, vel, src, recs = GCASES[name]()
    base = dict(refinearoundsrc=refine,
                grdrefpars=GridRefinementPars(downscalefactor=5 if grd.ndim == 2 else 3, noderadius=3))
    ep64 = ExtraParams(**base)
    epmx = ExtraParams(**base, precision="mixed")
    which = "gradvelandsrcloc" if grd.kind < 2 else "gradvel"
    obs, std = _obs(oracle, vel, grd, src, recs, ep64, 79)
    res = eik.eikgradient(vel, grd, src, recs, obs, std, which, extraparams=epmx)
    ref = oracle.gradient(vel, grd, src, recs, obs, std, which, extraparams=ep64)
    rmx = oracle.gradient(vel, grd, src, recs, obs, std, which, extraparams=ep64, precision="mixed")
    gv, ov, mv = res[0], ref[0], rmx[0]
    print(f"case={name} refine={refine}: mixed gradvel rel L2 {_rel_l2(gv, ov):.3e} vs FP64, "
          f"{_rel_l2(gv, mv):.3e} vs the mixed oracle; misfit rel {abs(res[-1] - ref[-1]) / ref[-1]:.3e}", flush=True)
    assert np.all(np.isfinite(gv))
    assert _rel_l2(gv, ov) <= 1e-3
    assert _rel_l2(gv, mv) <= 1e-7
    assert res[-1] == pytest.approx(rmx[-1], rel=1e-10)
    if which == "gradvelandsrcloc":
        gs, os_, ms = res[1], ref[1], rmx[1]
        print(f"  gradsrc rel L2 {_rel_l2(gs, os_):.3e} vs FP64, {_rel_l2(gs, ms):.3e} vs the mixed oracle")
        assert _rel_l2(gs, os_) <= 1e-3
        assert _rel_l2(gs, ms) <= 1e-7
