"""smoothgradient on the GPU (k_smooth) against the oracle's og_smoothgradient.

eikgradient smooths the summed gradient with a Gaussian kernel of odd size smoothgradkern when it is
> 0 (eikmaingrad.jl:145-147 -> smoothgradient, utils.jl:73-96: zero padding, the border band of
(l-1)/2 nodes copied unfiltered).  The GPU's smoothed gradient is compared with the oracle's whole
eikgradient (which smooths the same way) and with og_smoothgradient applied to the GPU's own
unsmoothed gradient, for kernels 3, 5 and 7 in 2D and 3D, FP64: the first within 1e-7 relative L2
(the adjoints' summation orders differ), the second to rounding (1e-12).
"""
import numpy as np
import pytest

from cases import random2d, random3d, smooth_field
from eikonalsolvers.grids import ExtraParams

pytestmark = pytest.mark.gpu

CASES = {
    "rand2d": lambda: random2d(seed=15, n=(70, 52), nsrc=3, nrec=10),
    "rand3d": lambda: random3d(seed=16, n=(26, 22, 18), nsrc=2, nrec=9),
}


def _rel_l2(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("kern", [3, 5, 7])
def test_smoothgradient_matches_oracle(eik, oracle, name, kern):
    grd, vel, src, recs = CASES[name]()
    print(f"case={name} smoothgradkern={kern} grid={grd.grsize} nsrc={len(src)}", flush=True)
    velobs = vel * (1.0 + 0.03 * smooth_field(grd.grsize, 91, lo=-3, hi=3, base=0.0, amp=1.0))
    obs = oracle.traveltime(velobs, grd, src, recs)
    std = [np.full(len(p), 0.01) for p in obs]
    ep = ExtraParams(smoothgradkern=kern)
    gg, mg = eik.eikgradient(vel, grd, src, recs, obs, std, "gradvel", extraparams=ep)
    go, mo = oracle.gradient(vel, grd, src, recs, obs, std, "gradvel", extraparams=ep)
    # the same gradient unsmoothed, smoothed by the oracle's smoothgradient
    g0, _ = eik.eikgradient(vel, grd, src, recs, obs, std, "gradvel", extraparams=ExtraParams())
    gs = oracle.smoothgradient(grd, kern, g0)
    e1, e2 = _rel_l2(gg, go), _rel_l2(gg, gs)
    print(f"  smoothed gradvel rel L2 vs oracle eikgradient {e1:.3e}, vs og_smoothgradient(GPU raw) {e2:.3e}; "
          f"smoothing changed it by {_rel_l2(g0, gg):.3e}", flush=True)
    assert mg == pytest.approx(mo, rel=1e-10)
    assert e1 <= 1e-7  # the raw FP64 gradients agree to ~1e-9 (the adjoint sums in another order)
    assert e2 <= 1e-12
    assert _rel_l2(g0, gg) > 1e-3  # the kernel did act
