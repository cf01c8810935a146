"""The 3D benchmark configuration on the GPU: 201x201x101 (h = 1), velocity 3 + 0.02 z plus a
Gaussian random field (std 0.2, correlation 15 cells), sources at depth 60-90, 64 top-face
receivers; observed picks from a second perturbed model plus N(0, 0.01^2) noise, std 0.01.

* FP64 :gradvel against the oracle on the first two sources (the oracle's serial source loop,
  whose per-worker buffer carries across sources as the reference's calcgradsomesrc does);
* the FP32-stencil / FP64-adjoint mode against FP64 on all 16 sources: receivers within 1e-5
  relative, the gradient within 1e-3 relative L2, and the grids within 1e-4 relative on all but a
  handful of nodes (the heap's near-tie cascades make single nodes jump; see test_gpu_forward.py).
"""
import numpy as np
import pytest

from eikonalsolvers.grids import ExtraParams
from eikonalsolvers.synthetic import c2_3d, c2_observed

pytestmark = pytest.mark.gpu


def _rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


@pytest.fixture(scope="module")
def c2(eik):
    grd, vel, src, recs, velobs = c2_3d()
    mobs = eik.EikModel(velobs, grd)
    pobs = mobs.traveltime(src, recs, extraparams=ExtraParams())
    mobs.close()
    obs, std = c2_observed(pobs, len(src))
    model = eik.EikModel(vel, grd)
    yield grd, vel, src, recs, obs, std, model
    model.close()


def test_c2_fp64_gradvel_matches_oracle(c2, oracle):
    grd, vel, src, recs, obs, std, model = c2
    print(f"C2 grid={grd.grsize} sources 0-1 of {len(src)}", flush=True)
    gv, mg = model.gradient(src[:2], recs[:2], obs[:2], std[:2], "gradvel", extraparams=ExtraParams())
    ov, mo = oracle.gradient(vel, grd, src[:2], recs[:2], obs[:2], std[:2], "gradvel", extraparams=ExtraParams(),
                             nthreads=1)
    e = _rel_l2(gv, ov)
    print(f"  misfit gpu {mg:.12e} oracle {mo:.12e}; gradvel rel L2 {e:.3e}", flush=True)
    assert mg == pytest.approx(mo, rel=1e-10)
    assert e <= 1e-7


def test_c2_mixed_against_fp64(c2):
    grd, vel, src, recs, obs, std, model = c2
    p64, t64 = model.traveltime(src, recs, returntt=True, extraparams=ExtraParams())
    pmx, tmx = model.traveltime(src, recs, returntt=True, extraparams=ExtraParams(precision="mixed"))
    rp = max(float(np.max(np.abs(a - b) / np.abs(b))) for a, b in zip(pmx, p64))
    t64, tmx = np.asarray(t64), np.asarray(tmx)
    rg = np.abs(tmx - t64) / np.maximum(np.abs(t64), 1e-300)
    nbad = int((rg > 1e-4).sum())
    g64, m64 = model.gradient(src, recs, obs, std, "gradvel", extraparams=ExtraParams())
    gmx, mmx = model.gradient(src, recs, obs, std, "gradvel", extraparams=ExtraParams(precision="mixed"))
    eg = _rel_l2(gmx, g64)
    print(f"C2 mixed vs FP64: picks max rel {rp:.3e}; grid max rel {rg.max():.3e}, nodes > 1e-4: {nbad} of {rg.size}, "
          f"grid rel L2 {_rel_l2(tmx, t64):.3e}; gradvel rel L2 {eg:.3e}; misfit rel {abs(mmx - m64) / m64:.3e}",
          flush=True)
    assert rp <= 1e-5
    assert nbad <= 1e-5 * rg.size
    assert _rel_l2(tmx, t64) <= 1e-5
    assert eg <= 1e-3
    assert mmx == pytest.approx(m64, rel=1e-5)
