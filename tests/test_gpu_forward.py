"""GPU forward parity: libeik (through the C ABI) against the CPU oracle on the same inputs.

FP64 mode must agree with the oracle's heap FMM to rounding level (the fixed point of the
replayed update is the heap's answer).  Mixed precision (FP64 traveltimes and heap order, the
stencil's quadratic increment in FP32, FP64 within 4 nodes of the source box) must equal the
oracle's restatement of that arithmetic the same way, and the FP64 oracle within the stated
tolerances on every node: receivers 1e-5 relative, grids 1e-4 relative.

The "fp32" precision (FP32 traveltimes) is a throughput mode.  The heap FMM is discontinuous at
acceptance-order near-ties: where a node's tentative value and the value of the neighbour the heap
pops next differ by less than FP32 resolution, FP32 storage can take the other branch and the cone
downstream differs by up to ~1e-2 (random3d seed 2, source 0, node (3,11,17): tentative 3.1343110
vs neighbour 3.13431045, 1.7e-7 apart; 103 of 40320 nodes then differ by more than 1e-4).  That
mode is checked in relative L2, on 99.5 % of the nodes and strictly at the receivers.
"""
import numpy as np
import pytest

from cases import c1_model, random2d, random3d, sphere2d
from eikonalsolvers.grids import ExtraParams, GridRefinementPars

pytestmark = pytest.mark.gpu

CASES = {
    "c1": c1_model,
    "rand2d": random2d,
    "rand3d": random3d,
    "sph2d": sphere2d,
}


def _run(eik, oracle, name, refine, precision, fast=False):
    grd, vel, src, recs = CASES[name]()
    ep = ExtraParams(refinearoundsrc=refine, precision=precision, fast=fast)
    print(f"case={name} refine={refine} precision={precision} grid={grd.grsize} nsrc={len(src)}", flush=True)
    pg, tg = eik.eiktraveltime(vel, grd, src, recs, returntt=True, extraparams=ep)
    po, to = oracle.traveltime(vel, grd, src, recs, returntt=True, extraparams=ep)
    return pg, tg, po, to


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("refine", [False, True])
@pytest.mark.parametrize("fast", [False, True])
def test_forward_fp64_matches_oracle(eik, oracle, name, refine, fast):
    """Default FP64 replays the heap bit for bit (measured difference 0; checked at 1e-12).  With
    fast=True changes below 1e-12 relative stop propagating; a suppressed change can flip a
    decision at an exact heap tie, so that mode is held to the stated grid/receiver tolerances."""
    pg, tg, po, to = _run(eik, oracle, name, refine, "fp64", fast)
    for s in range(len(tg)):
        rel = np.abs(tg[s] - to[s]) / np.maximum(np.abs(to[s]), 1e-300)
        relp = np.abs(pg[s] - po[s]) / np.abs(po[s])
        nbad = int(np.sum(rel > 0))
        print(f"  fast={fast} src {s}: max rel grid diff {rel.max():.3e}, nodes differing: {nbad}/{rel.size}, "
              f"picks {relp.max():.3e}")
        if not fast:
            assert rel.max() <= 1e-12
            np.testing.assert_allclose(pg[s], po[s], rtol=1e-12, atol=0)
        else:
            assert rel.max() <= 1e-4
            assert relp.max() <= 1e-5


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("refine", [False, True])
def test_forward_mixed_within_tolerance(eik, oracle, name, refine):
    pg, tg, po, to = _run(eik, oracle, name, refine, "mixed")
    grd, vel, src, recs = CASES[name]()
    pm, tm = oracle.traveltime(vel, grd, src, recs, returntt=True, extraparams=ExtraParams(refinearoundsrc=refine),
                               precision="mixed")
    for s in range(len(tg)):
        rel = np.abs(tg[s] - to[s]) / np.maximum(np.abs(to[s]), 1e-300)
        relp = np.abs(pg[s] - po[s]) / np.abs(po[s])
        relm = np.abs(tg[s] - tm[s]) / np.maximum(np.abs(tm[s]), 1e-300)
        print(f"  src {s}: vs FP64 grid max rel {rel.max():.3e}, picks {relp.max():.3e}; "
              f"vs the mixed oracle grid max rel {relm.max():.3e}")
        assert relm.max() <= 1e-12
        assert rel.max() <= 1e-4
        assert relp.max() <= 1e-5


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("refine", [False, True])
def test_forward_fp32_within_tolerance(eik, oracle, name, refine):
    pg, tg, po, to = _run(eik, oracle, name, refine, "fp32")
    for s in range(len(tg)):
        rel = np.abs(tg[s] - to[s]) / np.maximum(np.abs(to[s]), 1e-300)
        relp = np.abs(pg[s] - po[s]) / np.abs(po[s])
        l2 = np.linalg.norm(tg[s] - to[s]) / np.linalg.norm(to[s])
        q = float(np.quantile(rel, 0.995))
        print(f"  src {s}: grid max rel {rel.max():.3e}, 99.5% {q:.3e}, rel L2 {l2:.3e}; "
              f"picks max rel {relp.max():.3e}")
        assert l2 <= 1e-4
        assert q <= 1e-4
        assert relp.max() <= 1e-5


def test_settling_stays_within_fp64_tolerance(eik, oracle, monkeypatch):
    """The settling rounds (entered only when a solve circulates rounding-level changes instead of
    converging) stop changes below 1e-12 relative from propagating; forced on from the first
    round, the result must still match the heap FMM far below the stated tolerances."""
    monkeypatch.setenv("EIK_SETTLE_ROUND", "0")
    grd, vel, src, recs = CASES["rand3d"]()
    ep = ExtraParams(refinearoundsrc=True)
    m = eik.EikModel(vel, grd)
    pg, tg = m.traveltime(src, recs, returntt=True, extraparams=ep)
    stage = m.ctx.last_stats()["settle_stage"]
    m.close()
    po, to = oracle.traveltime(vel, grd, src, recs, returntt=True, extraparams=ep)
    for s in range(len(tg)):
        rel = np.abs(tg[s] - to[s]) / np.abs(to[s])
        print(f"  settle stage {stage}; src {s}: grid max rel {rel.max():.3e}")
        assert rel.max() <= 1e-9
        np.testing.assert_allclose(pg[s], po[s], rtol=1e-9, atol=0)
    assert stage >= 1
