"""Traveltime and gradient calls of 3 x 2^22 source-nodes or more split their sources into two halves
solved concurrently on two streams (the context's own and a second one it creates); results must
not depend on that.  FP64, 3D 101x101x51 (h = 1, velocity 3 + 0.02 z plus a Gaussian random
field), 25 sources (13 + 12 on the two streams), 16 top-face receivers each, bit for bit:

* picks and traveltime grids against single-source calls (never split);
* the gradient (misfit, velocity and source-location gradients) against the same call on a
  context held to one stream (eik_ctx_set_streams): a gradient is not a per-source sum of
  single-source gradients, since a source node's value carries over from the previous source
  (calcgradsomesrc), across the two halves too.
"""
import numpy as np
import pytest

from eikonalsolvers.grids import ExtraParams, Grid3DCart
from eikonalsolvers.synthetic import gaussian_random_field, normal, uniform

pytestmark = pytest.mark.gpu

NSRC = 25


@pytest.fixture(scope="module")
def case(eik):
    grd = Grid3DCart(hgrid=1.0, cooinit=(0.0, 0.0, 0.0), grsize=(101, 101, 51))
    z = grd.z[None, None, :]
    vel = np.clip(3.0 + 0.02 * z + gaussian_random_field(grd.grsize, 7301, 8.0, 0.2), 1.5, 6.0)
    u = uniform(7302, 3 * NSRC).reshape(NSRC, 3)
    src = np.column_stack([5.0 + 90.0 * u[:, 0], 5.0 + 90.0 * u[:, 1], 15.0 + 25.0 * u[:, 2]])
    lat = np.linspace(8.0, 92.0, 4)
    rec = np.array([[x, y, 0.0] for y in lat for x in lat])
    recs = [rec.copy() for _ in range(NSRC)]
    assert NSRC * grd.npts >= 3 << 22 > grd.npts  # the whole call splits, a single source does not
    model = eik.EikModel(vel, grd)
    yield grd, vel, src, recs, model
    model.close()


def test_split_traveltime_matches_single_source_calls(case):
    grd, vel, src, recs, model = case
    print(f"split traveltime grid={grd.grsize} nsrc={NSRC}", flush=True)
    ep = ExtraParams()
    picks, grids = model.traveltime(src, recs, returntt=True, extraparams=ep)
    st = model.ctx.last_stats()
    assert 0.0 < st["all_kernel_busy_ms"] <= st["_kernel_ms"] + st["fine_kernel_ms"] + 1e-3
    for s in range(NSRC):
        p1, g1 = model.traveltime(src[s:s + 1], recs[s:s + 1], returntt=True, extraparams=ep)
        assert np.array_equal(picks[s], p1[0]), s
        assert np.array_equal(grids[s], g1[0]), s


def test_split_gradient_matches_one_stream(eik, case):
    grd, vel, src, recs, model = case
    print(f"split gradient grid={grd.grsize} nsrc={NSRC}", flush=True)
    ep = ExtraParams()
    picks = model.traveltime(src, recs, extraparams=ep)
    noise = normal(7303, NSRC * len(recs[0])).reshape(NSRC, -1) * 0.01
    obs = [picks[s] + noise[s] for s in range(NSRC)]
    std = [np.full(len(recs[s]), 0.01) for s in range(NSRC)]
    ctx1 = eik.Context()
    ctx1.set_streams(1)
    model1 = eik.EikModel(vel, grd, ctx=ctx1)
    try:
        for which in ("gradvelandsrcloc", "gradvel"):
            out2 = model.gradient(src, recs, obs, std, which, extraparams=ep)
            out1 = model1.gradient(src, recs, obs, std, which, extraparams=ep)
            for a, b in zip(out2, out1):
                assert np.array_equal(np.asarray(a), np.asarray(b)), which
        gv = out2[0]
        assert np.any(gv != 0.0)
        with pytest.raises(Exception):
            ctx1.set_streams(3)
    finally:
        model1.close()
        ctx1.close()
