"""Edge cases of the forward and gradient path through the C ABI (the reference's own tests use
on-node and off-node sources, receivers on the grid border and ragged receiver sets):

* no sources, and sources with no receivers, among others with some (ragged sets);
* sources exactly on a node (one-corner source box), on the grid border and in a corner, and
  receivers on the upper grid edges (the interpolation cell clamped back): FP64 against the
  oracle, bit for bit;
* the reference's validation: positions outside the grid, non-positive velocity, unknown
  whichgrad, an even smoothing kernel;
* a large batch (C2 grid, 64 sources in one call) against the same sources in four calls of 16:
  every source's solve is independent of the others in its launches, so picks are identical.
"""
import numpy as np
import pytest

from cases import random2d, random3d, smooth_field
from eikonalsolvers.grids import ExtraParams, Grid2DCart, Grid3DCart

pytestmark = pytest.mark.gpu


def test_empty_and_ragged(eik, oracle):
    grd, vel, src, recs = random2d(seed=21, n=(50, 40), nsrc=3, nrec=6)
    assert eik.eiktraveltime(vel, grd, np.zeros((0, 2)), []) == []
    recs = [recs[0], np.zeros((0, 2)), recs[2][:1]]
    pg = eik.eiktraveltime(vel, grd, src, recs, extraparams=ExtraParams())
    po = oracle.traveltime(vel, grd, src, recs, extraparams=ExtraParams())
    assert [len(p) for p in pg] == [6, 0, 1]
    for a, b in zip(pg, po):
        np.testing.assert_array_equal(a, b)
    obs = [p + 0.01 for p in po]
    std = [np.full(len(p), 0.02) for p in po]
    gv, gs, m = eik.eikgradient(vel, grd, src, recs, obs, std, "gradvelandsrcloc", extraparams=ExtraParams())
    ov, os_, mo = oracle.gradient(vel, grd, src, recs, obs, std, "gradvelandsrcloc", extraparams=ExtraParams(),
                                  nthreads=1)
    assert m == pytest.approx(mo, rel=1e-12)
    assert np.linalg.norm(gv - ov) <= 1e-10 * np.linalg.norm(ov)
    np.testing.assert_allclose(gs, os_, rtol=1e-9, atol=1e-12 * np.abs(os_).max())


@pytest.mark.parametrize("refine", [False, True])
@pytest.mark.parametrize("nd", [2, 3])
def test_border_and_on_node_positions(eik, oracle, nd, refine):
    if nd == 2:
        grd = Grid2DCart(hgrid=0.5, cooinit=(0.0, 0.0), grsize=(40, 32))
        vel = smooth_field(grd.grsize, 22)
        L = np.array([19.5, 15.5])
        src = np.array([[5.0, 4.5], [0.0, 7.3], [19.5, 15.5], [10.25, 0.0]])  # on node, border, corner, border
        rec = np.array([[19.5, 3.2], [7.7, 15.5], [19.5, 15.5], [0.0, 0.0], [12.3, 8.9]])
    else:
        grd = Grid3DCart(hgrid=1.0, cooinit=(0.0, 0.0, 0.0), grsize=(20, 18, 16))
        vel = smooth_field(grd.grsize, 23, passes=4)
        src = np.array([[5.0, 4.0, 3.0], [0.0, 8.2, 7.1], [19.0, 17.0, 15.0], [11.3, 0.0, 15.0]])
        rec = np.array([[19.0, 3.2, 0.0], [7.7, 17.0, 0.0], [19.0, 17.0, 15.0], [0.0, 0.0, 0.0], [12.3, 8.9, 4.4]])
    recs = [rec.copy() for _ in src]
    ep = ExtraParams(refinearoundsrc=refine)
    pg, tg = eik.eiktraveltime(vel, grd, src, recs, returntt=True, extraparams=ep)
    po, to = oracle.traveltime(vel, grd, src, recs, returntt=True, extraparams=ep)
    for s in range(len(src)):
        np.testing.assert_array_equal(pg[s], po[s])
        rel = np.abs(tg[s] - to[s]) / np.maximum(np.abs(to[s]), 1e-300)
        assert rel.max() <= 1e-12


def test_validation_errors(eik):
    grd, vel, src, recs = random2d(seed=24, n=(30, 20), nsrc=1, nrec=3)
    with pytest.raises(AssertionError):
        eik.eiktraveltime(vel, grd, np.array([[-0.5, 3.0]]), recs)
    with pytest.raises(AssertionError):
        eik.eiktraveltime(vel, grd, src, [np.array([[3.0, 99.0]])])
    bad = vel.copy()
    bad[3, 4] = 0.0
    with pytest.raises(AssertionError):
        eik.eiktraveltime(bad, grd, src, recs)
    obs = [np.ones(3)]
    std = [np.ones(3)]
    with pytest.raises(AssertionError):
        eik.eikgradient(vel, grd, src, recs, obs, std, "gradall")
    with pytest.raises(ValueError):
        eik.eikgradient(vel, grd, src, recs, obs, std, "gradvel", extraparams=ExtraParams(smoothgradkern=4))


def test_large_batch_equals_smaller_calls(eik):
    from eikonalsolvers.synthetic import c2_3d
    grd, vel, src, recs, _ = c2_3d(nsrc=64)
    m = eik.EikModel(vel, grd)
    ep = ExtraParams()
    big = m.traveltime(src, recs, extraparams=ep)
    parts = []
    for k in range(4):
        parts += m.traveltime(src[16 * k:16 * (k + 1)], recs[16 * k:16 * (k + 1)], extraparams=ep)
    print(f"64 sources of the 201x201x101 grid in one call: {m.ctx.last_stats()['rounds']:.0f} rounds")
    for a, b in zip(big, parts):
        np.testing.assert_array_equal(a, b)
    m.close()


@pytest.mark.parametrize("refine", [False, True])
@pytest.mark.parametrize("nd", [2, 3])
def test_symmetric_constant_velocity(eik, oracle, nd, refine):
    """Constant velocity with sources on nodes and halfway between fine nodes: mirror-symmetric
    fronts, so many heap batches hold several equal values, and updates can land exactly on the
    value of the node that triggered them (they pop in the heap's next batch).  Without refinement
    the result is the heap's bit for bit; with it, the refinement stop can fall inside a batch of
    equal values, whose order inside the heap is not reproduced (every tied coarse node is
    transferred): measured 4e-14 relative, checked at 1e-12."""
    if nd == 2:
        grd = Grid2DCart(hgrid=1.0, cooinit=(0.0, 0.0), grsize=(41, 33))
        src = np.array([[20.0, 16.0], [20.5, 16.5], [10.3, 7.0]])
        rec = np.array([[0.0, 0.0], [40.0, 32.0], [5.5, 30.0], [33.3, 2.2]])
    else:
        grd = Grid3DCart(hgrid=1.0, cooinit=(0.0, 0.0, 0.0), grsize=(21, 19, 17))
        src = np.array([[10.0, 9.0, 8.0], [10.5, 9.5, 8.5], [4.3, 9.0, 8.0]])
        rec = np.array([[0.0, 0.0, 0.0], [20.0, 18.0, 16.0], [3.3, 17.0, 0.0], [15.5, 2.2, 9.9]])
    vel = np.full(grd.grsize, 3.0)
    recs = [rec.copy() for _ in src]
    ep = ExtraParams(refinearoundsrc=refine)
    pg, tg = eik.eiktraveltime(vel, grd, src, recs, returntt=True, extraparams=ep)
    po, to = oracle.traveltime(vel, grd, src, recs, returntt=True, extraparams=ep)
    for s in range(len(src)):
        rel = np.abs(tg[s] - to[s]) / np.maximum(np.abs(to[s]), 1e-300)
        print(f"{nd}D refine={refine} src {src[s]}: max rel grid diff {rel.max():.3e}, "
              f"picks {np.max(np.abs(pg[s] - po[s]) / po[s]):.3e}")
        if refine:
            np.testing.assert_allclose(pg[s], po[s], rtol=1e-12, atol=0)
        else:
            np.testing.assert_array_equal(pg[s], po[s])
        assert rel.max() <= 1e-12


@pytest.mark.parametrize("refine", [False, True])
def test_gradient_on_node_and_halfway_sources(eik, oracle, refine):
    """The adjoint at on-node, border and halfway-between-fine-nodes sources (3D), against the
    oracle's gradient (the derivative codes and the system order come from the same replay)."""
    grd = Grid3DCart(hgrid=1.0, cooinit=(0.0, 0.0, 0.0), grsize=(20, 18, 16))
    vel = smooth_field(grd.grsize, 25, passes=4)
    src = np.array([[5.0, 4.3, 3.0], [0.0, 8.2, 7.1], [9.5, 8.5, 7.5], [11.3, 0.0, 15.0]])
    rec = np.array([[19.0, 3.2, 0.0], [7.7, 17.0, 0.0], [12.3, 8.9, 4.4], [2.0, 2.0, 0.0]])
    recs = [rec.copy() for _ in src]
    ep = ExtraParams(refinearoundsrc=refine)
    picks = oracle.traveltime(vel * 1.02, grd, src, recs, extraparams=ep)
    std = [np.full(len(p), 0.01) for p in picks]
    gv, gs, m = eik.eikgradient(vel, grd, src, recs, picks, std, "gradvelandsrcloc", extraparams=ep)
    ov, os_, mo = oracle.gradient(vel, grd, src, recs, picks, std, "gradvelandsrcloc", extraparams=ep, nthreads=1)
    ev = np.linalg.norm(gv - ov) / np.linalg.norm(ov)
    es = np.linalg.norm(gs - os_) / np.linalg.norm(os_)
    print(f"refine={refine}: misfit {m:.10e} vs {mo:.10e}, gradvel rel L2 {ev:.2e}, gradsrc rel L2 {es:.2e}")
    assert m == pytest.approx(mo, rel=1e-12)
    assert ev <= 1e-9 and es <= 1e-9


@pytest.mark.parametrize("df,radius", [(1, 3), (2, 0), (7, 2), (5, 8)])
@pytest.mark.parametrize("nd", [2, 3])
def test_refinement_parameters_and_small_grids(eik, oracle, nd, df, radius):
    """Unusual GridRefinementPars (downscalefactor 1..7, noderadius 0..8: boxes clipped by the
    model on both sides) on small grids with sharp velocity contrasts (many non-causal 2nd-order
    updates, i.e. cascade keys): FP64 against the oracle, bit for bit."""
    from eikonalsolvers.grids import GridRefinementPars
    rng = np.random.default_rng(31 + df + 10 * radius + nd)
    n = (9, 7) if nd == 2 else (7, 6, 5)
    grd = (Grid2DCart(hgrid=1.0, cooinit=(0.0, 0.0), grsize=n) if nd == 2
           else Grid3DCart(hgrid=1.0, cooinit=(0.0, 0.0, 0.0), grsize=n))
    vel = np.where(rng.uniform(size=n) < 0.5, 1.5, 6.0)
    L = np.array([k - 1 for k in n], dtype=float)
    src = rng.uniform(0.0, 1.0, size=(3, nd)) * L
    recs = [rng.uniform(0.0, 1.0, size=(5, nd)) * L for _ in range(3)]
    ep = ExtraParams(refinearoundsrc=True, grdrefpars=GridRefinementPars(downscalefactor=df, noderadius=radius))
    pg, tg = eik.eiktraveltime(vel, grd, src, recs, returntt=True, extraparams=ep)
    po, to = oracle.traveltime(vel, grd, src, recs, returntt=True, extraparams=ep)
    for s in range(3):
        rel = np.abs(tg[s] - to[s]) / np.maximum(np.abs(to[s]), 1e-300)
        assert rel.max() <= 1e-12, (s, rel.max())
        np.testing.assert_allclose(pg[s], po[s], rtol=1e-12, atol=0)
