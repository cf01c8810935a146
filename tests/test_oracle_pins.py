"""Pins of the oracle that the reference's tables leave open: closed forms, invariants and brute
force on tiny inputs, each chosen so that a plausible slip (wrong sign, dropped term, swapped
index, wrong tie rule) fails."""
import numpy as np
import pytest
from scipy import ndimage

from eikonalsolvers.grids import ExtraParams, Grid2DCart, Grid3DCart

NOREF = ExtraParams(refinearoundsrc=False)


def _rand_vel(shape, seed):
    rng = np.random.default_rng(seed)
    return 1.5 + 3.0 * rng.random(shape)


# ---------------------------------------------------------------- heap FMM vs brute force
@pytest.mark.parametrize("seed", range(6))
def test_heap_matches_bruteforce_2d(oracle, seed):
    grd = Grid2DCart(hgrid=1.0, cooinit=(0.0, 0.0), grsize=(10, 9))
    vel = _rand_vel(grd.grsize, seed)
    rng = np.random.default_rng(100 + seed)
    src = rng.uniform(0.5, 7.5, size=2)
    tt_h, order_h, _, nsp, nacc = oracle.fmm_record(vel, grd, src, NOREF)
    tt_b, order_b = oracle.fmm_bruteforce(vel, grd, src)
    assert np.array_equal(tt_h, tt_b)
    assert nacc == grd.npts
    # same order up to permutations inside equal-value batches
    assert np.array_equal(tt_h[order_h], tt_b[order_b])


@pytest.mark.parametrize("seed", range(4))
def test_heap_matches_bruteforce_3d(oracle, seed):
    grd = Grid3DCart(hgrid=1.0, cooinit=(0.0, 0.0, 0.0), grsize=(6, 6, 5))
    vel = _rand_vel(grd.grsize, seed + 50)
    src = np.array([2.3, 2.6, 1.7]) + 0.1 * seed
    tt_h, order_h, _, _, _ = oracle.fmm_record(vel, grd, src, NOREF)
    tt_b, order_b = oracle.fmm_bruteforce(vel, grd, src)
    assert np.array_equal(tt_h, tt_b)
    assert np.array_equal(tt_h[order_h], tt_b[order_b])


def test_heap_order_is_monotone_and_codes_point_upwind(oracle):
    grd = Grid2DCart(hgrid=0.5, cooinit=(0.0, 0.0), grsize=(40, 30))
    vel = 2.0 + 0.05 * np.arange(30)[None, :] * np.ones((40, 1))
    tt, order, codes, nsp, nacc = oracle.fmm_record(vel, grd, np.array([7.3, 5.1]), NOREF)
    seq = tt[order[:nacc]]
    assert np.all(np.diff(seq[nsp:]) >= 0)
    # every non-source node's stencil uses only nodes accepted before it
    pos = np.empty(grd.npts, dtype=np.int64)
    pos[order[:nacc]] = np.arange(nacc)
    nx = grd.grsize[0]
    for l in order[nsp:nacc]:
        i, j = l % nx, l // nx
        for ax, c in enumerate(codes[l]):
            if c == 0:
                continue
            step = -np.sign(c)
            for p in range(1, abs(c) + 1):
                q = [i, j]
                q[ax] += step * p
                assert pos[q[0] + nx * q[1]] < pos[l]


# ---------------------------------------------------------------- stencil closed forms
def _stencil_state(shape):
    return np.full(shape, np.inf), np.zeros(shape, dtype=np.uint8)


def test_stencil_1d_first_order(oracle):
    grd = Grid2DCart(hgrid=0.7, cooinit=(0.0, 0.0), grsize=(5, 5))
    vel = 2.5 * np.ones(grd.grsize)
    tt, st = _stencil_state(grd.grsize)
    tt[1, 2], st[1, 2] = 3.0, 2  # one accepted neighbour on the -x side, 2-away not accepted
    t, codes = oracle.stencil(grd, vel, tt, st, (2, 2))
    assert t == pytest.approx(3.0 + 0.7 / 2.5, rel=1e-15)
    assert list(codes) == [1, 0]


def test_stencil_2d_first_order_equal_neighbours(oracle):
    grd = Grid2DCart(hgrid=1.0, cooinit=(0.0, 0.0), grsize=(5, 5))
    vel = 2.0 * np.ones(grd.grsize)
    tt, st = _stencil_state(grd.grsize)
    tt[3, 2], st[3, 2] = 1.0, 2  # +x side
    tt[2, 1], st[2, 1] = 1.0, 2  # -y side
    t, codes = oracle.stencil(grd, vel, tt, st, (2, 2))
    assert t == pytest.approx(1.0 + 0.5 / np.sqrt(2.0), rel=1e-14)
    assert list(codes) == [-1, 1]


def test_stencil_second_order_one_axis(oracle):
    grd = Grid2DCart(hgrid=1.0, cooinit=(0.0, 0.0), grsize=(7, 7))
    vel = 4.0 * np.ones(grd.grsize)
    tt, st = _stencil_state(grd.grsize)
    tt[2, 3], st[2, 3] = 2.0, 2   # -x neighbour
    tt[1, 3], st[1, 3] = 1.8, 2   # -x 2-away, t2 <= t1 -> 2nd order
    t, codes = oracle.stencil(grd, vel, tt, st, (3, 3))
    assert t == pytest.approx((4 * 2.0 - 1.8) / 3 + 2 * 1.0 * 0.25 / 3, rel=1e-14)
    assert list(codes) == [2, 0]


def test_stencil_boundary_plane_blocks_second_order(oracle):
    """isonb2nd tests the first neighbour against ANY boundary plane (eikutils.jl:137)."""
    grd = Grid2DCart(hgrid=1.0, cooinit=(0.0, 0.0), grsize=(7, 7))
    vel = 4.0 * np.ones(grd.grsize)
    tt, st = _stencil_state(grd.grsize)
    tt[2, 0], st[2, 0] = 2.0, 2   # node (3,0) on the y=0 face: x-neighbour is on a boundary plane
    tt[1, 0], st[1, 0] = 1.8, 2
    t, codes = oracle.stencil(grd, vel, tt, st, (3, 0))
    assert t == pytest.approx(2.0 + 0.25, rel=1e-15)
    assert list(codes) == [1, 0]


def test_stencil_stale_second_order_quirk(oracle):
    """+x side sets 2nd order; the smaller -x side has no usable 2-away node: the value keeps the
    +x t2 while the code says 1st order (eikforward2D.jl:172-225)."""
    grd = Grid2DCart(hgrid=1.0, cooinit=(0.0, 0.0), grsize=(7, 7))
    vel = 1.0 * np.ones(grd.grsize)
    tt, st = _stencil_state(grd.grsize)
    tt[2, 3], st[2, 3] = 5.0, 2   # +x neighbour of (1,3)
    tt[3, 3], st[3, 3] = 4.9, 2   # +x 2-away
    tt[0, 3], st[0, 3] = 4.0, 2   # -x neighbour (its 2-away is outside)
    t, codes = oracle.stencil(grd, vel, tt, st, (1, 3))
    tmpa2 = 1.0 / 3.0 * (4.0 * 4.0 - 4.9)
    assert t == pytest.approx(tmpa2 + 2.0 / 3.0, rel=1e-14)
    assert list(codes) == [1, 0]


def test_stencil_negative_discriminant_falls_back(oracle):
    grd = Grid2DCart(hgrid=1.0, cooinit=(0.0, 0.0), grsize=(5, 5))
    vel = 10.0 * np.ones(grd.grsize)  # s*h = 0.1
    tt, st = _stencil_state(grd.grsize)
    tt[3, 2], st[3, 2] = 0.0, 2
    tt[2, 3], st[2, 3] = 5.0, 2   # far apart: (tx-ty)^2 > 2 s^2 h^2
    oracle.reset_warnings()
    t, codes = oracle.stencil(grd, vel, tt, st, (2, 2))
    assert t == pytest.approx(0.1, rel=1e-14)
    assert list(codes) == [-1, 0]
    assert oracle.warning_count() == 1


def test_velocity_scaling_invariance(oracle):
    grd = Grid2DCart(hgrid=1.0, cooinit=(0.0, 0.0), grsize=(30, 25))
    vel = _rand_vel(grd.grsize, 7)
    src = np.array([[11.3, 7.9]])
    rec = [np.array([[2.5, 20.5], [28.0, 1.0]])]
    _, t1 = oracle.traveltime(vel, grd, src, rec, returntt=True, extraparams=NOREF)
    _, t2 = oracle.traveltime(2.0 * vel, grd, src, rec, returntt=True, extraparams=NOREF)
    np.testing.assert_allclose(t2[0], t1[0] / 2.0, rtol=1e-13)


# ---------------------------------------------------------------- receivers, taper, smoothing
def test_trilinear_f101_weight(oracle):
    """trilinear_interp uses xd*(1-yd)*(1-zd) for corner f101 (eikforward3D.jl:102,133)."""
    grd = Grid3DCart(hgrid=1.0, cooinit=(0.0, 0.0, 0.0), grsize=(8, 8, 8))
    vel = 2.0 * np.ones(grd.grsize)
    src = np.array([[1.0, 1.0, 1.0]])
    pt = np.array([[3.25, 4.5, 2.75]])
    picks, tt = oracle.traveltime(vel, grd, src, [pt], returntt=True, extraparams=NOREF)
    f = tt[0]
    xd, yd, zd = 0.25, 0.5, 0.75
    i, j, k = 3, 4, 2
    w = {(0, 0, 0): (1 - xd) * (1 - yd) * (1 - zd), (1, 0, 0): xd * (1 - yd) * (1 - zd),
         (0, 1, 0): (1 - xd) * yd * (1 - zd), (0, 0, 1): (1 - xd) * (1 - yd) * zd,
         (1, 1, 0): xd * yd * (1 - zd), (0, 1, 1): (1 - xd) * yd * zd,
         (1, 0, 1): xd * (1 - yd) * (1 - zd), (1, 1, 1): xd * yd * zd}
    expect = sum(f[i + a, j + b, k + c] * wt for (a, b, c), wt in w.items())
    assert picks[0][0] == pytest.approx(expect, rel=1e-14)
    correct = expect - f[i + 1, j, k + 1] * (xd * (1 - yd) * (1 - zd)) + f[i + 1, j, k + 1] * (xd * (1 - yd) * zd)
    assert abs(picks[0][0] - correct) > 1e-6


def test_bilinear_upper_edge(oracle):
    grd = Grid2DCart(hgrid=0.5, cooinit=(1.0, 2.0), grsize=(10, 8))
    vel = 3.0 * np.ones(grd.grsize)
    src = np.array([[2.0, 3.0]])
    pt = np.array([[grd.x[-1], grd.y[-1]]])
    picks, tt = oracle.traveltime(vel, grd, src, [pt], returntt=True, extraparams=NOREF)
    assert picks[0][0] == pytest.approx(tt[0][-1, -1], rel=1e-15)


def test_taper_closed_form(oracle):
    """smoothgradaroundsrc!: gradient x r/rmax within rmax = radius*h (utils.jl:121-172)."""
    grd = Grid2DCart(hgrid=20.0, cooinit=(0.0, 0.0), grsize=(30, 24))
    vel = 2500.0 + 20.0 * np.arange(24)[None, :] * np.ones((30, 1))
    src = np.array([[301.7, 222.9]])
    rec = [np.array([[50.0, 40.0], [500.0, 60.0], [300.0, 440.0]])]
    obs = [np.array([0.1, 0.2, 0.15])]
    std = [np.array([0.01, 0.01, 0.01])]
    g0, _ = oracle.gradient(vel, grd, src, rec, obs, std, "gradvel",
                            extraparams=ExtraParams(refinearoundsrc=False, radiussmoothgradsrc=0))
    g3, _ = oracle.gradient(vel, grd, src, rec, obs, std, "gradvel",
                            extraparams=ExtraParams(refinearoundsrc=False, radiussmoothgradsrc=3))
    X, Y = np.meshgrid(grd.x, grd.y, indexing="ij")
    r = np.sqrt((X - src[0, 0]) ** 2 + (Y - src[0, 1]) ** 2)
    att = np.where(r <= 60.0, r / 60.0, 1.0)
    np.testing.assert_allclose(g3, g0 * att, rtol=1e-13, atol=1e-300)


def test_smoothgradient_is_zero_padded_gaussian(oracle):
    grd = Grid2DCart(hgrid=1.0, cooinit=(0.0, 0.0), grsize=(20, 15))
    img = np.random.default_rng(3).standard_normal(grd.grsize)
    out = oracle.smoothgradient(grd, 5, img)
    l = 5
    s = l / 2 - np.sqrt(l / 2)
    ax = np.arange(l) - l // 2
    R2 = ax[:, None] ** 2 + ax[None, :] ** 2
    ker = np.exp(-R2 / (2 * s))
    ker /= ker.sum()
    ref = ndimage.correlate(img, ker, mode="constant", cval=0.0)
    np.testing.assert_allclose(out, ref, rtol=1e-12, atol=1e-14)


def test_adjoint_backsolve_vs_dense_gradient(oracle):
    """Discrete adjoint vs a dense linearisation: perturb vel along a random direction and compare
    the directional derivative of the misfit (FD, central) with <gradvel, direction>."""
    grd = Grid2DCart(hgrid=10.0, cooinit=(0.0, 0.0), grsize=(24, 18))
    rng = np.random.default_rng(11)
    vel = 2000.0 + 300.0 * rng.random(grd.grsize)
    src = np.array([[117.3, 84.1]])
    rec = [np.column_stack([np.linspace(5, 225, 9), 3.0 * np.ones(9)])]
    obs = [np.full(9, 0.05)]
    std = [np.full(9, 1e-3)]
    ep = ExtraParams(refinearoundsrc=False, radiussmoothgradsrc=0)
    g, _ = oracle.gradient(vel, grd, src, rec, obs, std, "gradvel", extraparams=ep)
    dirn = rng.standard_normal(grd.grsize)
    eps = 1e-3
    mp = oracle.misfit(vel + eps * dirn, grd, src, rec, obs, std, extraparams=ep)
    mm = oracle.misfit(vel - eps * dirn, grd, src, rec, obs, std, extraparams=ep)
    assert (mp - mm) / (2 * eps) == pytest.approx(float(np.sum(g * dirn)), rel=1e-4)
