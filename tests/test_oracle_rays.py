"""Pins of the ray-tracing restatement (oracle/rays.py), no GPU.

Interpolations.jl (the reference's interpolator) is not available, so its cubic B-spline with the
default Line(OnGrid()) boundary is pinned by what fixes it: the interpolant takes the data values
at the nodes, reproduces linear functions (value and gradient) everywhere, and its coefficients
have a zero second difference at both ends of every line.  The ray loop is pinned by the
constant-velocity case, where the traveltime is the distance to the source over v and rays are
straight segments from receiver to source: on a Cartesian grid and on a Grid2DSphere (straight in
x = r cos(theta), y = r sin(theta)).
"""
import math

import numpy as np
import pytest

from eikonalsolvers.grids import Grid2DCart, Grid2DSphere, Grid3DCart
from oracle import rays as R


@pytest.mark.parametrize("shape", [(6, 5), (5, 4, 6)])
def test_interpolates_nodes_and_linear_functions(shape):
    rng = np.random.default_rng(11)
    y = rng.standard_normal(shape)
    c = R.bspline_coefs(y)
    o, h = [0.5] * len(shape), [0.7] * len(shape)
    for idx in np.ndindex(*shape):
        x = [o[d] + idx[d] * h[d] for d in range(len(shape))]
        assert abs(R.evaluate(c, o, h, x) - y[idx]) <= 1e-12
    slope = np.array([0.3, -1.2, 2.5][:len(shape)])
    grids = np.meshgrid(*[o[d] + h[d] * np.arange(shape[d]) for d in range(len(shape))], indexing="ij")
    lin = 1.5 + sum(slope[d] * grids[d] for d in range(len(shape)))
    cl = R.bspline_coefs(lin)
    for _ in range(20):
        x = [o[d] + rng.uniform(0, (shape[d] - 1) * h[d]) for d in range(len(shape))]
        assert abs(R.evaluate(cl, o, h, x) - (1.5 + float(np.dot(slope, x)))) <= 1e-12
        for d in range(len(shape)):
            assert abs(R.evaluate(cl, o, h, x, d) - slope[d]) <= 1e-12
    # Line(OnGrid()) boundary rows along every axis
    for ax in range(len(shape)):
        m = np.moveaxis(c, ax, 0)
        np.testing.assert_allclose(m[0] - 2 * m[1] + m[2], 0.0, atol=1e-12)
        np.testing.assert_allclose(m[-1] - 2 * m[-2] + m[-3], 0.0, atol=1e-12)


def _seg_dist(p, a, b):
    p, a, b = map(np.asarray, (p, a, b))
    ab = b - a
    t = np.clip(np.dot(p - a, ab) / np.dot(ab, ab), 0.0, 1.0)
    return float(np.linalg.norm(p - (a + t * ab)))


def test_straight_rays_constant_velocity_cartesian():
    grd = Grid2DCart(hgrid=1.0, cooinit=(0.0, 0.0), grsize=(41, 31))
    src = np.array([20.3, 15.7])
    X, Y = np.meshgrid(np.arange(41.0), np.arange(31.0), indexing="ij")
    tt = np.hypot(X - src[0], Y - src[1]) / 2.0
    recs = np.array([[2.0, 3.0], [38.5, 28.0], [5.0, 29.0]])
    rays = R.trace(grd, tt, src, recs, steplen=0.1, max_steps=20000)
    for rec, ray in zip(recs, rays):
        np.testing.assert_array_equal(ray[0], rec)
        np.testing.assert_array_equal(ray[-1], src)
        assert max(_seg_dist(p, rec, src) for p in ray) <= 0.05
        d = float(np.linalg.norm(rec - src))
        assert abs(len(ray) - ((d - math.sqrt(2.0)) / 0.1 + 2)) <= 2


def test_straight_rays_constant_velocity_sphere2d():
    grd = Grid2DSphere(Δr=2.0, Δθ=0.02, grsize=(40, 60), cooinit=(6000.0, 30.0))
    src = np.array([6041.0, 30.61])
    r = 6000.0 + 2.0 * np.arange(40)
    th = 30.0 + 0.02 * np.arange(60)
    Rr, T = np.meshgrid(r, th, indexing="ij")
    tt = np.sqrt(Rr ** 2 + src[0] ** 2 - 2 * Rr * src[0] * np.cos(np.radians(T - src[1]))) / 5.0
    recs = np.array([[6078.0, 30.05], [6010.0, 31.1]])
    rays = R.trace(grd, tt, src, recs, steplen=0.5, max_steps=20000)

    def cart(p):
        return np.array([p[0] * math.cos(math.radians(p[1])), p[0] * math.sin(math.radians(p[1]))])

    for rec, ray in zip(recs, rays):
        a, b = cart(rec), cart(src)
        assert max(_seg_dist(cart(p), a, b) for p in ray) <= 0.3
        np.testing.assert_array_equal(ray[-1], src)


def test_grid3dsphere_raises():
    from eikonalsolvers.grids import Grid3DSphere
    g = Grid3DSphere(Δr=1.0, Δθ=1.0, Δφ=1.0, cooinit=(100.0, 40.0, 10.0), grsize=(4, 4, 4))
    with pytest.raises(ValueError):
        R.trace(g, np.zeros((4, 4, 4)), np.array([101.0, 41.0, 11.0]), np.array([[102.0, 42.0, 12.0]]))
