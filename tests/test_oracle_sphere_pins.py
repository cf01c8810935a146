"""Pins of the oracle's spherical forward (Grid2DSphere, Grid3DSphere) by what the mathematics fixes.

The metric under test: Grid2DSphere spacings (dr, r_i * deg2rad(dtheta)) (eikforward2D.jl:117-118);
Grid3DSphere (dr, r_i * deg2rad(dtheta), r_i * sin(theta_j) * deg2rad(dphi)) -- the engine's own
3D semantics (the reference's 3D sphere drops sin(theta), eikforward3D.jl:170-171, and does not
run).  Source-box distances: law of cosines in degrees (eikutils.jl:47-73); 3D chords.

Checks, in a constant-velocity model where the exact traveltime is chord distance / v:
  * along the source's radial line the heap FMM is exact to rounding: the radial stencil is 1-D
    (the angular neighbours pop later) and 2nd-order differences of a linear function are exact;
  * under grid refinement the maximum error falls towards 0 at first order (the point source
    limits the order), from below 2 % of the largest traveltime;
  * t(2v) = t(v)/2 bit for bit (scaling by a power of two commutes with rounding).
Each check is run again with a known mistake planted in the oracle (degrees for radians, the
node radius dropped from the arc, sin(theta) dropped in 3D, degrees in the source-box law of
cosines) and must then fail: the pins detect exactly the slips a restatement could make.
"""
import numpy as np
import pytest

from eikonalsolvers.grids import ExtraParams, Grid2DSphere, Grid3DSphere

V0 = 5.0


def chord2d(r1, t1, r2, t2):
    return np.sqrt(r1 ** 2 + r2 ** 2 - 2.0 * r1 * r2 * np.cos(np.deg2rad(t1 - t2)))


def cart3(r, t, p):
    t = np.deg2rad(t)
    p = np.deg2rad(p)
    return np.stack([r * np.sin(t) * np.cos(p), r * np.sin(t) * np.sin(p), r * np.cos(t)], -1)


# 2D: r in [6203, 6371] km, theta in [44, 46] deg: the angular arc (~4.4 km per 0.04 deg) is
# comparable to dr = 4 km.  3D: theta around 30 deg, so sin(theta) = 1/2 matters.
def sphere2d(O, level, src, refine, vel_scale=1.0):
    f = 2 ** level
    nr, nt, dr, dt = 42 * f + 1, 50 * f + 1, 4.0 / f, 0.04 / f
    g = Grid2DSphere(Δr=dr, Δθ=dt, grsize=(nr, nt), cooinit=(6203.0, 44.0))
    vel = np.full((nr, nt), V0 * vel_scale)
    _, tt = O.traveltime(vel, g, src, [np.array([[6300.0, 45.0]])], returntt=True,
                         extraparams=ExtraParams(refinearoundsrc=refine))
    r = 6203.0 + np.arange(nr) * dr
    t = 44.0 + np.arange(nt) * dt
    R, T = np.meshgrid(r, t, indexing="ij")
    return tt[0], chord2d(R, T, src[0, 0], src[0, 1]) / V0


def sphere3d(O, level, src, vel_scale=1.0):
    f = 2 ** level
    nr, nt, nph = 24 * f + 1, 40 * f + 1, 40 * f + 1
    dr, dt, dp = 4.0 / f, 0.05 / f, 0.1 / f
    g = Grid3DSphere(Δr=dr, Δθ=dt, Δφ=dp, grsize=(nr, nt, nph), cooinit=(6275.0, 29.0, 10.0))
    vel = np.full((nr, nt, nph), V0 * vel_scale)
    _, tt = O.traveltime(vel, g, src, [np.array([[6300.0, 30.0, 11.0]])], returntt=True,
                         extraparams=ExtraParams(refinearoundsrc=False))
    r = 6275.0 + np.arange(nr) * dr
    t = 29.0 + np.arange(nt) * dt
    p = 10.0 + np.arange(nph) * dp
    R, T, P = np.meshgrid(r, t, p, indexing="ij")
    ex = np.linalg.norm(cart3(R, T, P) - cart3(*src[0]), axis=-1) / V0
    return tt[0].reshape((nr, nt, nph), order="F"), ex


SRC2_ON = np.array([[6203.0 + 20 * 4.0, 44.0 + 25 * 0.04]])       # on a node at every level
SRC2_OFF = np.array([[6203.0 + 20.37 * 4.0, 44.0 + 25.61 * 0.04]])  # off-node at every level
SRC3_ON = np.array([[6275.0 + 12 * 4.0, 29.0 + 20 * 0.05, 10.0 + 20 * 0.1]])
SRC3_OFF = np.array([[6275.0 + 12.3 * 4.0, 29.0 + 20.45 * 0.05, 10.0 + 19.6 * 0.1]])


def radial_line_2d(O, refine):
    tt, ex = sphere2d(O, 0, SRC2_ON, refine)
    return np.max(np.abs(tt[:, 25] - ex[:, 25]) / np.max(ex))


def radial_line_3d(O):
    tt, ex = sphere3d(O, 0, SRC3_ON)
    return np.max(np.abs(tt[:, 20, 20] - ex[:, 20, 20]) / np.max(ex))


def converges_2d(O, src, refine):
    errs = []
    for lev in range(3):
        tt, ex = sphere2d(O, lev, src, refine)
        f = 2 ** lev
        errs.append(np.max(np.abs(tt[::f, ::f] - ex[::f, ::f])) / np.max(ex))
    return errs


def converges_3d(O, src):
    errs = []
    for lev in range(2):
        tt, ex = sphere3d(O, lev, src)
        f = 2 ** lev
        errs.append(np.max(np.abs(tt[::f, ::f, ::f] - ex[::f, ::f, ::f])) / np.max(ex))
    return errs


def first_order(errs, first):
    """Errors start below `first`, fall with every halving of the spacings, and overall at least
    like h^0.7 (an off-node source's position inside its cell differs between levels, so single
    steps vary: measured 0.0118, 0.0048, 0.0036 for the 2D off-node source)."""
    overall = 0.6 ** (len(errs) - 1)
    return errs[0] < first and all(b < a for a, b in zip(errs, errs[1:])) and errs[-1] <= overall * errs[0]


# ------------------------------------------------------------------ the oracle as written

@pytest.mark.parametrize("refine", [False, True])
def test_sphere2d_radial_line_is_exact(oracle, refine):
    # no refinement: rounding of ~16 s traveltimes; refinement: the fine box's own node positions
    assert radial_line_2d(oracle, refine) <= (1e-13 if not refine else 1e-11)


def test_sphere3d_radial_line_is_exact(oracle):
    assert radial_line_3d(oracle) <= 1e-13


@pytest.mark.parametrize("refine", [False, True])
@pytest.mark.parametrize("src", [SRC2_ON, SRC2_OFF], ids=["on-node", "off-node"])
def test_sphere2d_converges_to_chord_over_v(oracle, src, refine):
    errs = converges_2d(oracle, src, refine)
    assert first_order(errs, 0.02), errs
    assert errs[-1] < 4e-3, errs


@pytest.mark.parametrize("src", [SRC3_ON, SRC3_OFF], ids=["on-node", "off-node"])
def test_sphere3d_converges_to_chord_over_v(oracle, src):
    errs = converges_3d(oracle, src)
    assert first_order(errs, 0.025), errs
    assert errs[-1] < 0.012, errs


def test_sphere_velocity_scaling_is_exact(oracle):
    for refine in (False, True):
        a, _ = sphere2d(oracle, 0, SRC2_OFF, refine)
        b, _ = sphere2d(oracle, 0, SRC2_OFF, refine, vel_scale=2.0)
        np.testing.assert_array_equal(a, 2.0 * b)
    a, _ = sphere3d(oracle, 0, SRC3_OFF)
    b, _ = sphere3d(oracle, 0, SRC3_OFF, vel_scale=2.0)
    np.testing.assert_array_equal(a, 2.0 * b)


# ------------------------------------------------------------------ planted mistakes are caught

def pins_2d_hold(O):
    ok = radial_line_2d(O, False) <= 1e-13 and radial_line_2d(O, True) <= 1e-11
    for src in (SRC2_ON, SRC2_OFF):
        for refine in (False, True):
            ok = ok and first_order(converges_2d(O, src, refine), 0.02)
    return ok


def pins_3d_hold(O):
    ok = radial_line_3d(O) <= 1e-13
    for src in (SRC3_ON, SRC3_OFF):
        ok = ok and first_order(converges_3d(O, src), 0.025)
    return ok


@pytest.mark.parametrize("mistake", ["MISTAKE_DEGREES", "MISTAKE_NO_RADIUS", "MISTAKE_INIT_DEGREES"])
def test_sphere2d_pins_catch_planted_mistakes(oracle, mistake):
    with oracle.planted_mistake(getattr(oracle, mistake)):
        assert not pins_2d_hold(oracle)
    assert pins_2d_hold(oracle)  # and the hook is off again


@pytest.mark.parametrize("mistake", ["MISTAKE_DEGREES", "MISTAKE_NO_RADIUS", "MISTAKE_NO_SINTHETA"])
def test_sphere3d_pins_catch_planted_mistakes(oracle, mistake):
    with oracle.planted_mistake(getattr(oracle, mistake)):
        assert not pins_3d_hold(oracle)
