"""savemodelvtk (ext/SaveGrid2VTKExt.jl:17-421): XML VTK datasets written and read back.

WriteVTK.jl (the reference's writer) is not available, so the files are checked against what
the reference builds before handing over to it: the coordinate vectors (node coordinates, or
cell vertices x - h/2 ... x[end] + h/2), the spherical-to-Cartesian conversion with the
reference's formulas (evaluated here point by point in plain Python on small grids), the field
values in x-fastest order, the split of an extra dimension into "name  #i" fields, and the
error for an unknown kind.  Byte-level equality with WriteVTK's output is not claimed.
"""
import math
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from eikonalsolvers.grids import Grid2DCart, Grid2DSphere, Grid3DCart, Grid3DSphere
from eikonalsolvers.vtk import readvtk, savemodelvtk


def _field(shape, seed):
    return np.random.default_rng(seed).standard_normal(shape)


@pytest.mark.parametrize("fmt", ["binary", "ascii"])
def test_cartesian_points_and_cells(tmp_path, fmt):
    g3 = Grid3DCart(hgrid=0.5, cooinit=(1.0, 2.0, -1.0), grsize=(5, 4, 3))
    tt = _field(g3.grsize, 1)
    multi = _field(g3.grsize + (2,), 2)
    files = savemodelvtk(str(tmp_path / "m3"), g3, {"tt": tt, "g": multi}, fmt=fmt)
    assert files[0].endswith("m3.vtr")
    ET.parse(files[0])  # well-formed XML
    d = readvtk(files[0])
    assert d["type"] == "RectilinearGrid" and d["extent"] == [0, 4, 0, 3, 0, 2]
    for c, a in zip(d["coords"], g3.axes()):
        np.testing.assert_array_equal(c, a)
    np.testing.assert_array_equal(d["pointdata"]["tt"], tt.ravel(order="F"))
    np.testing.assert_array_equal(d["pointdata"]["g  #2"], multi[..., 1].ravel(order="F"))
    # cells: vertices at half spacings, one more per axis
    files = savemodelvtk(str(tmp_path / "c3.vtr"), g3, {"vel": tt}, kind="cells", fmt=fmt)
    d = readvtk(files[0])
    assert d["extent"] == [0, 5, 0, 4, 0, 3]
    np.testing.assert_array_equal(d["coords"][0], np.array([0.75, 1.25, 1.75, 2.25, 2.75, 3.25]))
    np.testing.assert_array_equal(d["celldata"]["vel"], tt.ravel(order="F"))
    g2 = Grid2DCart(hgrid=2.0, cooinit=(0.0, 0.0), grsize=(3, 2))
    d = readvtk(savemodelvtk(str(tmp_path / "m2"), g2, {"v": np.ones((3, 2))}, fmt=fmt)[0])
    assert d["extent"] == [0, 2, 0, 1, 0, 0]
    np.testing.assert_array_equal(d["coords"][2], [0.0])


def test_sphere_points_and_cells(tmp_path):
    g = Grid3DSphere(Δr=2.0, Δθ=1.5, Δφ=2.5, cooinit=(6300.0, 40.0, 10.0), grsize=(3, 4, 2))
    v = _field(g.grsize, 3)
    d = readvtk(savemodelvtk(str(tmp_path / "s3"), g, {"v": v})[0])
    assert d["type"] == "StructuredGrid" and d["extent"] == [0, 2, 0, 3, 0, 1]
    pts = d["points"]
    k = 0
    for iph in range(2):
        for ith in range(4):
            for ir in range(3):
                r, th, ph = g.r[ir], g.θ[ith], g.φ[iph]
                exp = (r * math.sin(math.radians(th)) * math.cos(math.radians(ph)),
                       r * math.sin(math.radians(th)) * math.sin(math.radians(ph)), r * math.cos(math.radians(th)))
                np.testing.assert_allclose(pts[k], exp, rtol=1e-15, atol=1e-9)
                k += 1
    np.testing.assert_array_equal(d["pointdata"]["v"], v.ravel(order="F"))
    # cell vertices: low faces use c[i] - d/2 (c[n-1] + d/2 last), interior c[i-1] + d/2
    d = readvtk(savemodelvtk(str(tmp_path / "s3c"), g, {"v": v}, kind="cells")[0])
    assert d["extent"] == [0, 3, 0, 4, 0, 2]

    def vert(c, dd, i, low):
        n = len(c)
        if low:
            return c[i] - dd / 2 if i < n else c[n - 1] + dd / 2
        return c[0] - dd / 2 if i == 0 else c[i - 1] + dd / 2

    k = 0
    for c2 in range(3):
        for c1 in range(5):
            for c0 in range(4):
                low = c0 == 0 or c1 == 0 or c2 == 0
                r = vert(g.r, g.Δr, c0, low)
                th = vert(g.θ, g.Δθ, c1, low)
                ph = vert(g.φ, g.Δφ, c2, low)
                exp = (r * math.sin(math.radians(th)) * math.cos(math.radians(ph)),
                       r * math.sin(math.radians(th)) * math.sin(math.radians(ph)), r * math.cos(math.radians(th)))
                np.testing.assert_allclose(d["points"][k], exp, rtol=1e-15, atol=1e-9)
                k += 1
    np.testing.assert_array_equal(d["celldata"]["v"], v.ravel(order="F"))
    g2 = Grid2DSphere(Δr=1.0, Δθ=2.0, grsize=(3, 3), cooinit=(100.0, 30.0))
    d = readvtk(savemodelvtk(str(tmp_path / "s2"), g2, {"v": np.ones((3, 3))})[0])
    np.testing.assert_allclose(d["points"][4], (101.0 * math.sin(math.radians(32.0)),
                                                101.0 * math.cos(math.radians(32.0)), 0.0), rtol=1e-15)


def test_wrong_kind_and_shape(tmp_path):
    g = Grid2DCart(hgrid=1.0, cooinit=(0.0, 0.0), grsize=(3, 2))
    with pytest.raises(ValueError):
        savemodelvtk(str(tmp_path / "x"), g, {"v": np.ones((3, 2))}, kind="nodes")
    with pytest.raises(AssertionError):
        savemodelvtk(str(tmp_path / "x"), g, {"v": np.ones((2, 3))})
