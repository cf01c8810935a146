"""VTK output of grid fields: ``savemodelvtk`` (ext/SaveGrid2VTKExt.jl:17-421).

The reference writes through WriteVTK.jl (a package extension, not vendored); this module writes
the same XML VTK datasets itself:

* ``Grid2DCart`` / ``Grid3DCart``: a RectilinearGrid (``.vtr``) on the node coordinates
  (``kind="points"``) or on the cell vertices ``x - h/2 ... x[end] + h/2`` (``kind="cells"``);
* ``Grid2DSphere`` / ``Grid3DSphere``: a StructuredGrid (``.vts``) whose points are the nodes
  (``kind="points"``) or the cell vertices (``kind="cells"``) converted to Cartesian coordinates,
  ``x = r sin(theta), y = r cos(theta)`` in 2D and ``x = r sin(theta) cos(phi), y = r sin(theta)
  sin(phi), z = r cos(theta)`` in 3D (angles in degrees).  Vertex coordinates follow the
  reference's construction: vertices on the low faces take ``c[i] - d/2`` per axis (``c[n] + d/2``
  for the last index), interior vertices ``c[i-1] + d/2``.

``datadict`` maps names to arrays of the grid's shape (point data) or of the cell shape, which is
the same here (one cell per node); an array with one more dimension is written as one field per
slice named ``"name  #i"`` (i from 1), like the reference.  Arrays are written x fastest (the
column-major order of the reference), FP64, base64 binary by default or ASCII.  Returns the list
of written files (``vtk_save``).  Any other ``kind`` raises ``ValueError`` (the reference's
``error("savevtk(): Wrong argument 'kind'.")``).
"""
from __future__ import annotations

import base64
import os
import struct

import numpy as np

from .grids import Grid2DCart, Grid2DSphere, Grid3DCart, Grid3DSphere


def _encode(a: np.ndarray, fmt: str) -> str:
    a = np.ascontiguousarray(a, dtype="<f8")
    if fmt == "ascii":
        return " ".join(repr(float(v)) for v in a.ravel())
    raw = a.tobytes()
    return base64.b64encode(struct.pack("<Q", len(raw)) + raw).decode("ascii")


def _array(name, a, fmt, ncomp=1):
    nc = f' NumberOfComponents="{ncomp}"' if ncomp > 1 else ""
    return (f'<DataArray type="Float64" Name="{name}"{nc} format="{fmt}">\n{_encode(a, fmt)}\n</DataArray>\n')


def _fields(datadict, ndim, shape, fmt):
    out = []
    for key, arr in datadict.items():
        a = np.asarray(arr, dtype=np.float64)
        if a.ndim > ndim:
            for i in range(a.shape[-1]):
                sl = a[..., i]
                assert sl.shape == shape, f"size of '{key}' does not match the grid"
                out.append(_array(f"{key}  #{i + 1}", sl.ravel(order="F"), fmt))
        else:
            assert a.shape == shape, f"size of '{key}' does not match the grid"
            out.append(_array(str(key), a.ravel(order="F"), fmt))
    return "".join(out)


def _vertex_axis(c, d, low_face):
    """Cell-vertex coordinates (n + 1) along one axis of a spherical grid."""
    n = len(c)
    v = np.empty(n + 1)
    if low_face:
        v[:n] = c - d / 2.0
        v[n] = c[n - 1] + d / 2.0
    else:
        v[0] = c[0] - d / 2.0
        v[1:] = c + d / 2.0
    return v


def _sph_points(axes, spacing, cells):
    nd = len(axes)
    if not cells:
        grids = np.meshgrid(*axes, indexing="ij")
    else:
        # per vertex: the "minus" form on the low faces, the "plus" form inside
        shape = tuple(len(a) + 1 for a in axes)
        grids = [np.empty(shape) for _ in range(nd)]
        lo = [_vertex_axis(np.asarray(a), s, True) for a, s in zip(axes, spacing)]
        hi = [_vertex_axis(np.asarray(a), s, False) for a, s in zip(axes, spacing)]
        idx = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
        on_low = np.zeros(shape, dtype=bool)
        for ax in range(nd):
            on_low |= idx[ax] == 0
        for ax in range(nd):
            grids[ax] = np.where(on_low, lo[ax][idx[ax]], hi[ax][idx[ax]])
    rad = np.deg2rad
    if nd == 2:
        r, th = grids
        x, y, z = r * np.sin(rad(th)), r * np.cos(rad(th)), np.zeros_like(r)
    else:
        r, th, ph = grids
        x = r * np.sin(rad(th)) * np.cos(rad(ph))
        y = r * np.sin(rad(th)) * np.sin(rad(ph))
        z = r * np.cos(rad(th))
    pts = np.stack([x.ravel(order="F"), y.ravel(order="F"), z.ravel(order="F")], axis=1)
    return pts, grids[0].shape


def savemodelvtk(flname: str, grd, datadict: dict, *, kind: str = "points", fmt: str = "binary") -> list:
    if kind not in ("points", "cells"):
        raise ValueError("savevtk(): Wrong argument 'kind'.")
    assert fmt in ("binary", "ascii")
    cells = kind == "cells"
    nd = grd.ndim
    shape = tuple(grd.grsize)
    if isinstance(grd, (Grid2DCart, Grid3DCart)):
        ext, vtype = ".vtr", "RectilinearGrid"
        h = grd.hgrid
        axes = [np.asarray(a, dtype=np.float64) for a in grd.axes()]
        if cells:
            axes = [np.concatenate([a - h / 2.0, [a[-1] + h / 2.0]]) for a in axes]
        coords = axes + [np.zeros(1)] * (3 - nd)
        npts = tuple(len(a) for a in coords)
        geom = ("<Coordinates>\n" + "".join(_array(n, a, fmt) for n, a in zip(("x", "y", "z"), coords)) +
                "</Coordinates>\n")
    elif isinstance(grd, (Grid2DSphere, Grid3DSphere)):
        ext, vtype = ".vts", "StructuredGrid"
        pts, pshape = _sph_points([np.asarray(a, dtype=np.float64) for a in grd.axes()], grd.spacing3[:nd], cells)
        npts = tuple(pshape) + (1,) * (3 - nd)
        geom = "<Points>\n" + _array("Points", pts, fmt, ncomp=3) + "</Points>\n"
    else:
        raise TypeError(f"savemodelvtk: unsupported grid type {type(grd).__name__}")
    extent = " ".join(f"0 {n - 1}" for n in npts)
    data = _fields(datadict, nd, shape, fmt)
    section = "CellData" if cells else "PointData"
    path = flname if flname.endswith(ext) else flname + ext
    with open(path, "w") as f:
        f.write('<?xml version="1.0"?>\n')
        f.write(f'<VTKFile type="{vtype}" version="1.0" byte_order="LittleEndian" header_type="UInt64">\n')
        f.write(f'<{vtype} WholeExtent="{extent}">\n<Piece Extent="{extent}">\n')
        f.write(f"<{section}>\n{data}</{section}>\n")
        f.write(geom)
        f.write(f"</Piece>\n</{vtype}>\n</VTKFile>\n")
    return [os.path.abspath(path)]


def readvtk(path: str) -> dict:
    """Read back a file written by ``savemodelvtk``: {"type", "extent", "coords" or "points",
    "pointdata", "celldata"} with FP64 arrays (flat, x fastest)."""
    import xml.etree.ElementTree as ET

    root = ET.parse(path).getroot()
    ds = root[0]

    def arr(e):
        txt = (e.text or "").strip()
        if e.get("format") == "ascii":
            return np.array([float(v) for v in txt.split()])
        raw = base64.b64decode(txt)
        n = struct.unpack("<Q", raw[:8])[0]
        return np.frombuffer(raw[8:8 + n], dtype="<f8").copy()

    piece = ds.find("Piece")
    out = {"type": ds.tag, "extent": [int(v) for v in ds.get("WholeExtent").split()], "pointdata": {},
           "celldata": {}}
    for sec, key in (("PointData", "pointdata"), ("CellData", "celldata")):
        e = piece.find(sec)
        if e is not None:
            for a in e.findall("DataArray"):
                out[key][a.get("Name")] = arr(a)
    c = piece.find("Coordinates")
    if c is not None:
        out["coords"] = [arr(a) for a in c.findall("DataArray")]
    p = piece.find("Points")
    if p is not None:
        out["points"] = arr(p.find("DataArray")).reshape(-1, 3)
    return out
