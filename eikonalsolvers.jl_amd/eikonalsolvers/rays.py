"""Ray tracing on traveltime grids: ``tracerays`` (src/EikSolv/eikutils.jl:201-328).

``tracerays(grd, ttime, coordsrc, coordrec, steplen=0.01)`` returns, for every source, the list
of its receivers' rays, each an ``npoints x ndim`` array that starts at the receiver and ends at
the source, like the reference.  The rays descend the gradient of the cubic B-spline interpolant
of the source's traveltime grid (``ttime[s]``, e.g. from ``eiktraveltime(..., returntt=True)``);
all rays of all sources are traced at once on the GPU (``eik_tracerays``).  Grid3DSphere raises
``ValueError`` as the reference does.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .api import _grid_struct, default_context


def tracerays(grd, ttime, coordsrc, coordrec, *, steplen: float = 0.01, ctx=None):
    ctx = ctx or default_context()
    nd = grd.ndim
    cs = np.ascontiguousarray(np.asarray(coordsrc, dtype=np.float64).reshape(-1, nd))
    nsrc = cs.shape[0]
    assert len(ttime) == nsrc and len(coordrec) == nsrc, "one traveltime grid and one receiver set per source"
    recs = [np.asarray(r, dtype=np.float64).reshape(-1, nd) for r in coordrec]
    off = np.zeros(nsrc + 1, dtype=np.int64)
    off[1:] = np.cumsum([len(r) for r in recs])
    packed = np.ascontiguousarray(np.concatenate(recs) if recs else np.zeros((0, nd)))
    N = int(np.prod(grd.grsize))
    tt = np.empty(max(nsrc * N, 1))
    for s in range(nsrc):
        a = np.asarray(ttime[s], dtype=np.float64)
        assert a.shape == tuple(grd.grsize), "size(ttime[s])==grsize"
        tt[s * N:(s + 1) * N] = a.ravel(order="F")
    nrec = int(off[-1])
    g = _grid_struct(grd)
    L = _lib.load()
    ray_off = np.zeros(nrec + 1, dtype=np.int64)
    rc = L.eik_tracerays(ctx.handle, ctypes.byref(g), nsrc, _lib.ptr(tt), _lib.ptr(cs), _lib.ptr(off, ctypes.c_int64),
                         _lib.ptr(packed), float(steplen), _lib.ptr(ray_off, ctypes.c_int64), None, 0)
    _lib.check(rc, ctx.handle)
    total = int(ray_off[-1])
    pts = np.zeros((max(total, 1), nd))
    rc = L.eik_tracerays(ctx.handle, ctypes.byref(g), nsrc, _lib.ptr(tt), _lib.ptr(cs), _lib.ptr(off, ctypes.c_int64),
                         _lib.ptr(packed), float(steplen), _lib.ptr(ray_off, ctypes.c_int64), _lib.ptr(pts), total)
    _lib.check(rc, ctx.handle)
    out = []
    for s in range(nsrc):
        out.append([pts[ray_off[r]:ray_off[r + 1]].copy() for r in range(off[s], off[s + 1])])
    return out
