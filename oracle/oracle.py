"""ctypes wrapper of the CPU oracle (liboracle.so) -- TEST INFRASTRUCTURE ONLY.

The oracle is a plain-C FP64 restatement of EikonalSolvers.jl v0.15.3 (see eik_oracle.c for the
file:line citations).  Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may use it, and only as the checker or as the timed CPU baseline.

Argument conventions mirror the reference API: ``vel`` has shape ``grsize`` (index 1 fastest),
``coordsrc`` is ``nsrc x ndim``, ``coordrec`` / ``pickobs`` / ``stdobs`` are per-source lists.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "liboracle.so")
_lib = None


class OgGrid(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("ndim", ctypes.c_int32), ("n", ctypes.c_int64 * 3),
                ("d", ctypes.c_double * 3), ("o", ctypes.c_double * 3)]


class OgParams(ctypes.Structure):
    _fields_ = [("refine", ctypes.c_int32), ("df", ctypes.c_int32), ("noderadius", ctypes.c_int32),
                ("radiussmooth", ctypes.c_int32), ("smoothkern", ctypes.c_int32),
                ("nthreads", ctypes.c_int32), ("precision", ctypes.c_int32)]


def build() -> str:
    """Compile liboracle.so with the Makefile beside this file."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIBPATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIBPATH):
            build()
        L = ctypes.CDLL(_LIBPATH)
        P = ctypes.POINTER
        dp = P(ctypes.c_double)
        ip = P(ctypes.c_int64)
        L.og_last_error.restype = ctypes.c_char_p
        L.og_warning_count.restype = ctypes.c_int64
        L.og_traveltime.argtypes = [P(OgGrid), dp, ctypes.c_int64, dp, ip, dp, P(OgParams), dp, dp]
        L.og_gradient.argtypes = [P(OgGrid), dp, ctypes.c_int64, dp, ip, dp, dp, dp, ctypes.c_int,
                                  P(OgParams), dp, dp, dp]
        L.og_misfit.argtypes = [P(OgGrid), dp, ctypes.c_int64, dp, ip, dp, dp, dp, P(OgParams), dp]
        L.og_fmm_record.argtypes = [P(OgGrid), dp, dp, P(OgParams), dp, ip, P(ctypes.c_int8), ip]
        L.og_fmm_record.restype = ctypes.c_int64
        L.og_fmm_bruteforce.argtypes = [P(OgGrid), dp, dp, dp, ip]
        L.og_stencil.argtypes = [P(OgGrid), dp, dp, P(ctypes.c_uint8), ip, P(ctypes.c_int32)]
        L.og_stencil.restype = ctypes.c_double
        L.og_smoothgradient.argtypes = [P(OgGrid), ctypes.c_int32, dp, dp]
        L.og_plant_mistake.argtypes = [ctypes.c_int]
        L.og_plant_mistake.restype = ctypes.c_int
        _lib = L
    return _lib


class OracleError(RuntimeError):
    pass


def _ptr(a, ct=ctypes.c_double):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def grid_struct(grd) -> OgGrid:
    g = OgGrid()
    g.kind = grd.kind
    g.ndim = grd.ndim
    for d in range(3):
        g.n[d] = grd.grsize[d] if d < grd.ndim else 1
        g.d[d] = grd.spacing3[d]
        g.o[d] = grd.cooinit[d] if d < grd.ndim else 0.0
    return g


PRECISIONS = {"fp64": 0, "mixed": 1}


def params_struct(extraparams=None, nthreads: int = 1, precision: str = "fp64") -> OgParams:
    """og_params from ExtraParams.  The oracle's arithmetic is chosen by `precision` alone ("fp64",
    the reference, by default; "mixed", the engine's mixed precision restated), never by
    ExtraParams.precision, so a test states which oracle it compares against."""
    p = OgParams()
    if extraparams is None:
        p.refine, p.df, p.noderadius, p.radiussmooth, p.smoothkern = 1, 5, 3, 3, 0
    else:
        p.refine = int(bool(extraparams.refinearoundsrc))
        p.df = int(extraparams.grdrefpars.downscalefactor)
        p.noderadius = int(extraparams.grdrefpars.noderadius)
        p.radiussmooth = int(extraparams.radiussmoothgradsrc)
        p.smoothkern = int(extraparams.smoothgradkern)
    p.nthreads = int(nthreads)
    p.precision = PRECISIONS[precision]
    return p


def _vel(vel, grd):
    v = np.asarray(vel, dtype=np.float64)
    assert v.shape == tuple(grd.grsize), f"vel shape {v.shape} != grsize {grd.grsize}"
    return np.ascontiguousarray(v.ravel(order="F"))


def _pack(coordsrc, coordrec, ndim):
    cs = np.ascontiguousarray(np.asarray(coordsrc, dtype=np.float64).reshape(-1, ndim))
    nsrc = cs.shape[0]
    if len(coordrec) != nsrc:
        raise AssertionError("size(coordsrc,1)==length(coordrec)")
    recs = [np.asarray(r, dtype=np.float64).reshape(-1, ndim) for r in coordrec]
    off = np.zeros(nsrc + 1, dtype=np.int64)
    off[1:] = np.cumsum([r.shape[0] for r in recs])
    packed = np.ascontiguousarray(np.concatenate(recs, axis=0) if nsrc else np.zeros((0, ndim)))
    return cs, off, packed


def _err():
    raise OracleError(lib().og_last_error().decode())


def traveltime(vel, grd, coordsrc, coordrec, returntt=False, extraparams=None, nthreads=1, precision="fp64"):
    L = lib()
    g = grid_struct(grd)
    p = params_struct(extraparams, nthreads, precision)
    v = _vel(vel, grd)
    cs, off, packed = _pack(coordsrc, coordrec, grd.ndim)
    nsrc = cs.shape[0]
    picks = np.zeros(max(int(off[-1]), 1))
    N = grd.npts
    tt = np.zeros(nsrc * N) if returntt else None
    rc = L.og_traveltime(ctypes.byref(g), _ptr(v), nsrc, _ptr(cs), _ptr(off, ctypes.c_int64), _ptr(packed),
                         ctypes.byref(p), _ptr(picks), _ptr(tt) if tt is not None else None)
    if rc != 0:
        _err()
    out = [picks[off[s]:off[s + 1]].copy() for s in range(nsrc)]
    if returntt:
        grids = [tt[s * N:(s + 1) * N].reshape(grd.grsize, order="F").copy() for s in range(nsrc)]
        return out, grids
    return out


def gradient(vel, grd, coordsrc, coordrec, pickobs, stdobs, whichgrad, extraparams=None, nthreads=1,
             precision="fp64"):
    L = lib()
    wg = {"gradvel": 1, "gradsrcloc": 2, "gradvelandsrcloc": 3}[str(whichgrad).lstrip(":")]
    g = grid_struct(grd)
    p = params_struct(extraparams, nthreads, precision)
    v = _vel(vel, grd)
    cs, off, packed = _pack(coordsrc, coordrec, grd.ndim)
    nsrc = cs.shape[0]
    po = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).ravel() for x in pickobs]))
    so = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).ravel() for x in stdobs]))
    gv = np.zeros(grd.npts)
    gs = np.zeros((max(nsrc, 1), grd.ndim))
    misf = np.zeros(1)
    rc = L.og_gradient(ctypes.byref(g), _ptr(v), nsrc, _ptr(cs), _ptr(off, ctypes.c_int64), _ptr(packed),
                       _ptr(po), _ptr(so), wg, ctypes.byref(p), _ptr(gv), _ptr(gs), _ptr(misf))
    if rc != 0:
        _err()
    gvel = gv.reshape(grd.grsize, order="F")
    gsrc = gs[:nsrc].copy()
    if wg == 1:
        return gvel, float(misf[0])
    if wg == 2:
        return gsrc, float(misf[0])
    return gvel, gsrc, float(misf[0])


def misfit(vel, grd, coordsrc, coordrec, pickobs, stdobs, extraparams=None, nthreads=1, precision="fp64"):
    L = lib()
    g = grid_struct(grd)
    p = params_struct(extraparams, nthreads, precision)
    v = _vel(vel, grd)
    cs, off, packed = _pack(coordsrc, coordrec, grd.ndim)
    po = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).ravel() for x in pickobs]))
    so = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).ravel() for x in stdobs]))
    m = np.zeros(1)
    rc = L.og_misfit(ctypes.byref(g), _ptr(v), cs.shape[0], _ptr(cs), _ptr(off, ctypes.c_int64), _ptr(packed),
                     _ptr(po), _ptr(so), ctypes.byref(p), _ptr(m))
    if rc != 0:
        _err()
    return float(m[0])


def fmm_record(vel, grd, src, extraparams=None, precision="fp64"):
    """Single-source FMM: (tt grid, acceptance order, codes (N, ndim), n source points, n accepted)."""
    L = lib()
    g = grid_struct(grd)
    p = params_struct(extraparams, 1, precision)
    v = _vel(vel, grd)
    N = grd.npts
    tt = np.zeros(N)
    order = np.zeros(N, dtype=np.int64)
    codes = np.zeros(N * grd.ndim, dtype=np.int8)
    ns = np.zeros(1, dtype=np.int64)
    s = np.ascontiguousarray(np.asarray(src, dtype=np.float64).ravel())
    nacc = L.og_fmm_record(ctypes.byref(g), _ptr(v), _ptr(s), ctypes.byref(p), _ptr(tt),
                           _ptr(order, ctypes.c_int64), _ptr(codes, ctypes.c_int8), _ptr(ns, ctypes.c_int64))
    if nacc < 0:
        _err()
    return tt, order, codes.reshape(N, grd.ndim), int(ns[0]), int(nacc)


def fmm_bruteforce(vel, grd, src):
    L = lib()
    g = grid_struct(grd)
    v = _vel(vel, grd)
    N = grd.npts
    tt = np.zeros(N)
    order = np.zeros(N, dtype=np.int64)
    s = np.ascontiguousarray(np.asarray(src, dtype=np.float64).ravel())
    rc = L.og_fmm_bruteforce(ctypes.byref(g), _ptr(v), _ptr(s), _ptr(tt), _ptr(order, ctypes.c_int64))
    if rc != 0:
        _err()
    return tt, order


def stencil(grd, vel, tt, status, ijk):
    """Local 2nd-order update at node ijk (0-based) given flat tt/status arrays (Fortran order)."""
    L = lib()
    g = grid_struct(grd)
    v = _vel(vel, grd)
    t = np.ascontiguousarray(np.asarray(tt, dtype=np.float64).ravel(order="F"))
    st = np.ascontiguousarray(np.asarray(status, dtype=np.uint8).ravel(order="F"))
    p = np.zeros(3, dtype=np.int64)
    p[:len(ijk)] = ijk
    codes = np.zeros(3, dtype=np.int32)
    val = L.og_stencil(ctypes.byref(g), _ptr(v), _ptr(t), _ptr(st, ctypes.c_uint8), _ptr(p, ctypes.c_int64),
                       _ptr(codes, ctypes.c_int32))
    return float(val), codes[:grd.ndim].copy()


def smoothgradient(grd, l, grad):
    L = lib()
    g = grid_struct(grd)
    x = np.ascontiguousarray(np.asarray(grad, dtype=np.float64).ravel(order="F"))
    y = np.zeros_like(x)
    if L.og_smoothgradient(ctypes.byref(g), int(l), _ptr(x), _ptr(y)) != 0:
        _err()
    return y.reshape(grd.grsize, order="F")


def warning_count() -> int:
    return int(lib().og_warning_count())


def reset_warnings():
    lib().og_reset_warnings()


# Test hook: known mistakes that can be planted in the oracle's spherical forward, so that a test
# can show that its pins detect each of them (eik_oracle.h, og_plant_mistake).
MISTAKE_DEGREES = 1
MISTAKE_NO_RADIUS = 2
MISTAKE_NO_SINTHETA = 3
MISTAKE_INIT_DEGREES = 4


class planted_mistake:
    """Context manager: run the oracle with one planted mistake, then restore it."""

    def __init__(self, which: int):
        self.which = which

    def __enter__(self):
        self.prev = lib().og_plant_mistake(self.which)
        return self

    def __exit__(self, *exc):
        lib().og_plant_mistake(self.prev)
        return False
