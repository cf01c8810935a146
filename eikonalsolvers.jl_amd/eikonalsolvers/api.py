"""Public API of EikonalSolvers.jl on the MI355X engine.

Same names, argument meaning and exceptions as the reference:

* ``eiktraveltime``  -- src/EikSolv/eikmainfwd.jl:33-143
* ``eikgradient``    -- src/EikSolv/eikmaingrad.jl:31-164
* ``eikttimemisfit`` -- src/EikSolv/eikmaingrad.jl:1302-1328

plus ``EikModel``, a velocity model kept resident in HBM for repeated calls (the reference
re-sends ``vel`` on every call).  Arrays use the reference's indexing: ``vel[i, j(, k)]`` of shape
``grd.grsize``; ``coordsrc`` is ``nsrc x ndim``; ``coordrec``, ``pickobs`` and ``stdobs`` are
per-source lists.
"""
from __future__ import annotations

import ctypes
import threading

import numpy as np

from . import _lib
from .grids import PRECISIONS, ExtraParams

_WHICH = {"gradvel": 1, "gradsrcloc": 2, "gradvelandsrcloc": 3}


class Context:
    """One HIP device + stream + device scratch pools (eik_ctx)."""

    def __init__(self, device: int = 0):
        L = _lib.load()
        h = ctypes.c_void_p()
        rc = L.eik_ctx_create(int(device), ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(f"eik_ctx_create(device={device}) failed with {rc}: no usable GPU?")
        self.handle = h
        self.device = device

    def set_streams(self, nstreams: int):
        """Streams large calls may split their sources over (2, the default, or 1)."""
        _lib.check(_lib.load().eik_ctx_set_streams(self.handle, int(nstreams)), self.handle)

    def close(self):
        if self.handle:
            _lib.load().eik_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def last_stats(self):
        out = np.zeros(19)
        _lib.load().eik_last_stats(self.handle, _lib.ptr(out), 19)
        return {"total_ms": out[0], "fine_ms": out[1], "coarse_ms": out[2], "adjoint_ms": out[3],
                "tile_activations": out[4], "rounds": out[5], "_kernel_ms": out[6], "tile_sweeps": out[7],
                "node_updates": out[8], "changed_updates": out[9], "key_only_updates": out[10],
                "settle_stage": int(out[11]), "verify_rounds": int(out[12]), "launches": out[13],
                "fine_kernel_ms": out[14], "fine_launches": out[15], "fine_tile_activations": out[16],
                "kernel_busy_ms": out[17], "all_kernel_busy_ms": out[18]}


_default_ctx = None
_ctx_lock = threading.Lock()


def default_context() -> Context:
    global _default_ctx
    with _ctx_lock:
        if _default_ctx is None:
            _default_ctx = Context(0)
        return _default_ctx


def _grid_struct(grd) -> _lib.EikGrid:
    g = _lib.EikGrid()
    g.kind = grd.kind
    g.ndim = grd.ndim
    for d in range(3):
        g.n[d] = grd.grsize[d] if d < grd.ndim else 1
        g.d[d] = grd.spacing3[d]
        g.o[d] = grd.cooinit[d] if d < grd.ndim else 0.0
    return g


def _params_struct(extraparams) -> _lib.EikParams:
    ep = extraparams if extraparams is not None else ExtraParams()
    p = _lib.EikParams()
    p.refinearoundsrc = int(bool(ep.refinearoundsrc))
    p.downscalefactor = int(ep.grdrefpars.downscalefactor)
    p.noderadius = int(ep.grdrefpars.noderadius)
    p.radiussmoothgradsrc = int(ep.radiussmoothgradsrc)
    p.smoothgradkern = int(ep.smoothgradkern)
    p.precision = PRECISIONS[ep.precision]
    p.fast = int(bool(ep.fast))
    return p


def _vel_array(vel, grd) -> np.ndarray:
    v = np.asarray(vel, dtype=np.float64)
    assert 2 <= v.ndim <= 3, "2<=ndims(vel)<=3"
    assert v.shape == tuple(grd.grsize), f"size(vel) {v.shape} != grd.grsize {grd.grsize}"
    return np.ascontiguousarray(v.ravel(order="F"))


def _pack(coordsrc, coordrec, ndim):
    cs = np.asarray(coordsrc, dtype=np.float64)
    if cs.ndim == 1:
        cs = cs.reshape(1, -1)
    assert cs.ndim == 2 and cs.shape[1] == ndim, "coordsrc must be nsrc x ndim"
    cs = np.ascontiguousarray(cs)
    nsrc = cs.shape[0]
    assert nsrc == len(coordrec), "size(coordsrc,1)==length(coordrec)"
    recs = []
    for r in coordrec:
        a = np.asarray(r, dtype=np.float64)
        if a.ndim == 1:
            a = a.reshape(1, -1)
        assert a.ndim == 2 and a.shape[1] == ndim, "coordrec[s] must be nrec x ndim"
        recs.append(a)
    off = np.zeros(nsrc + 1, dtype=np.int64)
    off[1:] = np.cumsum([a.shape[0] for a in recs]) if nsrc else []
    packed = np.ascontiguousarray(np.concatenate(recs, axis=0)) if nsrc else np.zeros((0, ndim))
    return cs, off, packed


class EikModel:
    """A velocity model resident in HBM (eik_model)."""

    def __init__(self, vel, grd, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        self.grd = grd
        v = _vel_array(vel, grd)
        g = _grid_struct(grd)
        h = ctypes.c_void_p()
        rc = _lib.load().eik_model_create(self.ctx.handle, ctypes.byref(g), _lib.ptr(v), ctypes.byref(h))
        _lib.check(rc, self.ctx.handle)
        self.handle = h

    def update_vel(self, vel):
        v = _vel_array(vel, self.grd)
        _lib.check(_lib.load().eik_model_update_vel(self.handle, _lib.ptr(v)), self.ctx.handle)

    def close(self):
        if getattr(self, "handle", None):
            _lib.load().eik_model_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ calls
    def traveltime(self, coordsrc, coordrec, *, returntt=False, extraparams=None):
        grd = self.grd
        cs, off, packed = _pack(coordsrc, coordrec, grd.ndim)
        nsrc = cs.shape[0]
        p = _params_struct(extraparams)
        picks = np.zeros(max(int(off[-1]), 1))
        tt = np.zeros(nsrc * grd.npts) if returntt else None
        rc = _lib.load().eik_traveltime(self.ctx.handle, self.handle, nsrc, _lib.ptr(cs),
                                        _lib.ptr(off, ctypes.c_int64), _lib.ptr(packed), ctypes.byref(p),
                                        _lib.ptr(picks), _lib.ptr(tt) if tt is not None else None)
        _lib.check(rc, self.ctx.handle)
        out = [picks[off[s]:off[s + 1]].copy() for s in range(nsrc)]
        if returntt:
            N = grd.npts
            return out, [tt[s * N:(s + 1) * N].reshape(grd.grsize, order="F").copy() for s in range(nsrc)]
        return out

    def gradient(self, coordsrc, coordrec, pickobs, stdobs, whichgrad, *, extraparams=None):
        grd = self.grd
        wg = _WHICH.get(str(whichgrad).lstrip(":"))
        assert wg is not None, "Symbol 'whichgrad' not in (:gradvel, :gradsrcloc, :gradvelandsrcloc) "
        cs, off, packed = _pack(coordsrc, coordrec, grd.ndim)
        nsrc = cs.shape[0]
        assert len(coordrec) == len(pickobs), "length(coordrec)==length(pickobs)"
        for r in range(len(coordrec)):
            assert np.asarray(coordrec[r]).reshape(-1, grd.ndim).shape[0] == np.asarray(pickobs[r]).size, \
                "size(coordrec[r],1)==size(pickobs[r],1)"
            assert np.asarray(stdobs[r]).shape == np.asarray(pickobs[r]).shape, "size(stdobs[r])==size(pickobs[r])"
        po = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).ravel() for x in pickobs])
                                  if nsrc else np.zeros(1))
        so = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).ravel() for x in stdobs])
                                  if nsrc else np.zeros(1))
        p = _params_struct(extraparams)
        gv = np.zeros(grd.npts)
        gs = np.zeros((max(nsrc, 1), grd.ndim))
        misf = np.zeros(1)
        rc = _lib.load().eik_gradient(self.ctx.handle, self.handle, nsrc, _lib.ptr(cs), _lib.ptr(off, ctypes.c_int64),
                                      _lib.ptr(packed), _lib.ptr(po), _lib.ptr(so), wg, ctypes.byref(p),
                                      _lib.ptr(gv), _lib.ptr(gs), _lib.ptr(misf))
        _lib.check(rc, self.ctx.handle)
        gvel = gv.reshape(grd.grsize, order="F")
        gsrc = gs[:nsrc].copy()
        if wg == 1:
            return gvel, float(misf[0])
        if wg == 2:
            return gsrc, float(misf[0])
        return gvel, gsrc, float(misf[0])

    def misfit(self, coordsrc, coordrec, pickobs, stdobs, *, extraparams=None):
        grd = self.grd
        cs, off, packed = _pack(coordsrc, coordrec, grd.ndim)
        assert len(coordrec) == len(pickobs) == len(stdobs), "length(coordrec)==length(pickobs)"
        po = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).ravel() for x in pickobs]
                                                 or [np.zeros(0)]))
        so = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).ravel() for x in stdobs]
                                                 or [np.zeros(0)]))
        assert po.size == so.size == int(off[-1]), "size(pickobs)==size(stdobs)==number of receivers"
        p = _params_struct(extraparams)
        m = np.zeros(1)
        rc = _lib.load().eik_misfit(self.ctx.handle, self.handle, cs.shape[0], _lib.ptr(cs),
                                    _lib.ptr(off, ctypes.c_int64), _lib.ptr(packed), _lib.ptr(po), _lib.ptr(so),
                                    ctypes.byref(p), _lib.ptr(m))
        _lib.check(rc, self.ctx.handle)
        return float(m[0])


def eiktraveltime(vel, grd, coordsrc, coordrec, *, returntt=False, extraparams=None):
    """Traveltimes at the receivers (and optionally on the whole grid) for every source.

    Mirrors ``eiktraveltime`` (eikmainfwd.jl:33-143): returns ``ttpicks`` (list of arrays) or
    ``(ttpicks, ttime)`` when ``returntt``.
    """
    m = EikModel(vel, grd)
    try:
        return m.traveltime(coordsrc, coordrec, returntt=returntt, extraparams=extraparams)
    finally:
        m.close()


def eikgradient(vel, grd, coordsrc, coordrec, pickobs, stdobs, whichgrad, *, extraparams=None):
    """Discrete-adjoint gradient of the misfit (eikmaingrad.jl:31-164).

    Returns ``(gradvel, misf)``, ``(gradsrcloc, misf)`` or ``(gradvel, gradsrcloc, misf)``.
    """
    m = EikModel(vel, grd)
    try:
        return m.gradient(coordsrc, coordrec, pickobs, stdobs, whichgrad, extraparams=extraparams)
    finally:
        m.close()


def eikttimemisfit(velmod, grd, coordsrc, coordrec, ttpicksobs, stdobs, *, extraparams=None):
    """0.5 * sum((t - t_obs)^2 / sigma^2) (eikmaingrad.jl:1302-1328)."""
    m = EikModel(velmod, grd)
    try:
        return m.misfit(coordsrc, coordrec, ttpicksobs, stdobs, extraparams=extraparams)
    finally:
        m.close()
