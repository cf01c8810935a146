"""Request loop on one GPU: continuous batching of traveltime and gradient requests.

No counterpart in the reference (it runs one call at a time).  ``Server`` wraps ``eik_server``
(include/eik.h): ``nslots`` execution slots, each a context with its own stream, take queued
requests as soon as a slot is free and the request's device working set fits in the memory
budget; admitted requests overlap on the device and each batches all of its sources into shared
launches.  Models are ``EikModel`` handles kept resident by the caller.

    srv = Server(nslots=4)
    r1 = srv.submit_traveltime(model_a, coordsrc, coordrec)
    r2 = srv.submit_gradient(model_b, coordsrc, coordrec, pickobs, stdobs, "gradvel")
    picks = r1.result(); gradvel, misf = r2.result()
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .api import _WHICH, _pack, _params_struct

REQ_TRAVELTIME = 1
REQ_GRADIENT = 2


class Request:
    """A submitted request.  Holds its input and output arrays until ``result()``."""

    def __init__(self, server, rid, kind, model, nsrc, off, arrays, outputs, whichgrad):
        self._server = server
        self.id = rid
        self.kind = kind
        self.model = model
        self.nsrc = nsrc
        self._off = off
        self._arrays = arrays  # keeps the buffers the device side reads alive
        self._out = outputs
        self.whichgrad = whichgrad
        self.latency_ms = self.queue_ms = self.run_ms = None
        self.stats = None  # solver statistics of the call that ran the request (see result())
        self._done = False
        self._result = None

    def done(self) -> bool:
        """True once the request has finished (non-blocking, ``eik_poll``)."""
        if self._done:
            return True
        d = ctypes.c_int32(0)
        _lib.check(_lib.load().eik_poll(self._server.handle, self.id, ctypes.byref(d)), None, "eik_poll failed")
        return bool(d.value)

    def result(self):
        """Wait for completion and return what ``eiktraveltime`` / ``eikgradient`` return."""
        if self._done:
            return self._result
        t = np.zeros(3)
        st = np.zeros(14)
        err = ctypes.create_string_buffer(512)
        rc = _lib.load().eik_wait_stats(self._server.handle, self.id, _lib.ptr(t), _lib.ptr(st), 14, err, 512)
        self.latency_ms, self.queue_ms, self.run_ms = (float(x) for x in t)
        self.stats = {"rounds": int(st[5]), "settle_stage": int(st[11]), "verify_rounds": int(st[12]),
                      "kernel_ms": float(st[6]), "tile_activations": float(st[4]), "node_updates": float(st[8]),
                      "requests_in_call": int(st[13])}
        self._done = True
        _lib.check(rc, None, err.value.decode(errors="replace"))
        off = self._off
        if self.kind == REQ_TRAVELTIME:
            picks = self._out["picks"]
            self._result = [picks[off[s]:off[s + 1]].copy() for s in range(self.nsrc)]
        else:
            grd = self.model.grd
            gvel = self._out["gradvel"].reshape(grd.grsize, order="F")
            gsrc = self._out["gradsrc"][:self.nsrc].copy()
            misf = float(self._out["misf"][0])
            self._result = {1: (gvel, misf), 2: (gsrc, misf), 3: (gvel, gsrc, misf)}[self.whichgrad]
        self._arrays = None
        self._server._pending.pop(self.id, None)
        return self._result


def traveltime_batch(items, *, extraparams=None, ctx=None):
    """Several ``(model, coordsrc, coordrec)`` traveltime requests in shared launches
    (``eik_traveltime_batch``): the models share grid kind and dimension; returns each request's
    picks as ``eiktraveltime`` would."""
    from .api import default_context

    ctx = ctx or default_context()
    n = len(items)
    reqs = (_lib.EikRequest * max(n, 1))()
    keep, outs = [], []
    for i, (model, coordsrc, coordrec) in enumerate(items):
        grd = model.grd
        cs, off, packed = _pack(coordsrc, coordrec, grd.ndim)
        picks = np.zeros(max(int(off[-1]), 1))
        q = reqs[i]
        q.type = REQ_TRAVELTIME
        q.model = model.handle
        q.nsrc = cs.shape[0]
        q.coordsrc = _lib.ptr(cs)
        q.rec_off = _lib.ptr(off, ctypes.c_int64)
        q.coordrec = _lib.ptr(packed)
        q.params = _params_struct(extraparams)
        q.ttpicks = _lib.ptr(picks)
        keep.append((cs, off, packed))
        outs.append((picks, off, cs.shape[0]))
    rc = _lib.load().eik_traveltime_batch(ctx.handle, n, reqs)
    _lib.check(rc, ctx.handle)
    return [[p[o[s]:o[s + 1]].copy() for s in range(ns)] for p, o, ns in outs]


def request_bytes(model, nsrc: int, kind: str = "traveltime", *, extraparams=None) -> int:
    """The device working set (bytes) the server counts for a request of ``nsrc`` sources on
    ``model`` (``kind`` "traveltime" or "gradient"; ``eik_request_bytes``, no device work)."""
    req = _lib.EikRequest()
    req.type = REQ_TRAVELTIME if kind == "traveltime" else REQ_GRADIENT
    req.model = model.handle
    req.nsrc = int(nsrc)
    req.params = _params_struct(extraparams)
    out = ctypes.c_int64()
    _lib.check(_lib.load().eik_request_bytes(ctypes.byref(req), ctypes.byref(out)), None, "eik_request_bytes failed")
    return int(out.value)


class Server:
    """Continuous-batching request loop (``eik_server``) on one device."""

    def __init__(self, nslots: int = 4, mem_budget: int | None = None, device: int = 0):
        """``mem_budget``: device bytes the slots may hold; None (or 0): the device's free memory,
        read at every admission, less a reserve of max(2 GiB, 1/32 of the device)."""
        if mem_budget is None:
            mem_budget = 0
        h = ctypes.c_void_p()
        rc = _lib.load().eik_server_create(int(device), int(nslots), int(mem_budget), ctypes.byref(h))
        _lib.check(rc, None, "eik_server_create failed")
        self.handle = h
        self._pending = {}

    def _submit(self, req: _lib.EikRequest) -> int:
        rid = ctypes.c_int64()
        rc = _lib.load().eik_submit(self.handle, ctypes.byref(req), ctypes.byref(rid))
        _lib.check(rc, None, "eik_submit rejected the request")
        return int(rid.value)

    def submit_traveltime(self, model, coordsrc, coordrec, *, extraparams=None) -> Request:
        grd = model.grd
        cs, off, packed = _pack(coordsrc, coordrec, grd.ndim)
        nsrc = cs.shape[0]
        picks = np.zeros(max(int(off[-1]), 1))
        req = _lib.EikRequest()
        req.type = REQ_TRAVELTIME
        req.model = model.handle
        req.nsrc = nsrc
        req.coordsrc = _lib.ptr(cs)
        req.rec_off = _lib.ptr(off, ctypes.c_int64)
        req.coordrec = _lib.ptr(packed)
        req.params = _params_struct(extraparams)
        req.ttpicks = _lib.ptr(picks)
        rid = self._submit(req)
        r = Request(self, rid, REQ_TRAVELTIME, model, nsrc, off, (cs, off, packed), {"picks": picks}, 0)
        self._pending[rid] = r  # buffers stay alive until the request is waited for
        return r

    def submit_gradient(self, model, coordsrc, coordrec, pickobs, stdobs, whichgrad, *, extraparams=None) -> Request:
        grd = model.grd
        wg = _WHICH.get(str(whichgrad).lstrip(":"))
        assert wg is not None, "Symbol 'whichgrad' not in (:gradvel, :gradsrcloc, :gradvelandsrcloc) "
        cs, off, packed = _pack(coordsrc, coordrec, grd.ndim)
        nsrc = cs.shape[0]
        po = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).ravel() for x in pickobs]))
        so = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64).ravel() for x in stdobs]))
        assert po.size == off[-1] and so.size == off[-1], "one observation and one std per receiver"
        gv = np.zeros(grd.npts)
        gs = np.zeros((max(nsrc, 1), grd.ndim))
        misf = np.zeros(1)
        req = _lib.EikRequest()
        req.type = REQ_GRADIENT
        req.whichgrad = wg
        req.model = model.handle
        req.nsrc = nsrc
        req.coordsrc = _lib.ptr(cs)
        req.rec_off = _lib.ptr(off, ctypes.c_int64)
        req.coordrec = _lib.ptr(packed)
        req.pickobs = _lib.ptr(po)
        req.stdobs = _lib.ptr(so)
        req.params = _params_struct(extraparams)
        req.gradvel = _lib.ptr(gv)
        req.gradsrc = _lib.ptr(gs)
        req.misf = _lib.ptr(misf)
        rid = self._submit(req)
        r = Request(self, rid, REQ_GRADIENT, model, nsrc, off, (cs, off, packed, po, so),
                    {"gradvel": gv, "gradsrc": gs, "misf": misf}, wg)
        self._pending[rid] = r
        return r

    def stats(self) -> dict:
        out = np.zeros(10, dtype=np.int64)
        _lib.load().eik_server_stats(self.handle, _lib.ptr(out, ctypes.c_int64), 10)
        return {"queued": int(out[0]), "running": int(out[1]), "admitted_bytes": int(out[2]), "budget": int(out[3]),
                "shared_batches": int(out[4]), "requests_in_shared_batches": int(out[5]), "pool_bytes": int(out[6]),
                "rejected_over_budget": int(out[7]), "pool_releases": int(out[8]), "oom_retries": int(out[9])}

    def close(self):
        """Wait for every submitted request, then release the slots."""
        if getattr(self, "handle", None):
            _lib.load().eik_server_destroy(self.handle)  # returns once queued and running work is done
            self.handle = None
            self._pending.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
