"""Serving loop (eik_server): every request's result is the same alone and in the mix.

A mix of 2D/3D traveltime and gradient requests (the serving workload's distribution: shrunk, and
at full size with its largest request, 160^3 with 64 sources) is submitted at once to a 4-slot server, so requests are admitted while others run and their work
overlaps on the device; each result is compared with the same request run alone through the
plain API.  FP64: identical picks, gradients to 1e-12 relative.  Mixed precision: receivers 1e-5
relative, gradients 1e-3 relative L2 (the stated tolerances).
"""
import time

import numpy as np
import pytest

from eikonalsolvers.grids import ExtraParams, Grid3DCart
from eikonalsolvers.synthetic import c5_request, smooth_velocity, uniform

pytestmark = pytest.mark.gpu

NREQ = 10


def _alone(req, model, ep):
    if req["kind"] == "traveltime":
        return model.traveltime(req["src"], req["recs"], extraparams=ep)
    return model.gradient(req["src"], req["recs"], req["pickobs"], req["stdobs"], "gradvelandsrcloc",
                          extraparams=ep)


def _largest_gradient_request():
    """The largest request the serving mix draws: 3D 160^3 (h = 1), 64 sources, 128 top-face
    receivers, a gradient; drawn like the mix's requests (seeds 50999 / 51000 / 51001)."""
    n, nsrc, nrec = 160, 64, 128
    grd = Grid3DCart(hgrid=1.0, cooinit=(0.0, 0.0, 0.0), grsize=(n, n, n))
    L = np.full(3, (n - 1) * 1.0)
    src = (0.05 + 0.9 * uniform(51000, nsrc * 3).reshape(nsrc, 3)) * L
    rec = uniform(51001, nrec * 3).reshape(nrec, 3) * L
    rec[:, -1] = 0.0
    return {"id": -1, "kind": "gradient", "grd": grd, "vel": smooth_velocity(grd.grsize, 50999), "src": src,
            "recs": [rec.copy() for _ in range(nsrc)],
            "pickobs": [np.sqrt(((rec - s) ** 2).sum(1)) / 3.5 for s in src],
            "stdobs": [np.full(nrec, 0.01) for _ in range(nsrc)]}


def _mix(reqs, models, ep, eik, nslots=4):
    with eik.Server(nslots=nslots) as srv:
        t_sub, handles = [], []
        for r, m in zip(reqs, models):
            t_sub.append(time.perf_counter())
            if r["kind"] == "traveltime":
                handles.append(srv.submit_traveltime(m, r["src"], r["recs"], extraparams=ep))
            else:
                handles.append(srv.submit_gradient(m, r["src"], r["recs"], r["pickobs"], r["stdobs"],
                                                   "gradvelandsrcloc", extraparams=ep))
        out = [h.result() for h in handles]
        st = srv.stats()
    spans = [(t + h.queue_ms / 1e3, t + (h.queue_ms + h.run_ms) / 1e3) for t, h in zip(t_sub, handles)]
    n = len(reqs)
    overlaps = sum(1 for i in range(n) for j in range(i + 1, n) if spans[i][0] < spans[j][1] and spans[j][0] < spans[i][1])
    return out, handles, overlaps, st


def _same(reqs, alone, mixed, precision):
    for r, a, b in zip(reqs, alone, mixed):
        if r["kind"] == "traveltime":
            for pa, pb in zip(a, b):
                if precision == "fp64":
                    np.testing.assert_array_equal(pa, pb)
                else:
                    np.testing.assert_allclose(pb, pa, rtol=1e-5, atol=0)
        else:
            gva, gsa, ma = a
            gvb, gsb, mb = b
            tol = 1e-12 if precision == "fp64" else 1e-3
            assert np.linalg.norm(gvb - gva) <= tol * np.linalg.norm(gva)
            assert np.linalg.norm(gsb - gsa) <= tol * max(np.linalg.norm(gsa), 1e-300)
            assert mb == pytest.approx(ma, rel=1e-12 if precision == "fp64" else 1e-5)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_full_size_mix_same_alone_and_in_the_mix(eik, precision):
    """Full-size serving requests, including the largest the mix can draw (160^3, 64 sources, gradient),
    run concurrently with six full-size mix requests; each equals its run alone."""
    ep = ExtraParams(precision=precision)
    reqs = [_largest_gradient_request()] + [c5_request(i) for i in range(6)]
    print(f"precision={precision}: " + ", ".join(f"{r['kind'][0]}{r['grd'].grsize}x{len(r['src'])}" for r in reqs),
          flush=True)
    models = [eik.EikModel(r["vel"], r["grd"]) for r in reqs]
    alone = [_alone(r, m, ep) for r, m in zip(reqs, models)]
    mixed, handles, overlaps, st = _mix(reqs, models, ep, eik)
    print(f"  latencies ms: {[round(h.latency_ms, 1) for h in handles]}; overlapping pairs: {overlaps}; {st}",
          flush=True)
    assert overlaps > 0, "requests never ran concurrently"
    _same(reqs, alone, mixed, precision)
    for m in models:
        m.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_results_same_alone_and_in_the_mix(eik, precision):
    ep = ExtraParams(precision=precision)
    reqs = [c5_request(i, small=True) for i in range(NREQ)]
    models = [eik.EikModel(r["vel"], r["grd"]) for r in reqs]
    alone = [_alone(r, m, ep) for r, m in zip(reqs, models)]
    print(f"precision={precision}: " + ", ".join(f"{r['kind'][0]}{r['grd'].grsize}x{len(r['src'])}" for r in reqs),
          flush=True)
    with eik.Server(nslots=4) as srv:
        t_sub, handles = [], []
        for r, m in zip(reqs, models):
            t_sub.append(time.perf_counter())
            if r["kind"] == "traveltime":
                handles.append(srv.submit_traveltime(m, r["src"], r["recs"], extraparams=ep))
            else:
                handles.append(srv.submit_gradient(m, r["src"], r["recs"], r["pickobs"], r["stdobs"],
                                                   "gradvelandsrcloc", extraparams=ep))
        mixed = [h.result() for h in handles]
        st = srv.stats()
    # run intervals on the server clock, relative to each submission
    spans = [(t + h.queue_ms / 1e3, t + (h.queue_ms + h.run_ms) / 1e3) for t, h in zip(t_sub, handles)]
    overlaps = sum(1 for i in range(NREQ) for j in range(i + 1, NREQ)
                   if spans[i][0] < spans[j][1] and spans[j][0] < spans[i][1])
    print(f"  latencies ms: {[round(h.latency_ms, 1) for h in handles]}; overlapping pairs: {overlaps}; "
          f"{st}")
    assert overlaps > 0, "requests never ran concurrently"
    for r, a, b in zip(reqs, alone, mixed):
        if r["kind"] == "traveltime":
            for pa, pb in zip(a, b):
                if precision == "fp64":
                    np.testing.assert_array_equal(pa, pb)
                else:
                    np.testing.assert_allclose(pb, pa, rtol=1e-5, atol=0)
        else:
            gva, gsa, ma = a
            gvb, gsb, mb = b
            tol = 1e-12 if precision == "fp64" else 1e-3
            assert np.linalg.norm(gvb - gva) <= tol * np.linalg.norm(gva)
            assert np.linalg.norm(gsb - gsa) <= tol * max(np.linalg.norm(gsa), 1e-300)
            assert mb == pytest.approx(ma, rel=1e-12 if precision == "fp64" else 1e-5)
    for m in models:
        m.close()


def test_request_errors_are_reported_per_request(eik):
    reqs = [c5_request(i, small=True) for i in range(2)]
    models = [eik.EikModel(r["vel"], r["grd"]) for r in reqs]
    with eik.Server(nslots=2) as srv:
        bad_src = np.array(reqs[0]["src"], copy=True)
        bad_src[0, 0] = -1.0  # outside the grid: the reference's checksrcrecposition @assert
        h_bad = srv.submit_traveltime(models[0], bad_src, reqs[0]["recs"])
        h_ok = srv.submit_traveltime(models[1], reqs[1]["src"], reqs[1]["recs"])
        with pytest.raises(AssertionError):
            h_bad.result()
        assert len(h_ok.result()) == len(reqs[1]["src"])
    for m in models:
        m.close()


def _tt_requests(nd, count):
    out = []
    rid = 0
    while len(out) < count:
        r = c5_request(rid, small=True)
        rid += 1
        if r["kind"] == "traveltime" and r["grd"].ndim == nd:
            out.append(r)
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("nd", [2, 3])
def test_traveltime_batch_equals_separate_calls(eik, precision, nd):
    """Requests on different models (sizes and velocities differ) in one shared-launch call give
    the picks of separate calls: FP64 identical, mixed within the receiver tolerance."""
    ep = ExtraParams(precision=precision)
    reqs = _tt_requests(nd, 5)
    models = [eik.EikModel(r["vel"], r["grd"]) for r in reqs]
    alone = [m.traveltime(r["src"], r["recs"], extraparams=ep) for r, m in zip(reqs, models)]
    batched = eik.traveltime_batch([(m, r["src"], r["recs"]) for r, m in zip(reqs, models)], extraparams=ep)
    print(f"{nd}D {precision}: grids {[r['grd'].grsize[0] for r in reqs]}, sources {[len(r['src']) for r in reqs]}")
    for a, b in zip(alone, batched):
        for pa, pb in zip(a, b):
            if precision == "fp64":
                np.testing.assert_array_equal(pb, pa)
            else:
                np.testing.assert_allclose(pb, pa, rtol=1e-5, atol=0)
    for m in models:
        m.close()


def test_server_batches_queued_traveltime_requests(eik):
    """A one-slot server admits the first request and takes the compatible queued ones along into
    one shared-launch call; every result equals the request run alone."""
    ep = ExtraParams()
    reqs = _tt_requests(2, 8)
    models = [eik.EikModel(r["vel"], r["grd"]) for r in reqs]
    alone = [m.traveltime(r["src"], r["recs"], extraparams=ep) for r, m in zip(reqs, models)]
    with eik.Server(nslots=1) as srv:
        hs = [srv.submit_traveltime(m, r["src"], r["recs"], extraparams=ep) for r, m in zip(reqs, models)]
        t0 = time.perf_counter()
        while not all(h.done() for h in hs):  # eik_poll, without blocking
            assert time.perf_counter() - t0 < 60.0
            time.sleep(0.002)
        res = [h.result() for h in hs]
        st = srv.stats()
    print(f"server stats: {st}")
    assert st["shared_batches"] >= 1 and st["requests_in_shared_batches"] >= 2
    for a, b in zip(alone, res):
        for pa, pb in zip(a, b):
            np.testing.assert_array_equal(pb, pa)
    for m in models:
        m.close()


def _same_gradient_requests(eik, count):
    """`count` copies of one gradient request of the shrunk serving mix, on one model (equal
    working sets), and that request's result alone."""
    r = next(q for q in (c5_request(i, small=True) for i in range(40)) if q["kind"] == "gradient")
    m = eik.EikModel(r["vel"], r["grd"])
    return r, m, [r] * count


def test_budget_binds_requests_wait_for_memory(eik):
    """A budget of 1.5 working sets on a two-slot server: two requests never run at once (the
    second waits, or the idle slot's pools are released for it), the pools the slots hold stay
    within the budget, and every result equals the request alone (FP64: gradients to 1e-12
    relative, the order of the atomic reductions aside)."""
    from eikonalsolvers.serving import request_bytes

    ep = ExtraParams()
    r, m, reqs = _same_gradient_requests(eik, 4)
    b = request_bytes(m, len(r["src"]), "gradient", extraparams=ep)
    budget = int(1.5 * b)
    alone = _alone(r, m, ep)
    with eik.Server(nslots=2, mem_budget=budget) as srv:
        t_sub, hs = [], []
        for q in reqs:
            t_sub.append(time.perf_counter())
            hs.append(srv.submit_gradient(m, q["src"], q["recs"], q["pickobs"], q["stdobs"], "gradvelandsrcloc",
                                          extraparams=ep))
        out = [h.result() for h in hs]
        st = srv.stats()
    spans = sorted((t + h.queue_ms / 1e3, t + (h.queue_ms + h.run_ms) / 1e3) for t, h in zip(t_sub, hs))
    print(f"working set {b} B, budget {budget} B, stats {st}, spans {spans}")
    assert st["budget"] == budget
    assert all(spans[k + 1][0] >= spans[k][1] - 1e-3 for k in range(len(spans) - 1)), "requests overlapped"
    assert 0 < st["pool_bytes"] <= budget
    _same(reqs, [alone] * len(reqs), out, "fp64")  # gradients to 1e-12 (atomic reductions), misfit 1e-12
    m.close()


def test_request_larger_than_budget_fails_alone(eik):
    """A request whose working set exceeds the whole budget fails with EIK_ENOMEM (MemoryError)
    without running; a smaller request submitted with it runs."""
    from eikonalsolvers.serving import request_bytes

    ep = ExtraParams()
    r, m, _ = _same_gradient_requests(eik, 1)
    small = next(q for q in (c5_request(i, small=True) for i in range(40))
                 if q["kind"] == "traveltime" and q["grd"].ndim == 2)
    ms = eik.EikModel(small["vel"], small["grd"])
    b_big = request_bytes(m, len(r["src"]), "gradient", extraparams=ep)
    b_small = request_bytes(ms, len(small["src"]), "traveltime", extraparams=ep)
    assert b_small < b_big // 2
    with eik.Server(nslots=2, mem_budget=b_big // 2) as srv:
        h_big = srv.submit_gradient(m, r["src"], r["recs"], r["pickobs"], r["stdobs"], "gradvel", extraparams=ep)
        h_small = srv.submit_traveltime(ms, small["src"], small["recs"], extraparams=ep)
        with pytest.raises(MemoryError, match="budget"):
            h_big.result()
        got = h_small.result()
        st = srv.stats()
    assert st["rejected_over_budget"] == 1
    for pa, pb in zip(ms.traveltime(small["src"], small["recs"], extraparams=ep), got):
        np.testing.assert_array_equal(pb, pa)
    m.close()
    ms.close()


def test_default_budget_follows_free_device_memory(eik):
    """With no budget given the server admits against the device's free memory less the reserve
    (max(2 GiB, 1/32 of the device)), read at every admission."""
    import ctypes

    from eikonalsolvers import _lib

    free, total = ctypes.c_int64(), ctypes.c_int64()
    assert _lib.load().eik_mem_info(eik.default_context().handle, ctypes.byref(free), ctypes.byref(total)) == 0
    reserve = max(2 << 30, total.value // 32)
    r, m, reqs = _same_gradient_requests(eik, 2)
    with eik.Server(nslots=2) as srv:
        b0 = srv.stats()["budget"]
        hs = [srv.submit_gradient(m, q["src"], q["recs"], q["pickobs"], q["stdobs"], "gradvel") for q in reqs]
        for h in hs:
            h.result()
            assert h.stats["rounds"] > 0 and h.stats["requests_in_call"] == 1
        st = srv.stats()
    print(f"free {free.value}, total {total.value}, budget at creation {b0}, after {st}")
    assert 0 < b0 <= free.value - reserve + (64 << 20)
    assert 0 < st["budget"] <= total.value - reserve
    m.close()


def _tt_request_2d(n, nsrc, seed):
    """A 2D n x n traveltime request with nsrc sources and 16 top receivers each."""
    from eikonalsolvers.grids import Grid2DCart
    from eikonalsolvers.synthetic import uniform

    grd = Grid2DCart(hgrid=1.0, cooinit=(0.0, 0.0), grsize=(n, n))
    L = (n - 1) * 1.0
    src = (0.1 + 0.8 * uniform(seed, 2 * nsrc).reshape(nsrc, 2)) * L
    rec = np.column_stack([np.linspace(0.0, L, 16), np.zeros(16)])
    return grd, src, [rec.copy() for _ in range(nsrc)]


def test_shared_launch_takes_only_requests_within_twice_the_working_set(eik):
    """With the only slot busy, three traveltime requests on one model queue: the first is
    admitted with the one whose working set is within 2x of its own (one shared call), not the
    one 4x larger, which runs in a call of its own; results equal separate calls."""
    from eikonalsolvers.serving import request_bytes

    ep = ExtraParams()
    grd, src, recs = _tt_request_2d(300, 24, 61001)
    model = eik.EikModel(smooth_velocity(grd.grsize, 61000), grd)
    sizes = {"a": 4, "b": 7, "c": 24}
    b = {k: request_bytes(model, v, "traveltime", extraparams=ep) for k, v in sizes.items()}
    assert 1.0 < b["b"] / b["a"] <= 2.0 and b["c"] / b["a"] > 2.0 and b["c"] / b["b"] > 2.0
    r, m_blk, _ = _same_gradient_requests(eik, 1)
    with eik.Server(nslots=1) as srv:
        blk = srv.submit_gradient(m_blk, r["src"], r["recs"], r["pickobs"], r["stdobs"], "gradvel")
        hs = {k: srv.submit_traveltime(model, src[:v], recs[:v], extraparams=ep) for k, v in sizes.items()}
        out = {k: h.result() for k, h in hs.items()}
        blk.result()
    calls = {k: h.stats["requests_in_call"] for k, h in hs.items()}
    print(f"working sets {b}, requests per call {calls}")
    assert calls == {"a": 2, "b": 2, "c": 1}
    for k, v in sizes.items():
        for pa, pb in zip(model.traveltime(src[:v], recs[:v], extraparams=ep), out[k]):
            np.testing.assert_array_equal(pb, pa)
    model.close()
    m_blk.close()


@pytest.mark.parametrize("bypass_ms", [None, 0])
def test_blocked_request_is_passed_only_within_the_bypass_time(eik, monkeypatch, bypass_ms):
    """Two slots, a budget of 1.5 gradient working sets: while one gradient runs, a second waits
    for memory and a small request that fits is admitted ahead of it (default: a blocked request
    may be passed for 2 s).  With the bypass time set to 0 (EIK_SERVER_BYPASS_MS, read when the
    server is created) the small request waits until the blocked one has started."""
    from eikonalsolvers.serving import request_bytes

    if bypass_ms is not None:
        monkeypatch.setenv("EIK_SERVER_BYPASS_MS", str(bypass_ms))
    ep = ExtraParams()
    r, m, _ = _same_gradient_requests(eik, 1)
    w = request_bytes(m, len(r["src"]), "gradient", extraparams=ep)
    grd, src, recs = _tt_request_2d(120, 2, 62001)
    ms = eik.EikModel(smooth_velocity(grd.grsize, 62000), grd)
    assert request_bytes(ms, 2, "traveltime", extraparams=ep) <= 0.4 * w
    with eik.Server(nslots=2, mem_budget=int(1.5 * w)) as srv:
        t_sub = []
        hs = []
        for kind in ("g1", "g2", "small"):
            t_sub.append(time.perf_counter())
            if kind == "small":
                hs.append(srv.submit_traveltime(ms, src, recs, extraparams=ep))
            else:
                hs.append(srv.submit_gradient(m, r["src"], r["recs"], r["pickobs"], r["stdobs"], "gradvel",
                                              extraparams=ep))
        for h in hs:
            h.result()
    start = [t + h.queue_ms / 1e3 for t, h in zip(t_sub, hs)]
    print(f"bypass {bypass_ms}: starts after the first submission (ms) "
          f"{[round(1e3 * (x - t_sub[0]), 1) for x in start]}")
    if bypass_ms is None:
        assert start[2] < start[1], "the small request should pass the blocked gradient"
    else:
        assert start[2] >= start[1] - 1e-4, "the small request passed the blocked gradient"
    m.close()
    ms.close()
