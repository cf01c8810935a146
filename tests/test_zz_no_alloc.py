"""eik_traveltime_noalloc through the C ABI against eik_traveltime.

Two calls in a row, the second with a smaller input (fewer sources on a smaller model), must give
the picks and grids of eik_traveltime bit for bit (FP64: the same kernels on the same inputs) and
within the receiver tolerance in mixed precision; device free memory must not move across a
no-allocation call once the context, models and first call exist; a call whose buffers exceed the
scratch block is rejected with EIK_ENOMEM and a message naming the limit.
"""
import ctypes

import numpy as np
import pytest

from cases import random2d, random3d
from eikonalsolvers import _lib
from eikonalsolvers.api import _pack, _params_struct
from eikonalsolvers.grids import ExtraParams
from eikonalsolvers.synthetic import c2_3d

pytestmark = pytest.mark.gpu


def _call(name, model, src, recs, ep, returntt):
    L = _lib.load()
    cs, off, packed = _pack(src, recs, model.grd.ndim)
    p = _params_struct(ep)
    picks = np.zeros(max(int(off[-1]), 1))
    tt = np.zeros(cs.shape[0] * model.grd.npts) if returntt else None
    rc = getattr(L, name)(model.ctx.handle, model.handle, cs.shape[0], _lib.ptr(cs), _lib.ptr(off, ctypes.c_int64),
                          _lib.ptr(packed), ctypes.byref(p), _lib.ptr(picks),
                          _lib.ptr(tt) if returntt else None)
    return rc, picks[:int(off[-1])], tt


def _free_bytes(ctx):
    # through libeik's own HIP runtime (a process that imported torch may hold a second one)
    free, total = ctypes.c_int64(0), ctypes.c_int64(0)
    assert _lib.load().eik_mem_info(ctx.handle, ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("refine", [False, True])
def test_noalloc_equals_traveltime_on_two_calls(eik, precision, refine):
    ep = ExtraParams(refinearoundsrc=refine, precision=precision)
    big = random3d(seed=21, n=(40, 36, 30), nsrc=4, nrec=12)
    small = random2d(seed=22, n=(50, 40), nsrc=2, nrec=7)
    models = [eik.EikModel(c[1], c[0]) for c in (big, small)]
    try:
        for k, (case, model) in enumerate(zip((big, small), models)):
            grd, vel, src, recs = case
            print(f"call {k}: grid={grd.grsize} nsrc={len(src)} precision={precision} refine={refine}", flush=True)
            rc0, p0, t0 = _call("eik_traveltime", model, src, recs, ep, True)
            rc1, p1, t1 = _call("eik_traveltime_noalloc", model, src, recs, ep, True)
            assert rc0 == 0 and rc1 == 0, _lib.load().eik_last_error(model.ctx.handle)
            if precision == "fp64":
                np.testing.assert_array_equal(p1, p0)
                np.testing.assert_array_equal(t1, t0)
            else:
                np.testing.assert_allclose(p1, p0, rtol=1e-5, atol=0)
                assert np.linalg.norm(t1 - t0) <= 1e-5 * np.linalg.norm(t0)
    finally:
        for m in models:
            m.close()


def test_noalloc_allocates_no_device_memory(eik):
    grd, vel, src, recs = random3d(seed=23, n=(48, 40, 36), nsrc=6, nrec=10)
    model = eik.EikModel(vel, grd)
    try:
        ep = ExtraParams()
        rc, _, _ = _call("eik_traveltime_noalloc", model, src, recs, ep, True)  # warms the runtime
        assert rc == 0
        f0 = _free_bytes(model.ctx)
        for _ in range(3):
            rc, _, _ = _call("eik_traveltime_noalloc", model, src, recs, ep, True)
            assert rc == 0
        f1 = _free_bytes(model.ctx)
        print(f"device free bytes before {f0}, after three calls {f1}", flush=True)
        assert f1 == f0
    finally:
        model.close()


def test_noalloc_rejects_calls_beyond_the_scratch(eik):
    grd, vel, src, recs, _ = c2_3d(nsrc=4)
    model = eik.EikModel(vel, grd)
    try:
        rc, _, _ = _call("eik_traveltime_noalloc", model, src, recs, ExtraParams(), False)
        msg = _lib.load().eik_last_error(model.ctx.handle).decode()
        print(f"rc={rc}: {msg}", flush=True)
        assert rc == _lib.EIK_ENOMEM
        assert "128 MiB" in msg
        # one source of the same model fits
        rc, p, _ = _call("eik_traveltime_noalloc", model, src[:1], recs[:1], ExtraParams(), False)
        assert rc == 0
        rc0, p0, _ = _call("eik_traveltime", model, src[:1], recs[:1], ExtraParams(), False)
        np.testing.assert_array_equal(p, p0)
    finally:
        model.close()
