"""xt_trace_device_static (the trace entry point that allocates nothing): through
the C ABI, its hits equal xt_trace_device's on the same device-resident rays,
bit for bit, on two calls in a row (the second with fewer rays, over a larger
first call's leftovers in the shared scratch), and it refuses more rays than
its scratch holds."""
import ctypes

import numpy as np
import pytest

import xtracer_amd as xa
from tests.helpers import HipBuffers, mixed_rays, scene_path

MAX_RAYS = 1 << 20  # XT_TRACE_STATIC_MAX_RAYS


def _rays(o, d):
    r0 = np.zeros((len(o), 4), np.float32)
    r1 = np.zeros((len(o), 4), np.float32)
    r0[:, :3] = o
    r0[:, 3] = np.uint32(0xFFFFFFFF).view(np.float32)  # no origin primitive
    r1[:, :3] = d
    return r0, r1


@pytest.mark.gpu
@pytest.mark.parametrize("accel", [xa.ACCEL_KDTREE, xa.ACCEL_OCTREE])
def test_static_trace_equals_trace_device(accel):
    L = xa.lib()
    hip = HipBuffers()
    dev = xa.Device(0)
    try:
        s = xa.Scene(scene_path("icosphere", subdiv=5))
        sid = dev.upload(s)
        for n, seed in ((300_000, 1), (70_001, 2)):
            o, d = mixed_rays(None, n, seed=seed)
            r0, r1 = _rays(o, d)
            d0, d1 = hip.put(r0), hip.put(r1)
            outs = []
            for fn, extra in ((L.xt_trace_device, (None,)), (L.xt_trace_device_static, ())):
                hit0 = np.full((n, 4), np.nan, np.float32)
                flag0 = np.full(n, 7, np.uint32)
                dh, df = hip.put(hit0), hip.put(flag0)
                rc = fn(dev._d, sid, accel, d0, d1, dh, df, ctypes.c_uint64(n), *extra)
                assert rc == 0, L.xt_last_error(dev._d)
                outs.append((hip.get(dh, hit0), hip.get(df, flag0)))
            (h_ref, f_ref), (h_st, f_st) = outs
            assert np.array_equal(h_ref.view(np.uint32), h_st.view(np.uint32)), f"{n} rays: hits differ"
            assert np.array_equal(f_ref, f_st)
            assert np.isfinite(h_st[:, 0]).any() and (h_st[:, 1].view(np.uint32) != 0xFFFFFFFF).any()
        # more rays than the scratch holds: refused before any work
        rc = L.xt_trace_device_static(dev._d, sid, accel, d0, d1, d0, None, ctypes.c_uint64(MAX_RAYS + 1))
        assert rc == 1  # XT_E_INVALID
        dev.release(sid)
    finally:
        dev.close()
        hip.free()
