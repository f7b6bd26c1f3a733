"""kpt_trace_closest_counted_noalloc through the C-ABI: the counted
closest-hit trace whose event counters live in the library's fixed device
array instead of a buffer allocated per call.  Its hits and counters equal
kpt_trace_closest_counted's on two calls in a row, the second with fewer
rays (the shared array is cleared by every call), on an empty call, and
after a refused call while a TLAS rebuild is pending."""
import ctypes as C

import numpy as np
import pytest

from kittlespt_amd import abi, scenes
from kittlespt_amd.context import Context, DeviceBuffer

from common import random_rays

pytestmark = pytest.mark.gpu


def _noalloc(ctx, rays):
    L = abi.lib()
    f = L.kpt_trace_closest_counted_noalloc
    f.restype = C.c_int
    f.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(abi.KptTraceCounters)]
    rays = np.ascontiguousarray(rays, abi.RAY_DTYPE)
    n = len(rays)
    dr = DeviceBuffer.from_array(rays)
    dh = DeviceBuffer(n * abi.HIT_DTYPE.itemsize)
    cnt = abi.KptTraceCounters()
    try:
        abi.check(ctx.h, f(ctx.h, dr.ptr, dh.ptr, n, C.byref(cnt)))
        hits = dh.download(np.zeros(n, abi.HIT_DTYPE))
    finally:
        dr.free()
        dh.free()
    return hits, cnt.as_dict()


def _same_hits(a, b):
    return (np.array_equal(a["prim_id"], b["prim_id"]) and np.array_equal(a["inst_id"], b["inst_id"]) and
            np.array_equal(a["t"].view(np.uint32), b["t"].view(np.uint32)) and
            np.array_equal(a["bary"].view(np.uint32), b["bary"].view(np.uint32)))


def test_noalloc_counted_trace_equals_the_allocating_one():
    sc = scenes.scene_instanced_small()
    ctx = Context(0)
    try:
        ctx.upload(sc)
        for n, seed in ((60_000, 61), (7_000, 62)):
            print(f"[noalloc] {n} rays, seed {seed}", flush=True)
            rays = random_rays(n, (-10, -10, -10), (110, 110, 110), (50, 50, 50), 30.0, seed=seed)
            want, wc = ctx.trace_closest(rays, counted=True)
            got, gc = _noalloc(ctx, rays)
            assert _same_hits(got, want)
            assert gc == wc, (gc, wc)
            assert gc["rays"] == n and gc["tri_tests"] > 0 and gc["tlas_pops"] > 0
        print("[noalloc] empty call", flush=True)
        _, ec = _noalloc(ctx, np.zeros(0, abi.RAY_DTYPE))
        assert all(v == 0 for v in ec.values())
        # an instance edit leaves the TLAS to rebuild: refused, then fine once rebuilt
        m = np.array(sc.instance_model[0], np.float32).reshape(4, 4)
        m[3, :3] += np.float32(1.0)
        ctx.set_instance_transform(0, m.reshape(-1))
        rays = random_rays(5_000, (-10, -10, -10), (110, 110, 110), (50, 50, 50), 30.0, seed=63)
        with pytest.raises(abi.KptError) as ei:
            _noalloc(ctx, rays)
        assert ei.value.code == abi.KPT_ERR_STATE
        want, wc = ctx.trace_closest(rays, counted=True)        # rebuilds the TLAS
        got, gc = _noalloc(ctx, rays)
        assert _same_hits(got, want) and gc == wc
    finally:
        ctx.close()
