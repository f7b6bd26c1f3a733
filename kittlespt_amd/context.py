"""Python handle over one kpt_ctx (include/kpt_abi.h).

Device memory for the trace entry points comes straight from the HIP runtime
(hipMalloc / hipMemcpy through ctypes); no framework tensors are involved.
"""
from __future__ import annotations

import ctypes as C
import ctypes.util
import os

import numpy as np

from . import abi

_hip = None


def hip():
    global _hip
    if _hip is None:
        cands = [os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"),
                 ctypes.util.find_library("amdhip64") or "libamdhip64.so"]
        err = None
        for c in cands:
            try:
                _hip = C.CDLL(c)
                break
            except OSError as e:  # pragma: no cover
                err = e
        if _hip is None:
            raise RuntimeError(f"HIP runtime not found: {err}")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


H2D, D2H = 1, 2


class DeviceBuffer:
    def __init__(self, nbytes):
        self.ptr = C.c_void_p()
        self.nbytes = max(16, int(nbytes))
        rc = hip().hipMalloc(C.byref(self.ptr), self.nbytes)
        if rc != 0:
            raise RuntimeError(f"hipMalloc({self.nbytes}) failed: {rc}")

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        b.upload(a)
        return b

    def upload(self, a):
        a = np.ascontiguousarray(a)
        rc = hip().hipMemcpy(self.ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy H2D failed: {rc}")

    def download(self, out):
        rc = hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, out.nbytes, D2H)
        if rc != 0:
            raise RuntimeError(f"hipMemcpy D2H failed: {rc}")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Context:
    """One renderer context.  device=-2 makes a host-only context (builders
    and queries work, GPU entry points return KPT_ERR_NO_DEVICE)."""

    def __init__(self, device=-1):
        self.L = abi.lib()
        h = C.c_void_p()
        rc = self.L.kpt_ctx_create(device, C.byref(h))
        if rc != abi.KPT_OK:
            raise abi.KptError(rc, "kpt_ctx_create failed")
        self.h = h
        self.scene = None

    def close(self):
        if self.h:
            self.L.kpt_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, rc):
        return abi.check(self.h, rc)

    # ---- scene
    def upload(self, scene):
        desc, keep = scene.desc()
        self._ck(self.L.kpt_upload_scene(self.h, C.byref(desc)))
        self.scene = scene

    def info(self):
        out = (C.c_int64 * 8)()
        self._ck(self.L.kpt_scene_info(self.h, out))
        return list(out)

    def blas(self, mesh):
        n = int(self.scene.mesh_count[mesh])
        nodes = np.zeros(max(1, 2 * n - 1), abi.BVH2NODE_DTYPE)
        tri = np.zeros(max(1, n), np.int32)
        cnt = (C.c_int64 * 2)()
        self._ck(self.L.kpt_get_blas(self.h, mesh, nodes.ctypes.data_as(C.c_void_p), len(nodes),
                                     tri.ctypes.data_as(C.POINTER(C.c_int32)), len(tri), cnt))
        return nodes[:cnt[0]], tri[:cnt[1]]

    def tlas(self):
        nodes = np.zeros(2 * self.scene.num_instances + 2, abi.TLASNODE_DTYPE)
        cnt = C.c_int64()
        self._ck(self.L.kpt_get_tlas(self.h, nodes.ctypes.data_as(C.c_void_p), len(nodes), C.byref(cnt)))
        return nodes[:cnt.value]

    def set_instance_transform(self, i, model16):
        m = np.ascontiguousarray(model16, np.float32)
        self._ck(self.L.kpt_set_instance_transform(self.h, i, abi.fptr(m)))

    def set_instance_transforms(self, first, models):
        m = np.ascontiguousarray(models, np.float32).reshape(-1, 16)
        self._ck(self.L.kpt_set_instance_transforms(self.h, first, len(m), abi.fptr(m)))

    # ---- traversal (host arrays in, host arrays out)
    def trace_closest(self, rays, counted=False):
        rays = np.ascontiguousarray(rays, abi.RAY_DTYPE)
        n = len(rays)
        dr = DeviceBuffer.from_array(rays)
        dh = DeviceBuffer(n * abi.HIT_DTYPE.itemsize)
        cnt = abi.KptTraceCounters()
        if counted:
            self._ck(self.L.kpt_trace_closest_counted(self.h, dr.ptr, dh.ptr, n, C.byref(cnt)))
        else:
            self._ck(self.L.kpt_trace_closest(self.h, dr.ptr, dh.ptr, n))
        self._ck(self.L.kpt_sync(self.h))
        hits = dh.download(np.zeros(n, abi.HIT_DTYPE))
        return (hits, cnt.as_dict()) if counted else hits

    def trace_any(self, rays):
        rays = np.ascontiguousarray(rays, abi.RAY_DTYPE)
        n = len(rays)
        dr = DeviceBuffer.from_array(rays)
        do = DeviceBuffer(n * 4)
        self._ck(self.L.kpt_trace_any(self.h, dr.ptr, do.ptr, n))
        self._ck(self.L.kpt_sync(self.h))
        return do.download(np.zeros(n, np.int32))

    # ---- rendering
    def settings(self):
        s = abi.KptSettings()
        self._ck(self.L.kpt_get_settings(self.h, C.byref(s)))
        return s

    def set_settings(self, s):
        self._ck(self.L.kpt_set_settings(self.h, C.byref(s)))

    def set_view(self, proj16, view16):
        p = np.ascontiguousarray(proj16, np.float32)
        v = np.ascontiguousarray(view16, np.float32)
        self._ck(self.L.kpt_set_view(self.h, abi.fptr(p), abi.fptr(v)))

    def camera(self):
        c = abi.KptCamera()
        self._ck(self.L.kpt_get_camera(self.h, C.byref(c)))
        return c

    def resize(self, w, h):
        """Output size w x h; the render size is w x h, or w/2.5 x h/2.5 when
        the settings ask for upscaling (see resolution())."""
        self._ck(self.L.kpt_resize(self.h, w, h))
        self.width, self.height = w, h
        self.render_width, self.render_height = self.resolution()[:2]

    def resolution(self):
        """(render width, render height, output width, output height)."""
        r = (C.c_int32 * 4)()
        self._ck(self.L.kpt_get_resolution(self.h, r))
        return tuple(int(x) for x in r)

    def reset(self):
        self._ck(self.L.kpt_reset_accumulation(self.h))

    def render(self, spp=1):
        self._ck(self.L.kpt_render(self.h, spp))

    def frame_index(self):
        v = C.c_int32()
        self._ck(self.L.kpt_get_frame_index(self.h, C.byref(v)))
        return v.value

    def readback(self, which=0):
        """0 render, 1 accumulation, 2 G-buffer, 3 V-buffer (render size),
        4 output (output size)."""
        dt = np.uint32 if which == 2 else np.float32
        rw, rh, ow, oh = self.resolution()
        w, h = (ow, oh) if which == 4 else (rw, rh)
        out = np.zeros((h, w, 4), dt)
        self._ck(self.L.kpt_readback(self.h, which, out.ctypes.data_as(C.c_void_p), out.nbytes))
        return out

    def timing(self):
        out = (C.c_double * 4)()
        self._ck(self.L.kpt_last_timing(self.h, out))
        return list(out)

    def build_timing(self):
        """(ms, on_device) of the last upload's BLAS builds."""
        out = (C.c_double * 2)()
        self._ck(self.L.kpt_last_build_timing(self.h, out))
        return out[0], bool(out[1])

    TUNING_KEYS = ("refill", "chunk", "refill_min", "blocks_per_cu", "sort", "graph", "ray_sort", "tail",
                   "tlas_device", "blas_device", "stack", "block", "suspend", "overlap")

    def tuning(self):
        n = len(self.TUNING_KEYS)
        k = (C.c_int32 * n)()
        self._ck(self.L.kpt_get_tuning(self.h, k, n))
        return dict(zip(self.TUNING_KEYS, list(k)))

    def set_tuning(self, **kw):
        cur = self.tuning()
        cur.update(kw)
        n = len(self.TUNING_KEYS)
        k = (C.c_int32 * n)(*[int(cur[x]) for x in self.TUNING_KEYS])
        self._ck(self.L.kpt_set_tuning(self.h, k, n))

    def set_trace_counting(self, enable):
        self._ck(self.L.kpt_set_trace_counting(self.h, 1 if enable else 0))

    def render_counters(self):
        c = abi.KptTraceCounters()
        self._ck(self.L.kpt_get_render_counters(self.h, C.byref(c)))
        return c.as_dict()

    STAGES = {"trace": 0, "tail": 1, "shade": 2, "sky": 3, "cache_infer": 4, "cache_train": 5}

    def heat_map(self, rgba=False):
        """DebugData heat map of the last counted wavefront render's last
        frame (blas.cuh:39-42): (H, W, 2) int32 BLAS and TLAS node pops of
        each pixel's closest-hit traces, or with rgba=True the reference's
        colours (kernels.cu:442-443) as (H, W, 4) float32."""
        W, H = self.resolution()[:2]
        if rgba:
            out = np.zeros((H, W, 4), np.float32)
            self._ck(self.L.kpt_get_heat_map(self.h, None, out.ctypes.data_as(C.c_void_p), W * H))
        else:
            out = np.zeros((H, W, 2), np.int32)
            self._ck(self.L.kpt_get_heat_map(self.h, out.ctypes.data_as(C.c_void_p), None, W * H))
        return out

    def stage_stats(self, stage):
        """Per-stage measurements of the last wavefront render: 'trace' (the
        closest-hit launches) or 'tail' (the path-per-lane tail): HIP-event
        ms and launch count, and in counting mode the traversal events (and,
        for the tail, the paths and vertices it processed)."""
        st = abi.KptStageStats()
        self._ck(self.L.kpt_get_stage_stats(self.h, self.STAGES[stage], C.byref(st)))
        return st.as_dict()

    # ---- post chain
    def set_exposure(self, exposure):
        """Renderer::setExposure: a KptExposure (abi.KptExposure.camera(...))."""
        self._ck(self.L.kpt_set_exposure(self.h, C.byref(exposure)))

    def exposure(self):
        e = abi.KptExposure()
        self._ck(self.L.kpt_get_exposure(self.h, C.byref(e)))
        return e

    def film(self):
        f = abi.KptFilmState()
        self._ck(self.L.kpt_get_film(self.h, C.byref(f)))
        return f

    def set_frame_delta(self, ms):
        self._ck(self.L.kpt_set_frame_delta(self.h, float(ms)))

    # ---- batched views
    def render_views(self, views, spp=1):
        """views: list of (projection16, view16, width, height).  Returns a
        list of [H, W, 4] estimates (one wavefront over all views' pixels)."""
        arr = (abi.KptView * len(views))()
        total = 0
        for k, (proj, view, w, h) in enumerate(views):
            arr[k].projection[:] = [float(x) for x in np.asarray(proj, np.float32).reshape(-1)]
            arr[k].view[:] = [float(x) for x in np.asarray(view, np.float32).reshape(-1)]
            arr[k].width, arr[k].height = int(w), int(h)
            total += int(w) * int(h)
        out = DeviceBuffer(total * 16)
        self._ck(self.L.kpt_render_views(self.h, arr, len(views), spp, out.ptr))
        flat = out.download(np.zeros((total, 4), np.float32))
        res, off = [], 0
        for (_, _, w, h) in views:
            res.append(flat[off:off + w * h].reshape(h, w, 4))
            off += w * h
        return res

    # ---- radiance cache
    def cache_configure(self, enabled=False, train=False, train_samples=65536, train_steps=4, learning_rate=1e-3,
                        radiance_scale=1e4, seed=1, bounds=None, nee_at_vertex=1, wide_roughness=0.0):
        s = abi.KptCacheSettings()
        s.enabled, s.train, s.nee_at_vertex = int(enabled), int(train), int(nee_at_vertex)
        s.wide_roughness = float(wide_roughness)
        s.train_samples, s.train_steps = int(train_samples), int(train_steps)
        s.learning_rate, s.radiance_scale, s.seed = float(learning_rate), float(radiance_scale), int(seed)
        if bounds is not None:
            s.bounds_min[:] = [float(x) for x in bounds[0]]
            s.bounds_max[:] = [float(x) for x in bounds[1]]
        self._ck(self.L.kpt_cache_configure(self.h, C.byref(s)))
        return s

    def cache_settings(self):
        s = abi.KptCacheSettings()
        self._ck(self.L.kpt_cache_get_settings(self.h, C.byref(s)))
        return s

    def cache_weights(self):
        w = np.zeros(abi.CACHE_PARAMS, np.float32)
        self._ck(self.L.kpt_cache_get_weights(self.h, abi.fptr(w), len(w)))
        return w

    def cache_set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        self._ck(self.L.kpt_cache_set_weights(self.h, abi.fptr(w), len(w)))

    def cache_gradient(self):
        g = np.zeros(abi.CACHE_PARAMS, np.float32)
        self._ck(self.L.kpt_cache_get_gradient(self.h, abi.fptr(g), len(g)))
        return g

    def cache_infer(self, pos, dirs):
        pa = np.ascontiguousarray(pos, np.float32).reshape(-1, 4)
        p = DeviceBuffer.from_array(pa)
        d = DeviceBuffer.from_array(np.ascontiguousarray(dirs, np.float32).reshape(-1, 4))
        n = len(pa)
        out = DeviceBuffer(max(16, n * 16))
        self._ck(self.L.kpt_cache_infer(self.h, p.ptr, d.ptr, n, out.ptr))
        self._ck(self.L.kpt_sync(self.h))
        return out.download(np.zeros((n, 4), np.float32))

    def cache_train(self, pos, dirs, targets, steps=1):
        pa = np.ascontiguousarray(pos, np.float32).reshape(-1, 4)
        p = DeviceBuffer.from_array(pa)
        d = DeviceBuffer.from_array(np.ascontiguousarray(dirs, np.float32).reshape(-1, 4))
        t = DeviceBuffer.from_array(np.ascontiguousarray(targets, np.float32).reshape(-1, 4))
        loss = C.c_float(0.0)
        self._ck(self.L.kpt_cache_train(self.h, p.ptr, d.ptr, t.ptr, len(pa), steps, C.byref(loss)))
        return loss.value

    def cache_queries(self):
        """(positions [n,4], directions [n,4] with bit-cast path index in w,
        throughputs [n,4]) of the last frame, in slot order."""
        cnt = C.c_int64(0)
        self._ck(self.L.kpt_cache_get_queries(self.h, None, None, None, 0, C.byref(cnt)))
        n = cnt.value
        p, d, b = (np.zeros((n, 4), np.float32) for _ in range(3))
        self._ck(self.L.kpt_cache_get_queries(self.h, p.ctypes.data_as(C.c_void_p), d.ctypes.data_as(C.c_void_p),
                                              b.ctypes.data_as(C.c_void_p), n, C.byref(cnt)))
        return p, d, b

    def cache_stats(self):
        out = (C.c_double * 5)()
        self._ck(self.L.kpt_cache_stats(self.h, out))
        return dict(zip(("loss", "samples", "queries", "train_records", "steps"), list(out)))

    def estimate(self):
        """Demodulated radiance estimate of static accumulation: the running
        sum divided by the frame count, as the trace kernels compute it before
        the post chain modulates the render buffer in place."""
        acc = self.readback(1)[..., :3]
        return acc / np.float32(max(1, self.frame_index()))
