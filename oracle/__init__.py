"""ctypes binding of the CPU oracle (oracle/_build/libkpt_oracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the reference results are compared with; the product never imports
it.  Build with ``make -C oracle`` (or ``oracle.build()``).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from kittlespt_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
# KPT_ORACLE_LIB selects another build (the sanitizer build: make -C oracle sanitize)
LIB_PATH = os.environ.get("KPT_ORACLE_LIB") or os.path.join(HERE, "_build", "libkpt_oracle.so")
_lib = None

MODE_CLOSEST, MODE_ANY, MODE_BRUTE_CLOSEST, MODE_BRUTE_ANY, MODE_CLOSEST_ENTRY_BOUND = 0, 1, 2, 3, 4
FN = {"exp": 0, "exp2": 1, "log2": 2, "pow": 3, "sin": 4, "cos": 5, "tan": 6, "acos": 7, "log": 8}


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        L.orc_scene_create.restype = P
        L.orc_scene_create.argtypes = [C.POINTER(abi.KptSceneDesc)]
        L.orc_scene_destroy.argtypes = [P]
        L.orc_scene_info.argtypes = [P, C.POINTER(C.c_int64)]
        L.orc_get_blas.restype = C.c_int64
        L.orc_get_blas.argtypes = [P, C.c_int32, P, C.c_int64, C.POINTER(C.c_int32), C.c_int64]
        L.orc_get_tlas.restype = C.c_int64
        L.orc_get_tlas.argtypes = [P, P, C.c_int64]
        L.orc_trace.argtypes = [P, P, C.c_int64, C.c_int32, P, P, C.POINTER(abi.KptTraceCounters),
                                C.POINTER(C.c_int32), C.c_int32]
        L.orc_render_pixels.argtypes = [P, C.POINTER(abi.KptSettings), C.POINTER(abi.KptCamera), C.c_int32,
                                        C.c_int32, P, C.c_int64, C.c_int32, C.c_int32, C.c_int32, P, P, P,
                                        C.c_int32]
        L.orc_sample_pixels.argtypes = [P, C.POINTER(abi.KptSettings), C.POINTER(abi.KptCamera), C.c_int32,
                                        C.c_int32, P, C.c_int64, C.c_int32, P, C.c_int32]
        L.orc_dmath.argtypes = [C.c_int32, P, P, P, C.c_int64]
        L.orc_pcg_hash.restype = C.c_uint32
        L.orc_pcg_hash.argtypes = [C.c_uint32]
        L.orc_glm_inverse.argtypes = [P, P]
        L.orc_mat4_inverse.argtypes = [P, P]
        _lib = L
    return _lib


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p)


def threads_default():
    return max(1, min(16, os.cpu_count() or 1))


class OracleScene:
    """The oracle's copy of a scene (reference-layout trees built by its own
    restatement of BLASBuilder / TLASBuilder)."""

    def __init__(self, scene):
        desc, self._keep = scene.desc()
        self.h = lib().orc_scene_create(C.byref(desc))
        self.scene = scene

    def close(self):
        if self.h:
            lib().orc_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def info(self):
        out = (C.c_int64 * 8)()
        lib().orc_scene_info(self.h, out)
        return list(out)

    def blas(self, mesh):
        n = int(self.scene.mesh_count[mesh])
        nodes = np.zeros(max(1, 2 * n - 1), abi.BVH2NODE_DTYPE)
        tri = np.zeros(max(1, n), np.int32)
        cnt = lib().orc_get_blas(self.h, mesh, _ptr(nodes), len(nodes), tri.ctypes.data_as(C.POINTER(C.c_int32)),
                                 len(tri))
        if cnt < 0:
            raise RuntimeError("orc_get_blas failed")
        return nodes[:cnt], tri

    def tlas(self):
        nodes = np.zeros(2 * self.scene.num_instances + 2, abi.TLASNODE_DTYPE)
        cnt = lib().orc_get_tlas(self.h, _ptr(nodes), len(nodes))
        return nodes[:cnt]

    def trace(self, rays, mode=MODE_CLOSEST, threads=None):
        rays = np.ascontiguousarray(rays, abi.RAY_DTYPE)
        n = len(rays)
        hits = np.zeros(n, abi.HIT_DTYPE)
        occ = np.zeros(n, np.int32)
        cnt = abi.KptTraceCounters()
        ms = C.c_int32(0)
        lib().orc_trace(self.h, _ptr(rays), n, mode, _ptr(hits), _ptr(occ), C.byref(cnt), C.byref(ms),
                        threads or threads_default())
        return (hits if mode in (MODE_CLOSEST, MODE_BRUTE_CLOSEST, MODE_CLOSEST_ENTRY_BOUND) else occ), cnt.as_dict(), \
            ms.value

    def render_pixels(self, settings, camera, width, height, pixels, frame0, spp, double=False, threads=None,
                      want_buffers=False):
        pixels = np.ascontiguousarray(pixels, np.int32)
        n = len(pixels)
        rgb = np.zeros((n, 3), np.float32)
        g = np.zeros((n, 4), np.uint32) if want_buffers else None
        v = np.zeros((n, 4), np.float32) if want_buffers else None
        lib().orc_render_pixels(self.h, C.byref(settings), C.byref(camera), width, height, _ptr(pixels), n, frame0,
                                spp, 1 if double else 0, _ptr(rgb), _ptr(g) if g is not None else None,
                                _ptr(v) if v is not None else None, threads or threads_default())
        return (rgb, g, v) if want_buffers else rgb

    def render_cache_queries(self, settings, camera, width, height, frame, nee=False, wide_roughness=0.0,
                             threads=None):
        """Paths ending in the cache at bounce 1: (L_before [n,3], queries [n,12])."""
        n = width * height
        Lb = np.zeros((n, 3), np.float32)
        q = np.zeros((n, 12), np.float32)
        L = lib()
        L.orc_render_cache_queries.argtypes = [C.c_void_p, C.POINTER(abi.KptSettings), C.POINTER(abi.KptCamera),
                                               C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_void_p,
                                               C.c_void_p, C.c_int32]
        L.orc_render_cache_queries(self.h, C.byref(settings), C.byref(camera), width, height, frame, int(nee),
                                   float(wide_roughness), _ptr(Lb),
                                   _ptr(q), threads or threads_default())
        return Lb, q

    def set_instance_transform(self, i, model16):
        m = np.ascontiguousarray(model16, np.float32)
        L = lib()
        L.orc_set_instance_transform.argtypes = [C.c_void_p, C.c_int32, C.c_void_p]
        L.orc_set_instance_transform(self.h, i, _ptr(m))

    def render_temporal(self, settings, camera, width, height, frame, prev_accum, prev_gbuffer, threads=None):
        """One temporal frame; returns (render [H,W,4], gbuffer, vbuffer)."""
        pa = np.ascontiguousarray(prev_accum, np.float32)
        pg = np.ascontiguousarray(prev_gbuffer, np.uint32)
        r = np.zeros((height, width, 4), np.float32)
        g = np.zeros((height, width, 4), np.uint32)
        v = np.zeros((height, width, 4), np.float32)
        L = lib()
        L.orc_render_temporal.argtypes = [C.c_void_p, C.POINTER(abi.KptSettings), C.POINTER(abi.KptCamera),
                                          C.c_int32, C.c_int32, C.c_int32] + [C.c_void_p] * 5 + [C.c_int32]
        L.orc_render_temporal(self.h, C.byref(settings), C.byref(camera), width, height, frame, _ptr(pa), _ptr(pg),
                              _ptr(r), _ptr(g), _ptr(v), threads or threads_default())
        return r, g, v

    def render_temporal_pixels(self, settings, camera, width, height, frame, prev_accum, prev_gbuffer, pixels,
                               threads=None):
        """One temporal frame for the listed pixels only; returns (render
        [n,4], gbuffer [n,4], vbuffer [n,4])."""
        pa = np.ascontiguousarray(prev_accum, np.float32)
        pg = np.ascontiguousarray(prev_gbuffer, np.uint32)
        px = np.ascontiguousarray(pixels, np.int32)
        n = len(px)
        r = np.zeros((n, 4), np.float32)
        g = np.zeros((n, 4), np.uint32)
        v = np.zeros((n, 4), np.float32)
        f = _sig("orc_render_temporal_pixels", None,
                 [C.c_void_p, C.POINTER(abi.KptSettings), C.POINTER(abi.KptCamera), C.c_int32, C.c_int32,
                  C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.c_int32])
        f(self.h, C.byref(settings), C.byref(camera), width, height, frame, _ptr(pa), _ptr(pg), _ptr(px), n,
          _ptr(r), _ptr(g), _ptr(v), threads or threads_default())
        return r, g, v

    def sample_pixels(self, settings, camera, width, height, pixels, frame, threads=None):
        pixels = np.ascontiguousarray(pixels, np.int32)
        rgb = np.zeros((len(pixels), 3), np.float32)
        lib().orc_sample_pixels(self.h, C.byref(settings), C.byref(camera), width, height, _ptr(pixels),
                                len(pixels), frame, _ptr(rgb), threads or threads_default())
        return rgb


def dmath(fn, x, y=None):
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y if y is not None else np.zeros_like(x), np.float32)
    out = np.zeros_like(x)
    lib().orc_dmath(FN[fn], _ptr(x), _ptr(y), _ptr(out), len(x))
    return out


def pcg_hash(v):
    return int(lib().orc_pcg_hash(int(v) & 0xFFFFFFFF))


def glm_inverse(m16):
    m = np.ascontiguousarray(m16, np.float32)
    out = np.zeros(16, np.float32)
    lib().orc_glm_inverse(_ptr(m), _ptr(out))
    return out


def mat4_inverse(m16):
    m = np.ascontiguousarray(m16, np.float32)
    out = np.zeros(16, np.float32)
    lib().orc_mat4_inverse(_ptr(m), _ptr(out))
    return out


def _f32(a):
    return np.ascontiguousarray(a, np.float32)


def bsdf_eval(params, wo, wi, double=False):
    p, wo, wi = _f32(params), _f32(wo), _f32(wi)
    n = len(wo)
    f = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    L = lib()
    L.orc_bsdf_eval.argtypes = [C.c_void_p] * 3 + [C.c_int64, C.c_int32, C.c_void_p, C.c_void_p]
    L.orc_bsdf_eval(_ptr(p), _ptr(wo), _ptr(wi), n, 1 if double else 0, _ptr(f), _ptr(pdf))
    return f, pdf


def bsdf_sample(params, wo, u, double=False):
    p, wo, u = _f32(params), _f32(wo), _f32(u)
    n = len(wo)
    wi = np.zeros((n, 3), np.float32)
    f = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    fl = np.zeros(n, np.int32)
    L = lib()
    L.orc_bsdf_sample.argtypes = [C.c_void_p] * 3 + [C.c_int64, C.c_int32] + [C.c_void_p] * 4
    L.orc_bsdf_sample(_ptr(p), _ptr(wo), _ptr(u), n, 1 if double else 0, _ptr(wi), _ptr(f), _ptr(pdf), _ptr(fl))
    return wi, f, pdf, fl


def ggx_d(alpha, wm, double=False):
    wm = _f32(wm)
    out = np.zeros(len(wm), np.float32)
    L = lib()
    L.orc_ggx_d.argtypes = [C.c_float, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p]
    L.orc_ggx_d(alpha, _ptr(wm), len(wm), 1 if double else 0, _ptr(out))
    return out


def fresnel_dielectric(cos_t, ior, double=False):
    c, i = _f32(cos_t), _f32(ior)
    out = np.zeros(len(c), np.float32)
    L = lib()
    L.orc_fresnel_dielectric.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p]
    L.orc_fresnel_dielectric(_ptr(c), _ptr(i), len(c), 1 if double else 0, _ptr(out))
    return out


def sky_li(o, d, sun, emission=1.0, double=False):
    o, d, s = _f32(o), _f32(d), _f32(sun)
    out = np.zeros((len(o), 3), np.float32)
    L = lib()
    L.orc_sky_li.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_float, C.c_int32, C.c_void_p]
    L.orc_sky_li(_ptr(o), _ptr(d), len(o), _ptr(s), emission, 1 if double else 0, _ptr(out))
    return out


def camera_rays(cam, width, height, pixels):
    px = np.ascontiguousarray(pixels, np.int32)
    o = np.zeros((len(px), 3), np.float32)
    d = np.zeros((len(px), 3), np.float32)
    L = lib()
    L.orc_camera_rays.argtypes = [C.POINTER(abi.KptCamera), C.c_int32, C.c_int32, C.c_void_p, C.c_int64,
                                  C.c_void_p, C.c_void_p]
    L.orc_camera_rays(C.byref(cam), width, height, _ptr(px), len(px), _ptr(o), _ptr(d))
    return o, d


# ---------------------------------------------------------------- post chain
def _sig(name, res, args):
    f = getattr(lib(), name)
    f.restype = res
    f.argtypes = args
    return f


def post_frame(estimate, gbuffer, settings, exposure, film, frame_delta_ms=16.0, out_size=None):
    """One frame of the post chain (modulate, auto exposure, upscale, bloom,
    post) over the demodulated estimate [H,W,4] and gbuffer [H,W,4] u32 at the
    render size; out_size = (output width, output height) (default: the
    render size; the upscale runs when it differs or settings.upscale_enable
    is set).  exposure and film (KptExposure / KptFilmState) are updated in
    place.  Returns (render [H,W,4], output, bloom_mip0, histogram), the last
    two at the output size."""
    H, W = estimate.shape[:2]
    OW, OH = out_size if out_size is not None else (W, H)
    est = _f32(estimate)
    gb = np.ascontiguousarray(gbuffer, np.uint32)
    render = np.zeros((H, W, 4), np.float32)
    out = np.zeros((OH, OW, 4), np.float32)
    bloom = np.zeros((OH, OW, 4), np.float32)
    hist = np.zeros(256, np.uint64)
    f = _sig("orc_post_frame_sized", None, [C.c_void_p, C.c_void_p] + [C.c_int32] * 5 +
             [C.POINTER(abi.KptSettings), C.c_float, C.POINTER(abi.KptExposure), C.POINTER(abi.KptFilmState)] +
             [C.c_void_p] * 4)
    f(_ptr(est), _ptr(gb), W, H, OW, OH, int(settings.upscale_enable), C.byref(settings), frame_delta_ms,
      C.byref(exposure), C.byref(film), _ptr(render), _ptr(out), _ptr(bloom), _ptr(hist))
    return render, out, bloom, hist


def upscale_fsr(src, out_w, out_h, easu=False):
    """FSR 1.0 (the reference's USE_FSR build, upscale.cuh:13-408) of an
    [H, W, 4] image to [out_h, out_w, 4]: EASU then RCAS; with easu=True also
    returns the EASU pass alone."""
    s = _f32(src)
    H, W = s.shape[:2]
    out = np.zeros((out_h, out_w, 4), np.float32)
    e = np.zeros((out_h, out_w, 4), np.float32) if easu else None
    _sig("orc_upscale_fsr", None, [C.c_void_p] + [C.c_int32] * 4 + [C.c_void_p, C.c_void_p])(
        _ptr(s), W, H, out_w, out_h, _ptr(out), _ptr(e) if easu else None)
    return (out, e) if easu else out


def fsr_rcas(src):
    """The RCAS pass alone (upscale.cuh:316-378, sharpness 0) of an [H, W, 4] image."""
    s = _f32(src)
    H, W = s.shape[:2]
    out = np.zeros((H, W, 4), np.float32)
    _sig("orc_fsr_rcas", None, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p])(_ptr(s), W, H, _ptr(out))
    return out


def upscale_bilinear(src, out_w, out_h):
    """upscale (upscale.cuh:287-292) of an [H, W, 4] image to [out_h, out_w, 4]."""
    s = _f32(src)
    H, W = s.shape[:2]
    out = np.zeros((out_h, out_w, 4), np.float32)
    _sig("orc_upscale_bilinear", None, [C.c_void_p] + [C.c_int32] * 4 + [C.c_void_p])(
        _ptr(s), W, H, out_w, out_h, _ptr(out))
    return out


def set_exposure(exposure):
    film = abi.KptFilmState()
    _sig("orc_set_exposure", None, [C.POINTER(abi.KptExposure), C.POINTER(abi.KptFilmState)])(
        C.byref(exposure), C.byref(film))
    return film


def exposure_program(exposure, avg_lum):
    _sig("orc_exposure_program", None, [C.POINTER(abi.KptExposure), C.c_float])(C.byref(exposure), avg_lum)
    return exposure


def lum_to_bin(lum, black_ev=-10.0, white_ev=6.5):
    x = _f32(lum)
    out = np.zeros(len(x), np.uint32)
    _sig("orc_lum_to_bin", None, [C.c_void_p, C.c_int64, C.c_float, C.c_float, C.c_void_p])(
        _ptr(x), len(x), black_ev, white_ev, _ptr(out))
    return out


def meter_weight(width, height, xy):
    xy = np.ascontiguousarray(xy, np.int32)
    out = np.zeros(len(xy), np.float32)
    _sig("orc_meter_weight", None, [C.c_int32, C.c_int32, C.c_void_p, C.c_int64, C.c_void_p])(
        width, height, _ptr(xy), len(xy), _ptr(out))
    return out


def histogram(rgba, white_ev=6.5, black_ev=-10.0):
    img = _f32(rgba)
    H, W = img.shape[:2]
    bins = np.zeros(256, np.uint64)
    _sig("orc_histogram", None, [C.c_void_p, C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_void_p])(
        _ptr(img), W, H, white_ev, black_ev, _ptr(bins))
    return bins


def histogram_average(bins, white_ev=6.5, black_ev=-10.0, last=0.0, dt_ms=16.0):
    b = np.ascontiguousarray(bins, np.uint64)
    return _sig("orc_histogram_average", C.c_float, [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_float])(
        _ptr(b), white_ev, black_ev, last, dt_ms)


def post_effects(back):
    """computeEffects (kernels.cu:508-528): fisheye + chromatic aberration of
    an [H, W, 4] back buffer (row 0 = the buffer's first row); returns the
    [H, W, 4] output."""
    b = np.ascontiguousarray(np.asarray(back, np.float32))
    H, W = b.shape[:2]
    out = np.zeros_like(b)
    _sig("orc_post_effects", None, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p])(_ptr(b), W, H, _ptr(out))
    return out


def film(rgb, scalar=1.0, white_ev=6.5, black_ev=-10.0):
    x = _f32(rgb).reshape(-1, 3)
    out = np.zeros_like(x)
    _sig("orc_film", None, [C.c_void_p, C.c_int64, C.c_float, C.c_float, C.c_float, C.c_void_p])(
        _ptr(x), len(x), scalar, white_ev, black_ev, _ptr(out))
    return out


def agx_curve(x):
    x = _f32(x)
    out = np.zeros_like(x)
    _sig("orc_agx_curve", None, [C.c_void_p, C.c_int64, C.c_void_p])(_ptr(x), x.size, _ptr(out))
    return out


def srgb_encode(x):
    x = _f32(x)
    out = np.zeros_like(x)
    _sig("orc_srgb_encode", None, [C.c_void_p, C.c_int64, C.c_void_p])(_ptr(x), x.size, _ptr(out))
    return out


def yxy_roundtrip(rgb, scalar=1.0):
    x = _f32(rgb).reshape(-1, 3)
    yxy = np.zeros_like(x)
    back = np.zeros_like(x)
    _sig("orc_yxy_roundtrip", None, [C.c_void_p, C.c_int64, C.c_float, C.c_void_p, C.c_void_p])(
        _ptr(x), len(x), scalar, _ptr(yxy), _ptr(back))
    return yxy, back


def bilinear(rgba, xy, filter_alpha=True):
    img = _f32(rgba)
    H, W = img.shape[:2]
    xy = _f32(xy).reshape(-1, 2)
    out = np.zeros((len(xy), 4), np.float32)
    _sig("orc_bilinear", None, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p])(
        _ptr(img), W, H, _ptr(xy), len(xy), 1 if filter_alpha else 0, _ptr(out))
    return out


def bloom_down(src, karis=False):
    s = _f32(src)
    sh, sw = s.shape[:2]
    dst = np.zeros((sh >> 1, sw >> 1, 4), np.float32)
    _sig("orc_bloom_down", None, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_int32, C.c_int32, C.c_int32])(
        _ptr(s), sw, sh, _ptr(dst), sw >> 1, sh >> 1, 1 if karis else 0)
    return dst


def bloom_up(src, dst, blend=0.65):
    s = _f32(src)
    d = np.array(dst, np.float32, copy=True, order="C")
    sh, sw = s.shape[:2]
    dh, dw = d.shape[:2]
    _sig("orc_bloom_up", None, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_int32, C.c_int32, C.c_float])(
        _ptr(s), sw, sh, _ptr(d), dw, dh, blend)
    return d


def bloom_levels(width, height):
    return int(_sig("orc_bloom_levels", C.c_int32, [C.c_int32, C.c_int32])(width, height))


# ---------------------------------------------------------------- radiance cache network
def _enc(bmin, bmax, scale=1.0):
    bmin = np.asarray(bmin, np.float32)
    ext = np.asarray(bmax, np.float32) - bmin
    binv = np.float32(1.0) / ext
    return np.concatenate([bmin, binv, [np.float32(scale)]]).astype(np.float32)


def cache_encode(pos, dirs, bmin, bmax):
    p, d = _f32(pos).reshape(-1, 4), _f32(dirs).reshape(-1, 4)
    out = np.zeros((len(p), 32), np.float32)
    e = _enc(bmin, bmax)
    _sig("orc_cache_encode", None, [C.c_void_p] * 2 + [C.c_int64, C.c_void_p, C.c_void_p])(
        _ptr(p), _ptr(d), len(p), _ptr(e), _ptr(out))
    return out


def cache_forward(w, pos, dirs, bmin, bmax, mode=0):
    w, p, d = _f32(w), _f32(pos).reshape(-1, 4), _f32(dirs).reshape(-1, 4)
    out = np.zeros((len(p), 3), np.float64)
    e = _enc(bmin, bmax)
    _sig("orc_cache_forward", None, [C.c_void_p] * 3 + [C.c_int64, C.c_void_p, C.c_int32, C.c_void_p])(
        _ptr(w), _ptr(p), _ptr(d), len(p), _ptr(e), mode, _ptr(out))
    return out


def cache_loss_grad(w, pos, dirs, targets, bmin, bmax, mode=0, w_den=None, want_grad=True):
    w, p, d, t = _f32(w), _f32(pos).reshape(-1, 4), _f32(dirs).reshape(-1, 4), _f32(targets).reshape(-1, 4)
    wd = _f32(w_den) if w_den is not None else None
    g = np.zeros(len(w), np.float64) if want_grad else None
    e = _enc(bmin, bmax)
    f = _sig("orc_cache_loss_grad", C.c_double, [C.c_void_p] * 5 + [C.c_int64, C.c_void_p, C.c_int32, C.c_void_p])
    loss = f(_ptr(w), _ptr(wd) if wd is not None else None, _ptr(p), _ptr(d), _ptr(t), len(p), _ptr(e), mode,
             _ptr(g) if g is not None else None)
    return loss, g


def cache_adam(w, m, v, grad, step, lr):
    w, m, v = (np.array(a, np.float32, copy=True) for a in (w, m, v))
    g = _f32(grad)
    _sig("orc_cache_adam", None, [C.c_void_p] * 4 + [C.c_int64, C.c_float])(
        _ptr(w), _ptr(m), _ptr(v), _ptr(g), step, lr)
    return w, m, v


# ---------------------------------------------------------------- shading probes
def shade_probe(scene_handle, rays):
    """getSurfaceInteraction + Material::getBSDF at each ray's closest hit:
    [n, 24] floats (see orc_shade_probe)."""
    rays = np.ascontiguousarray(rays, abi.RAY_DTYPE)
    out = np.zeros((len(rays), 24), np.float32)
    _sig("orc_shade_probe", None, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p])(
        scene_handle.h, _ptr(rays), len(rays), _ptr(out))
    return out


def light_sample(scene_handle, light, from_pos, u):
    """Light::sampleLi of light `light` from positions [n, 3] with numbers
    [n, 2]: ([n, 12] point, normal, pdf, L, uv; [n] pdf_Li of that point)."""
    f, uu = _f32(from_pos).reshape(-1, 3), _f32(u).reshape(-1, 2)
    out = np.zeros((len(f), 12), np.float32)
    pli = np.zeros(len(f), np.float32)
    _sig("orc_light_sample", None, [C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p,
                                    C.c_void_p])(scene_handle.h, light, _ptr(f), _ptr(uu), len(f), _ptr(out),
                                                 _ptr(pli))
    return out, pli


def encode_unit(v):
    v = _f32(v).reshape(-1, 3)
    out = np.zeros((len(v), 2), np.float32)
    _sig("orc_encode_unit", None, [C.c_void_p, C.c_int64, C.c_void_p])(_ptr(v), len(v), _ptr(out))
    return out


def float_to_byte(v):
    v = _f32(v)
    out = np.zeros(v.size, np.uint32)
    _sig("orc_float_to_byte", None, [C.c_void_p, C.c_int64, C.c_void_p])(_ptr(v), v.size, _ptr(out))
    return out


OP_PHASES = ("traversal", "sky_primary", "sky_later", "sky_sun", "shade_primary", "shade_later", "camera_output")
OP_KINDS = ("arith", "sqrt", "transcendental", "minmax")


def sample_events(scene_handle, settings, camera, width, height, pixels, frame=0, threads=None):
    """Traversal events of one sample of each listed pixel in frame `frame`,
    every closest-hit and shadow ray of its path, traversed as the reference
    does: an (n, 6) int64 array of BLAS interior pops, BLAS leaf pops,
    triangle tests, BLAS entries, TLAS pops and rays traced."""
    px = np.ascontiguousarray(pixels, np.int32)
    out = np.zeros((len(px), 6), np.int64)
    _sig("orc_sample_events", None, [C.c_void_p, C.POINTER(abi.KptSettings), C.POINTER(abi.KptCamera), C.c_int32,
                                     C.c_int32, C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_void_p])(
        scene_handle.h, C.byref(settings), C.byref(camera), width, height, _ptr(px), len(px), frame,
        threads or threads_default(), _ptr(out))
    return out


def heat_map(scene_handle, settings, camera, width, height, pixels, frame=0, entry_bound=True):
    """Per-pixel DebugData of one sample (blas.cuh:39-42): an (n, 2) int64
    array of BLAS node pops and TLAS node pops over the listed pixels'
    closest-hit traces in frame `frame` (shadow rays not counted).
    entry_bound: BLASes entered with the closest hit so far, as the GPU
    kernels enter them; False: the reference's traversal."""
    px = np.ascontiguousarray(pixels, np.int32)
    out = np.zeros((len(px), 2), np.int64)
    _sig("orc_heat_map", None, [C.c_void_p, C.POINTER(abi.KptSettings), C.POINTER(abi.KptCamera), C.c_int32,
                                C.c_int32, C.c_void_p, C.c_int64, C.c_int32, C.c_int32, C.c_void_p])(
        scene_handle.h, C.byref(settings), C.byref(camera), width, height, _ptr(px), len(px), frame,
        1 if entry_bound else 0, _ptr(out))
    return out


def count_ops(scene_handle, settings, camera, width, height, pixels, frame=0):
    """Algorithmic FP32 op counts per sample (the oracle instantiated with a
    counting float): {phase: {kind: ops per sample}} over the listed pixels'
    samples of frame `frame`."""
    px = np.ascontiguousarray(pixels, np.int32)
    out = np.zeros(len(OP_PHASES) * len(OP_KINDS), np.uint64)
    _sig("orc_count_ops", None, [C.c_void_p, C.POINTER(abi.KptSettings), C.POINTER(abi.KptCamera), C.c_int32,
                                 C.c_int32, C.c_void_p, C.c_int64, C.c_int32, C.c_void_p])(
        scene_handle.h, C.byref(settings), C.byref(camera), width, height, _ptr(px), len(px), frame, _ptr(out))
    o = out.reshape(len(OP_PHASES), len(OP_KINDS)).astype(np.float64) / max(1, len(px))
    return {ph: dict(zip(OP_KINDS, map(float, o[i]))) for i, ph in enumerate(OP_PHASES)}
