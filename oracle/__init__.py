"""ctypes binding of the CPU oracle (oracle/_build/libkpt_oracle.so).

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the checker; the product never imports
it.  Build with ``make -C oracle`` (or ``oracle.build()``).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from kittlespt_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libkpt_oracle.so")
_lib = None

MODE_CLOSEST, MODE_ANY, MODE_BRUTE_CLOSEST, MODE_BRUTE_ANY = 0, 1, 2, 3
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
        return (hits if mode in (MODE_CLOSEST, MODE_BRUTE_CLOSEST) else occ), cnt.as_dict(), ms.value

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
