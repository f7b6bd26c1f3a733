"""ctypes view of include/kpt_abi.h and the loader of libkpt_hip.so.

The library is built in-tree (kittlespt_amd/_lib/libkpt_hip.so) by
``kittlespt_amd.build.build()``.  Loading fails loudly when it is missing:
there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib")
# KPT_LIB_PATH selects another build of the same library (the host-code
# sanitizer build the CPU tests can run against).
LIB_PATH = os.environ.get("KPT_LIB_PATH") or os.path.join(LIB_DIR, "libkpt_hip.so")

KPT_OK = 0
KPT_ERR_INVALID_ARG = 1
KPT_ERR_HIP = 2
KPT_ERR_OUT_OF_MEMORY = 3
KPT_ERR_STATE = 4
KPT_ERR_OVERFLOW = 5
KPT_ERR_NO_DEVICE = 6
KPT_ERR_UNSUPPORTED = 7

SEED_REFERENCE = 0
SEED_XOR = 1
INTEGRATOR_MEGAKERNEL = 0
INTEGRATOR_WAVEFRONT = 1


class KptMaterial(C.Structure):
    _fields_ = [
        ("albedo_texture_id", C.c_int32),
        ("albedo", C.c_float * 3),
        ("orm_texture_id", C.c_int32),
        ("metallic", C.c_float),
        ("roughness", C.c_float),
        ("transmission_texture_id", C.c_int32),
        ("transmission", C.c_float),
        ("ior", C.c_float),
        ("emission_texture_id", C.c_int32),
        ("emission_factor", C.c_float * 3),
        ("emission_scale_nits", C.c_float),
        ("normal_texture_id", C.c_int32),
        ("normal_scale", C.c_float),
    ]


class KptTextureDesc(C.Structure):
    _fields_ = [
        ("width", C.c_uint32),
        ("height", C.c_uint32),
        ("channels", C.c_uint32),
        ("reserved", C.c_uint32),
        ("offset", C.c_uint64),
    ]


class KptSceneDesc(C.Structure):
    _fields_ = [
        ("num_triangles", C.c_int32),
        ("positions", C.POINTER(C.c_float)),
        ("normals", C.POINTER(C.c_float)),
        ("uvs", C.POINTER(C.c_float)),
        ("material_ids", C.POINTER(C.c_int32)),
        ("num_meshes", C.c_int32),
        ("mesh_first", C.POINTER(C.c_int32)),
        ("mesh_count", C.POINTER(C.c_int32)),
        ("mesh_model", C.POINTER(C.c_float)),
        ("num_instances", C.c_int32),
        ("instance_mesh", C.POINTER(C.c_int32)),
        ("instance_model", C.POINTER(C.c_float)),
        ("num_materials", C.c_int32),
        ("materials", C.POINTER(KptMaterial)),
        ("num_textures", C.c_int32),
        ("textures", C.POINTER(KptTextureDesc)),
        ("pixels", C.POINTER(C.c_uint8)),
        ("num_pixel_bytes", C.c_uint64),
    ]


class KptRay(C.Structure):
    _fields_ = [("org", C.c_float * 3), ("tmin", C.c_float), ("dir", C.c_float * 3), ("tmax", C.c_float)]


class KptHit(C.Structure):
    _fields_ = [("t", C.c_float), ("prim_id", C.c_int32), ("inst_id", C.c_int32), ("bary", C.c_float * 3)]


class KptTraceCounters(C.Structure):
    _fields_ = [
        ("blas_interior_pops", C.c_uint64),
        ("blas_leaf_pops", C.c_uint64),
        ("tri_tests", C.c_uint64),
        ("blas_entries", C.c_uint64),
        ("tlas_pops", C.c_uint64),
        ("rays", C.c_uint64),
    ]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


class KptStageStats(C.Structure):
    _fields_ = [
        ("ms", C.c_double),
        ("launches", C.c_int32),
        ("reserved", C.c_int32),
        ("events", KptTraceCounters),
        ("paths", C.c_uint64),
        ("vertices", C.c_uint64),
    ]

    def as_dict(self):
        return {"ms": float(self.ms), "launches": int(self.launches), "events": self.events.as_dict(),
                "paths": int(self.paths), "vertices": int(self.vertices)}


class KptCamera(C.Structure):
    _fields_ = [
        ("inv_view", C.c_float * 16),
        ("inv_proj", C.c_float * 16),
        ("prev_inv_view", C.c_float * 16),
        ("prev_inv_proj", C.c_float * 16),
    ]


class KptSettings(C.Structure):
    _fields_ = [
        ("max_depth", C.c_int32),
        ("use_temporal_accumulation", C.c_int32),
        ("bloom_generate_bloom", C.c_int32),
        ("bloom_use_karis_average", C.c_int32),
        ("bloom_final_blend", C.c_float),
        ("bloom_internal_blend", C.c_float),
        ("enable_auto_exposure", C.c_int32),
        ("sun_angular_diameter_rad", C.c_float),
        ("sun_phi_rad", C.c_float),
        ("sun_theta_rad", C.c_float),
        ("sun_emission_factor", C.c_float),
        ("seed_mode", C.c_int32),
        ("integrator", C.c_int32),
        ("postprocess_enable_effects", C.c_int32),
        ("upscale_enable", C.c_int32),
        ("upscale_fsr", C.c_int32),
    ]


class KptExposure(C.Structure):
    _fields_ = [
        ("aperture_f_num", C.c_float),
        ("iso", C.c_float),
        ("shutter_speed_secs", C.c_float),
        ("max_iso", C.c_int32),
        ("min_iso", C.c_int32),
        ("min_shutter_secs", C.c_float),
        ("max_shutter_secs", C.c_float),
        ("ev_comp", C.c_float),
        ("white_point_ev", C.c_float),
        ("black_point_ev", C.c_float),
    ]

    @classmethod
    def camera(cls, aperture=2.8, iso=100, shutter=1.0 / 60.0, ev_comp=0.0, white=6.5, black=-10.0,
               max_iso=6400, min_iso=100, max_shutter=0.0333, min_shutter=0.00025):
        """ExposureValues(aperture, iso, shutter) with the reference's limits
        (pod_types.hpp:11-33) plus the setExposure extras."""
        return cls(aperture, float(iso), shutter, max_iso, min_iso, min_shutter, max_shutter, ev_comp, white, black)

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class KptFilmState(C.Structure):
    _fields_ = [
        ("luminance_exposure_scalar", C.c_float),
        ("white_point_ev", C.c_float),
        ("black_point_ev", C.c_float),
        ("scene_average_luminance", C.c_float),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class KptCacheSettings(C.Structure):
    _fields_ = [
        ("enabled", C.c_int32),
        ("train", C.c_int32),
        ("train_samples", C.c_int32),
        ("train_steps", C.c_int32),
        ("learning_rate", C.c_float),
        ("radiance_scale", C.c_float),
        ("seed", C.c_uint32),
        ("nee_at_vertex", C.c_int32),
        ("bounds_min", C.c_float * 3),
        ("bounds_max", C.c_float * 3),
        ("wide_roughness", C.c_float),
    ]


class KptView(C.Structure):
    _fields_ = [
        ("projection", C.c_float * 16),
        ("view", C.c_float * 16),
        ("width", C.c_int32),
        ("height", C.c_int32),
    ]


CACHE_PARAMS = 18624
# (out, in) of the cache layers; weights are stored per layer row-major [out][in]
CACHE_LAYERS = [(64, 32), (64, 64), (64, 64), (64, 64), (64, 64), (3, 64)]


# dtype matching kpt_ray for zero-copy numpy arrays of rays
RAY_DTYPE = np.dtype([("org", "<f4", 3), ("tmin", "<f4"), ("dir", "<f4", 3), ("tmax", "<f4")])
HIT_DTYPE = np.dtype([("t", "<f4"), ("prim_id", "<i4"), ("inst_id", "<i4"), ("bary", "<f4", 3)])
# reference BVH2Node / TLASNode layouts (blas.cuh:9-35, 209-238)
BVH2NODE_DTYPE = np.dtype([("first", "<i4"), ("count", "<u4"), ("pmin", "<f4", 3), ("pmax", "<f4", 3)])
TLASNODE_DTYPE = np.dtype([("pmin", "<f4", 3), ("pmax", "<f4", 3), ("lr", "<u4"), ("blas", "<i4")])
assert RAY_DTYPE.itemsize == C.sizeof(KptRay) == 32
assert HIT_DTYPE.itemsize == C.sizeof(KptHit) == 24
assert BVH2NODE_DTYPE.itemsize == 32 and TLASNODE_DTYPE.itemsize == 32

_P = C.c_void_p
_lib = None

# (name, restype, argtypes) of every entry point declared in include/kpt_abi.h
SIGNATURES = [
    ("kpt_get_abi_version", C.c_int, []),
    ("kpt_ctx_create", C.c_int, [C.c_int, C.POINTER(_P)]),
    ("kpt_ctx_destroy", C.c_int, [_P]),
    ("kpt_last_error", C.c_char_p, [_P]),
    ("kpt_upload_scene", C.c_int, [_P, C.POINTER(KptSceneDesc)]),
    ("kpt_set_instance_transform", C.c_int, [_P, C.c_int32, C.POINTER(C.c_float)]),
    ("kpt_set_instance_transforms", C.c_int, [_P, C.c_int32, C.c_int32, C.POINTER(C.c_float)]),
    ("kpt_set_material", C.c_int, [_P, C.c_int32, C.POINTER(KptMaterial)]),
    ("kpt_scene_info", C.c_int, [_P, C.POINTER(C.c_int64)]),
    ("kpt_get_blas", C.c_int, [_P, C.c_int32, _P, C.c_int64, C.POINTER(C.c_int32), C.c_int64, C.POINTER(C.c_int64)]),
    ("kpt_get_tlas", C.c_int, [_P, _P, C.c_int64, C.POINTER(C.c_int64)]),
    ("kpt_trace_closest", C.c_int, [_P, _P, _P, C.c_int64]),
    ("kpt_trace_any", C.c_int, [_P, _P, _P, C.c_int64]),
    ("kpt_trace_closest_counted", C.c_int, [_P, _P, _P, C.c_int64, C.POINTER(KptTraceCounters)]),
    ("kpt_trace_closest_counted_noalloc", C.c_int, [_P, _P, _P, C.c_int64, C.POINTER(KptTraceCounters)]),
    ("kpt_png_unfilter", C.c_int, [_P, C.c_int64, C.c_int64, C.c_int32, _P]),
    ("kpt_png_decode_rgb8", C.c_int, [_P, C.c_int64, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _P, C.c_int64]),
    ("kpt_write_png_rgba8", C.c_int, [C.c_char_p, _P, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    ("kpt_write_pfm", C.c_int, [C.c_char_p, _P, C.c_int32, C.c_int32, C.c_int32]),
    ("kpt_write_exr", C.c_int, [C.c_char_p, _P, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    ("kpt_measure_gather16", C.c_int, [_P, C.c_int64, _P, C.c_int64, _P]),
    ("kpt_set_settings", C.c_int, [_P, C.POINTER(KptSettings)]),
    ("kpt_get_settings", C.c_int, [_P, C.POINTER(KptSettings)]),
    ("kpt_set_view", C.c_int, [_P, C.POINTER(C.c_float), C.POINTER(C.c_float)]),
    ("kpt_set_camera", C.c_int, [_P, C.POINTER(KptCamera)]),
    ("kpt_get_camera", C.c_int, [_P, C.POINTER(KptCamera)]),
    ("kpt_resize", C.c_int, [_P, C.c_int32, C.c_int32]),
    ("kpt_reset_accumulation", C.c_int, [_P]),
    ("kpt_render", C.c_int, [_P, C.c_int32]),
    ("kpt_get_frame_index", C.c_int, [_P, C.POINTER(C.c_int32)]),
    ("kpt_readback", C.c_int, [_P, C.c_int32, _P, C.c_int64]),
    ("kpt_sync", C.c_int, [_P]),
    ("kpt_last_timing", C.c_int, [_P, C.POINTER(C.c_double)]),
    ("kpt_last_build_timing", C.c_int, [_P, C.POINTER(C.c_double)]),
    ("kpt_set_trace_counting", C.c_int, [_P, C.c_int32]),
    ("kpt_get_render_counters", C.c_int, [_P, C.POINTER(KptTraceCounters)]),
    ("kpt_get_heat_map", C.c_int, [_P, _P, _P, C.c_int64]),
    ("kpt_get_stage_stats", C.c_int, [_P, C.c_int32, C.POINTER(KptStageStats)]),
    ("kpt_get_resolution", C.c_int, [_P, C.POINTER(C.c_int32)]),
    ("kpt_set_tuning", C.c_int, [_P, C.POINTER(C.c_int32), C.c_int32]),
    ("kpt_get_tuning", C.c_int, [_P, C.POINTER(C.c_int32), C.c_int32]),
    ("kpt_set_exposure", C.c_int, [_P, C.POINTER(KptExposure)]),
    ("kpt_get_exposure", C.c_int, [_P, C.POINTER(KptExposure)]),
    ("kpt_get_film", C.c_int, [_P, C.POINTER(KptFilmState)]),
    ("kpt_set_frame_delta", C.c_int, [_P, C.c_float]),
    ("kpt_render_views", C.c_int, [_P, C.POINTER(KptView), C.c_int32, C.c_int32, _P]),
    ("kpt_render_views_host", C.c_int, [_P, C.POINTER(KptView), C.c_int32, C.c_int32, C.POINTER(C.c_float)]),
    ("kpt_cache_configure", C.c_int, [_P, C.POINTER(KptCacheSettings)]),
    ("kpt_cache_get_settings", C.c_int, [_P, C.POINTER(KptCacheSettings)]),
    ("kpt_cache_get_weights", C.c_int, [_P, C.POINTER(C.c_float), C.c_int64]),
    ("kpt_cache_set_weights", C.c_int, [_P, C.POINTER(C.c_float), C.c_int64]),
    ("kpt_cache_infer", C.c_int, [_P, _P, _P, C.c_int64, _P]),
    ("kpt_cache_train", C.c_int, [_P, _P, _P, _P, C.c_int64, C.c_int32, C.POINTER(C.c_float)]),
    ("kpt_cache_get_gradient", C.c_int, [_P, C.POINTER(C.c_float), C.c_int64]),
    ("kpt_cache_stats", C.c_int, [_P, C.POINTER(C.c_double)]),
    ("kpt_cache_get_queries", C.c_int, [_P, _P, _P, _P, C.c_int64, C.POINTER(C.c_int64)]),
]


class KptError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"kpt error {code}: {msg}")
        self.code = code


def lib():
    """Loads libkpt_hip.so (raises when it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with kittlespt_amd.build.build() "
                "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(ctx, rc):
    if rc != KPT_OK:
        msg = lib().kpt_last_error(ctx).decode() if ctx else ""
        raise KptError(rc, msg)
    return rc


def fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def iptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))
