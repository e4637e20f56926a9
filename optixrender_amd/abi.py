"""ctypes binding of the C ABI declared in include/oxr.h.

The product path is the HIP library ``liboxr.so`` built in-tree; importing
:func:`lib` raises if it is missing -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "liboxr.so"

OXR_FLAG_KEEP_PATH_RADIANCE = 1
OXR_FLAG_COUNT_TRAVERSAL = 2
OXR_FLAG_EAGER = 4
OXR_FLAG_STAGE_TIMING = 8
OXR_FLAG_NO_CAMERA_CULL = 16
OXR_FLAG_PRESENT = 32

# MaterialType (core/include/Scene.h:59-68)
DIFFUSE, CONDUCTOR, DIELECTRIC, PLASTIC, METAL_WORKFLOW, THIN_DIELECTRIC, CLEARCOAT_CONDUCTOR, DIFFUSE_TRANSMITTER = range(8)
# FilterType (Scene.h:24-29)
FILTER_BOX, FILTER_TENT, FILTER_TRIANGLE, FILTER_GAUSSIAN = range(4)


class Material(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("albedo", C.c_float * 3),
        ("roughness", C.c_float),
        ("anisotropy", C.c_float),
        ("eta", C.c_float * 3),
        ("k", C.c_float * 3),
        ("int_ior", C.c_float),
        ("ext_ior", C.c_float),
        ("nonlinear", C.c_int32),
        ("specular", C.c_float * 3),
        ("metallic", C.c_float),
        ("coat_weight", C.c_float),
        ("coat_roughness_u", C.c_float),
        ("coat_roughness_v", C.c_float),
        ("in_medium", C.c_int32),
        ("out_medium", C.c_int32),
        ("tex", C.c_int32 * 8),
    ]


def material(type=METAL_WORKFLOW, **kw) -> Material:
    """A Material with the reference's defaults (core/include/Scene.h:70-116)."""
    m = Material()
    m.type = type
    m.albedo[:] = (1.0, 1.0, 1.0)
    m.roughness = 0.1
    m.anisotropy = 0.0
    m.eta[:] = (0.14282, 0.37414, 1.43944)
    m.k[:] = (3.97472, 2.38066, 1.59981)
    m.int_ior = 1.5
    m.ext_ior = 1.0
    m.nonlinear = 1
    m.specular[:] = (1.0, 1.0, 1.0)
    m.metallic = 0.8
    m.coat_weight = 2.0
    m.coat_roughness_u = 0.1
    m.coat_roughness_v = 0.1
    m.in_medium = -1
    m.out_medium = -1
    m.tex[:] = [-1] * 8
    for key, val in kw.items():
        if not hasattr(m, key):
            raise AttributeError(key)
        if isinstance(val, (tuple, list, np.ndarray)):
            getattr(m, key)[:] = [float(x) for x in val]
        else:
            setattr(m, key, val)
    return m


class Light(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("double_side", C.c_int32),
        ("radiance", C.c_float * 3),
        ("position", C.c_float * 3),
        ("u", C.c_float * 3),
        ("v", C.c_float * 3),
        ("radius", C.c_float),
        ("medium", C.c_int32),
    ]


def light(type=1, **kw) -> Light:
    """A Light with the reference's defaults (Scene.h:143-155)."""
    lt = Light()
    lt.type = type
    lt.double_side = 0
    lt.radiance[:] = (3.0, 3.0, 3.0)
    lt.position[:] = (6.0, 2.0, 0.0)
    lt.u[:] = (0.0, 0.0, 6.0)
    lt.v[:] = (6.0, 0.0, 0.0)
    lt.radius = 3.0
    lt.medium = -1
    for key, val in kw.items():
        if isinstance(val, (tuple, list, np.ndarray)):
            getattr(lt, key)[:] = [float(x) for x in val]
        else:
            setattr(lt, key, val)
    return lt


class Medium(C.Structure):
    _fields_ = [
        ("type", C.c_int32),
        ("phase_type", C.c_int32),
        ("g", C.c_float * 3),
        ("sigma_s", C.c_float * 3),
        ("sigma_a", C.c_float * 3),
        ("scale", C.c_float),
    ]


def medium(**kw) -> Medium:
    """A Medium with the reference's defaults (Scene.h:33-55)."""
    m = Medium()
    m.type = 0
    m.phase_type = 0
    m.g[:] = (0.0, 0.0, 0.0)
    m.sigma_s[:] = (0.74, 0.88, 1.01)
    m.sigma_a[:] = (0.032, 0.17, 0.48)
    m.scale = 500.0
    for key, val in kw.items():
        if isinstance(val, (tuple, list, np.ndarray)):
            getattr(m, key)[:] = [float(x) for x in val]
        else:
            setattr(m, key, val)
    return m


class Camera(C.Structure):
    _fields_ = [
        ("from_", C.c_float * 3),
        ("at", C.c_float * 3),
        ("up", C.c_float * 3),
        ("medium", C.c_int32),
        ("cos_fovy", C.c_float),
    ]


def camera(frm, at, up=(0.0, 1.0, 0.0), medium=-1, cos_fovy=0.66) -> Camera:
    c = Camera()
    c.from_[:] = [float(x) for x in frm]
    c.at[:] = [float(x) for x in at]
    c.up[:] = [float(x) for x in up]
    c.medium = medium
    c.cos_fovy = cos_fovy
    return c


class Texture(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("rgba8", C.c_void_p)]


class SceneDesc(C.Structure):
    _fields_ = [
        ("n_tris", C.c_int64),
        ("positions", C.c_void_p),
        ("normals", C.c_void_p),
        ("uvs", C.c_void_p),
        ("mat_ids", C.c_void_p),
        ("n_materials", C.c_int32),
        ("materials", C.c_void_p),
        ("n_lights", C.c_int32),
        ("lights", C.c_void_p),
        ("n_media", C.c_int32),
        ("media", C.c_void_p),
        ("env_w", C.c_int32),
        ("env_h", C.c_int32),
        ("env_rgba", C.c_void_p),
        ("n_textures", C.c_int32),
        ("textures", C.c_void_p),
    ]


class RenderParams(C.Structure):
    _fields_ = [
        ("spp", C.c_int32),
        ("max_depth", C.c_int32),
        ("rr_start", C.c_int32),
        ("progressive", C.c_int32),
        ("filter", C.c_int32),
        ("denoise", C.c_int32),
        ("flags", C.c_uint32),
        ("batch_frames", C.c_int32),
    ]


class RenderStats(C.Structure):
    _fields_ = [
        ("render_ms", C.c_double),
        ("trace_ms", C.c_double),
        ("shade_ms", C.c_double),
        ("generate_ms", C.c_double),
        ("accumulate_ms", C.c_double),
        ("tonemap_ms", C.c_double),
        ("denoise_ms", C.c_double),
        ("extension_rays", C.c_uint64),
        ("shadow_rays", C.c_uint64),
        ("node_fetches", C.c_uint64),
        ("tri_tests", C.c_uint64),
        ("ext_node_fetches", C.c_uint64),
        ("ext_tri_tests", C.c_uint64),
        ("paths", C.c_uint64),
        ("frame_id", C.c_int32),
        ("kernel_launches", C.c_int32),
        ("iterations", C.c_int32),
        ("windows", C.c_int32),
        ("camera_rays_resolved", C.c_uint64),
        ("resolved_node_fetches", C.c_uint64),
        ("resolved_tri_tests", C.c_uint64),
        ("tail_extension_rays", C.c_uint64),
        ("tail_shadow_rays", C.c_uint64),
        ("tail_node_fetches", C.c_uint64),
        ("tail_tri_tests", C.c_uint64),
        ("trace_lane_visits", C.c_uint64),
        ("trace_wave_slots", C.c_uint64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class ServerConfig(C.Structure):
    _fields_ = [
        ("hbm_budget_bytes", C.c_size_t),
        ("max_active", C.c_int32),
        ("runners", C.c_int32),
        ("tick_paths", C.c_int64),
        ("max_frame_pixels", C.c_int64),
    ]


class Request(C.Structure):
    _fields_ = [
        ("scene_id", C.c_uint32),
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("spp", C.c_int32),
        ("max_depth", C.c_int32),
        ("rr_start", C.c_int32),
        ("filter", C.c_int32),
        ("denoise", C.c_int32),
        ("camera", Camera),
    ]


class ServerStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64), ("completed", C.c_uint64), ("cancelled", C.c_uint64), ("rejected", C.c_uint64),
        ("ticks", C.c_uint64), ("paths", C.c_uint64),
        ("busy_ms", C.c_double), ("wall_ms", C.c_double),
        ("p50_ms", C.c_double), ("p99_ms", C.c_double), ("mean_ms", C.c_double),
        ("active", C.c_int32), ("pending", C.c_int32),
        ("bytes_in_use", C.c_size_t), ("budget_bytes", C.c_size_t),
        ("overlapped_ticks", C.c_uint64), ("runners", C.c_int32), ("max_running", C.c_int32),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


class Tuning(C.Structure):
    _fields_ = [
        ("trace_blocks_per_cu", C.c_int32),
        ("shade_blocks_per_cu", C.c_int32),
        ("stack_lds", C.c_int32),
        ("iters_per_poll", C.c_int32),
        ("pool_paths", C.c_int64),
        ("window_samples", C.c_int64),
        ("tail_paths", C.c_int64),
    ]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


REQ_PENDING, REQ_ACTIVE, REQ_DONE, REQ_CANCELLED, REQ_FAILED = range(5)

_P = C.c_void_p
_I64 = C.c_int64
# name -> (restype, argtypes); every symbol include/oxr.h declares.
SIGNATURES = {
    "oxr_create": (C.c_int, [C.c_int, C.POINTER(C.c_void_p)]),
    "oxr_destroy": (C.c_int, [_P]),
    "oxr_last_error": (C.c_char_p, [_P]),
    "oxr_malloc": (C.c_int, [_P, C.c_size_t, C.POINTER(C.c_void_p)]),
    "oxr_trace_rays_noalloc": (C.c_int, [_P, _P, _I64, C.c_int, _P, _P]),
    "oxr_free": (C.c_int, [_P, _P]),
    "oxr_memcpy_h2d": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "oxr_memcpy_d2h": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "oxr_memcpy_d2d": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "oxr_synchronize": (C.c_int, [_P]),
    "oxr_scene_upload": (C.c_int, [_P, _P, C.POINTER(C.c_uint32)]),
    "oxr_scene_build_bvh": (C.c_int, [_P, C.c_uint32, C.POINTER(C.c_float)]),
    "oxr_scene_build_bvh_rounds": (C.c_int, [_P, C.c_uint32, C.c_int32, C.POINTER(C.c_float)]),
    "oxr_scene_release": (C.c_int, [_P, C.c_uint32]),
    "oxr_scene_select": (C.c_int, [_P, C.c_uint32]),
    "oxr_scene_download_bvh": (C.c_int, [_P, C.c_uint32, _P, _P, _P, _P, _P, _P]),
    "oxr_scene_download_nodes": (C.c_int, [_P, C.c_uint32, _P]),
    "oxr_scene_download_env_cache": (C.c_int, [_P, C.c_uint32, _P]),
    "oxr_trace_rays": (C.c_int, [_P, _P, _I64, C.c_int, _P, _P, _P]),
    "oxr_resize": (C.c_int, [_P, C.c_int, C.c_int]),
    "oxr_set_camera": (C.c_int, [_P, _P]),
    "oxr_render": (C.c_int, [_P, _P]),
    "oxr_get_stats": (C.c_int, [_P, _P]),
    "oxr_download_rgba8": (C.c_int, [_P, _P, C.c_size_t]),
    "oxr_present": (C.c_int, [_P, C.POINTER(C.POINTER(C.c_uint32)), C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                              C.POINTER(C.c_int32)]),
    "oxr_download_aov": (C.c_int, [_P, C.c_int, _P, C.c_size_t]),
    "oxr_download_depth": (C.c_int, [_P, _P, C.c_size_t]),
    "oxr_denoiser_load": (C.c_int, [_P, _P, C.c_int64]),
    "oxr_denoise_host": (C.c_int, [_P, C.c_int, C.c_int, _P, _P, _P, _P, _P]),
    "oxr_download_path_radiance": (C.c_int, [_P, _P, C.c_size_t]),
    "oxr_server_create": (C.c_int, [_P, _P, C.POINTER(C.c_void_p)]),
    "oxr_server_destroy": (C.c_int, [_P]),
    "oxr_server_last_error": (C.c_char_p, [_P]),
    "oxr_submit": (C.c_int, [_P, _P, C.POINTER(C.c_uint64)]),
    "oxr_poll": (C.c_int, [_P, C.c_uint64, C.POINTER(C.c_int)]),
    "oxr_wait": (C.c_int, [_P, C.c_uint64, C.c_int]),
    "oxr_cancel": (C.c_int, [_P, C.c_uint64]),
    "oxr_result_rgba8": (C.c_int, [_P, C.c_uint64, _P, C.c_size_t]),
    "oxr_result_color": (C.c_int, [_P, C.c_uint64, _P, C.c_size_t]),
    "oxr_release_request": (C.c_int, [_P, C.c_uint64]),
    "oxr_server_get_stats": (C.c_int, [_P, _P]),
    "oxr_request_times": (C.c_int, [_P, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "oxr_get_tuning": (C.c_int, [_P, _P]),
    "oxr_set_tuning": (C.c_int, [_P, _P]),
    "oxr_autotune": (C.c_int, [_P, _P, _P, C.POINTER(C.c_double)]),
    "oxr_tonemap_host": (C.c_int, [_P, _P, _I64, _P]),
    "oxr_kulla_conty": (C.c_int, [_P, _P]),
    "oxr_material_batch": (C.c_int, [_P, _P, C.c_int, _P, _I64, _P]),
}

_lib = None


def lib() -> C.CDLL:
    """Load liboxr.so; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: run `python -m optixrender_amd.build` (no CPU fallback)")
        from . import build as _build

        if not _build.STAMP.exists() or _build.STAMP.read_text().strip() != _build.source_hash():
            raise RuntimeError(f"{LIB_PATH} was built from other sources than the ones in the tree: rebuild")
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


class OxrError(RuntimeError):
    pass


def ptr(a: np.ndarray | None):
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data_as(C.c_void_p)
