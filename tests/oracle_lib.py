"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so) for the tests.

The oracle is the checker; the product never loads it.  Its structs have the
same layout as the C ABI's (test_abi.py checks), so the product's ctypes
structures are passed straight through.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from optixrender_amd import abi
from optixrender_amd.renderer import SceneData

ROOT = Path(__file__).resolve().parent.parent
ORACLE = ROOT / "oracle" / "_build" / "liboracle.so"
REF_GDT = ROOT / "oracle" / "_ref" / "ref_gdt"

_P = C.c_void_p


class OrcScene(C.Structure):
    _fields_ = [
        ("n_tris", C.c_int64), ("positions", _P), ("normals", _P), ("mat_ids", _P),
        ("n_materials", C.c_int32), ("materials", _P), ("n_lights", C.c_int32), ("lights", _P),
        ("n_media", C.c_int32), ("media", _P), ("env_w", C.c_int32), ("env_h", C.c_int32), ("env_rgba", _P),
        ("uvs", _P), ("n_textures", C.c_int32), ("textures", _P),
    ]


class OrcRenderParams(C.Structure):
    _fields_ = [
        ("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32), ("max_depth", C.c_int32),
        ("rr_start", C.c_int32), ("filter", C.c_int32), ("frame_base", C.c_int32), ("threads", C.c_int32),
        ("fp64_accum", C.c_int32), ("row_step", C.c_int32),
    ]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not ORACLE.exists():
            raise RuntimeError(f"{ORACLE} missing: run `make -C oracle`")
        L = C.CDLL(str(ORACLE))
        L.orc_lcg_init.restype = C.c_uint32
        L.orc_lcg_init.argtypes = [C.c_uint32, C.c_uint32]
        L.orc_lcg_next.restype = C.c_float
        L.orc_lcg_next.argtypes = [C.POINTER(C.c_uint32)]
        L.orc_morton_codes.argtypes = [C.c_int64, _P, _P, _P]
        L.orc_lbvh_build.argtypes = [C.c_int64, _P, C.c_int32, _P, _P, _P, _P, _P, _P, _P, _P]
        L.orc_bvh_create.restype = _P
        L.orc_bvh_create.argtypes = [C.c_int64, _P]
        L.orc_bvh_destroy.argtypes = [_P]
        L.orc_bvh_closest.argtypes = [_P, C.c_int64, _P, _P, _P]
        L.orc_bvh_any.argtypes = [_P, C.c_int64, _P, _P, _P]
        L.orc_brute_closest.argtypes = [C.c_int64, _P, C.c_int64, _P, _P]
        L.orc_kulla_conty.argtypes = [_P, _P]
        L.orc_sample_ggx.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float, _P]
        L.orc_fresnel_conductor.argtypes = [_P, _P, _P, _P, _P]
        L.orc_fresnel_dielectric.argtypes = [_P, _P, C.c_float]
        L.orc_fresnel_dielectric.restype = C.c_float
        L.orc_average_fresnel_conductor.argtypes = [_P, _P, _P]
        L.orc_average_fresnel_dielectric.argtypes = [C.c_float]
        L.orc_average_fresnel_dielectric.restype = C.c_float
        L.orc_ggx_d.argtypes = [_P, C.c_float, C.c_float]
        L.orc_ggx_d.restype = C.c_float
        L.orc_smith_g1.argtypes = [_P, _P, C.c_float, C.c_float]
        L.orc_smith_g1.restype = C.c_float
        L.orc_material_eval.argtypes = [_P, _P, C.c_int, _P, _P, _P, _P]
        L.orc_material_sample.argtypes = [_P, _P, C.c_int, _P, _P, _P, _P, _P]
        L.orc_medium_sample_distance.argtypes = [_P, C.c_float, _P, _P]
        L.orc_medium_eval_distance.argtypes = [_P, C.c_int, C.c_float, _P]
        L.orc_phase_eval.argtypes = [_P, _P, _P, _P]
        L.orc_phase_sample.argtypes = [_P, _P, _P, _P]
        L.orc_filter_jitter.argtypes = [C.c_int, C.c_float, C.c_float, _P]
        L.orc_env_cache.argtypes = [C.c_int, C.c_int, _P, _P]
        L.orc_bilinear.argtypes = [C.c_int, C.c_int, _P, C.c_float, C.c_float, _P]
        L.orc_texture_sample.argtypes = [C.c_int, C.c_int, _P, C.c_float, C.c_float, _P]
        L.orc_normal_from_tangent.argtypes = [_P, _P, _P]
        L.orc_tonemap.argtypes = [C.c_int64, _P, _P]
        L.orc_render.argtypes = [_P, _P, _P, _P, _P, _P, _P, _P]
        L.orc_render.restype = C.c_int
        L.orc_render_stats.argtypes = [_P, _P, _P, _P]
        L.orc_render_stats.restype = C.c_int
        _lib = L
    return _lib


def p(a):
    return abi.ptr(a)


def lcg_init(v0, v1):
    return lib().orc_lcg_init(v0, v1)


def lcg_draws(state, n):
    s = C.c_uint32(state)
    out = []
    for _ in range(n):
        f = lib().orc_lcg_next(C.byref(s))
        out.append((f, s.value))
    return out


def morton(pos):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1)
    n = pos.size // 9
    codes = np.zeros(n, np.uint32)
    cb = np.zeros(6, np.float32)
    lib().orc_morton_codes(n, p(pos), p(codes), p(cb))
    return codes, cb


def lbvh(pos, rounds=-1):
    """The oracle's hierarchy over the positions: rounds = 0 the plain LBVH,
    < 0 the build default (the LBVH restructured by treelets, relabelled)."""
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1)
    n = pos.size // 9
    codes = np.zeros(n, np.uint32)
    order = np.zeros(n, np.uint32)
    left = np.zeros(max(n - 1, 1), np.uint32)
    right = np.zeros(max(n - 1, 1), np.uint32)
    parent = np.zeros(max(2 * n - 1, 1), np.uint32)
    lo = np.zeros((max(2 * n - 1, 1), 3), np.float32)
    hi = np.zeros((max(2 * n - 1, 1), 3), np.float32)
    qnode = np.zeros((max(n - 1, 1), 8), np.uint32)
    lib().orc_lbvh_build(n, p(pos), rounds, p(codes), p(order), p(left), p(right), p(parent), p(lo), p(hi), p(qnode))
    ni = max(n - 1, 0)
    return dict(codes=codes, order=order, left=left[:ni], right=right[:ni], parent=parent[: 2 * n - 1],
                lo=lo[: 2 * n - 1], hi=hi[: 2 * n - 1], qnode=qnode[:ni])


class Bvh:
    def __init__(self, pos):
        self.pos = np.ascontiguousarray(pos, np.float32).reshape(-1)
        self.n = self.pos.size // 9
        self.h = lib().orc_bvh_create(self.n, p(self.pos))

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_bvh_destroy(self.h)
            self.h = None

    def closest(self, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.zeros((rays.shape[0], 4), np.float32)
        counts = np.zeros((rays.shape[0], 2), np.uint32)
        lib().orc_bvh_closest(self.h, rays.shape[0], p(rays), p(hits), p(counts))
        return hits, counts

    def any(self, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        occ = np.zeros(rays.shape[0], np.uint8)
        counts = np.zeros((rays.shape[0], 2), np.uint32)
        lib().orc_bvh_any(self.h, rays.shape[0], p(rays), p(occ), p(counts))
        return occ.astype(bool), counts


def brute_closest(pos, rays):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1)
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    hits = np.zeros((rays.shape[0], 4), np.float32)
    lib().orc_brute_closest(pos.size // 9, p(pos), rays.shape[0], p(rays), p(hits))
    return hits


def kulla_conty():
    b = np.zeros(64 * 64, np.float32)
    a = np.zeros(64, np.float32)
    lib().orc_kulla_conty(p(b), p(a))
    return b.reshape(64, 64), a


def _v3(x):
    return np.ascontiguousarray(x, np.float32).reshape(3)


def sample_ggx(au, av, u0, u1):
    h = np.zeros(3, np.float32)
    lib().orc_sample_ggx(au, av, u0, u1, p(h))
    return h


def fresnel_conductor(V, H, eta, k):
    out = np.zeros(3, np.float32)
    V, H, eta, k = _v3(V), _v3(H), _v3(eta), _v3(k)
    lib().orc_fresnel_conductor(p(V), p(H), p(eta), p(k), p(out))
    return out


def fresnel_dielectric(V, H, eta_inv):
    V, H = _v3(V), _v3(H)
    return float(lib().orc_fresnel_dielectric(p(V), p(H), eta_inv))


def average_fresnel_conductor(eta, k):
    out = np.zeros(3, np.float32)
    eta, k = _v3(eta), _v3(k)
    lib().orc_average_fresnel_conductor(p(eta), p(k), p(out))
    return out


def average_fresnel_dielectric(eta):
    return float(lib().orc_average_fresnel_dielectric(eta))


def ggx_d(H, au, av):
    H = _v3(H)
    return float(lib().orc_ggx_d(p(H), au, av))


def smith_g1(V, H, au, av):
    V, H = _v3(V), _v3(H)
    return float(lib().orc_smith_g1(p(V), p(H), au, av))


def material_eval(m, n, front, V, L):
    f = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    n = np.asarray(n, np.float32); V = np.asarray(V, np.float32); L = np.asarray(L, np.float32)
    lib().orc_material_eval(C.byref(m), p(n), int(front), p(V), p(L), p(f), p(pdf))
    return f, float(pdf[0])


def material_sample(m, n, front, V, state):
    f = np.zeros(3, np.float32)
    pdf = np.zeros(1, np.float32)
    L = np.zeros(3, np.float32)
    st = C.c_uint32(state)
    n = np.asarray(n, np.float32); V = np.asarray(V, np.float32)
    lib().orc_material_sample(C.byref(m), p(n), int(front), p(V), C.byref(st), p(f), p(pdf), p(L))
    return f, float(pdf[0]), L, st.value


def phase_eval(md, V, L):
    out = np.zeros(4, np.float32)
    lib().orc_phase_eval(C.byref(md), p(np.asarray(V, np.float32)), p(np.asarray(L, np.float32)), p(out))
    return out[:3], float(out[3])


def phase_sample(md, V, state):
    out = np.zeros(7, np.float32)
    st = C.c_uint32(state)
    lib().orc_phase_sample(C.byref(md), p(np.asarray(V, np.float32)), C.byref(st), p(out))
    return out[:3], float(out[3]), out[4:7], st.value


def medium_sample_distance(md, maxd, state):
    out = np.zeros(6, np.float32)
    st = C.c_uint32(state)
    lib().orc_medium_sample_distance(C.byref(md), maxd, C.byref(st), p(out))
    return out, st.value


def medium_eval_distance(md, scattered, d):
    out = np.zeros(4, np.float32)
    lib().orc_medium_eval_distance(C.byref(md), int(scattered), d, p(out))
    return out


def filter_jitter(t, u0, u1):
    out = np.zeros(2, np.float32)
    lib().orc_filter_jitter(t, u0, u1, p(out))
    return out


def env_cache(env):
    env = np.ascontiguousarray(env, np.float32)
    h, w = env.shape[:2]
    out = np.zeros((h, w, 4), np.float32)
    lib().orc_env_cache(w, h, p(env), p(out))
    return out


def bilinear(env, u, v):
    env = np.ascontiguousarray(env, np.float32)
    h, w = env.shape[:2]
    out = np.zeros(4, np.float32)
    lib().orc_bilinear(w, h, p(env), u, v, p(out))
    return out


def texture_sample(img, u, v):
    img = np.ascontiguousarray(img, np.uint8)
    h, w = img.shape[:2]
    out = np.zeros(4, np.float32)
    lib().orc_texture_sample(w, h, p(img), u, v, p(out))
    return out


def normal_from_tangent(n, t):
    out = np.zeros(3, np.float32)
    lib().orc_normal_from_tangent(p(np.asarray(n, np.float32)), p(np.asarray(t, np.float32)), p(out))
    return out


def tonemap(rgba):
    rgba = np.ascontiguousarray(rgba, np.float32).reshape(-1, 4)
    out = np.zeros(rgba.shape[0], np.uint32)
    lib().orc_tonemap(rgba.shape[0], p(rgba), p(out))
    return out


def _scene_struct(scene: SceneData):
    pos, nrm, mid, mats, lights, media, env = scene.arrays()
    s = OrcScene()
    s.n_tris = scene.n_tris
    s.positions, s.normals, s.mat_ids = p(pos), p(nrm), p(mid)
    s.n_materials, s.materials = len(scene.materials), C.cast(mats, _P)
    s.n_lights, s.lights = len(scene.lights), C.cast(lights, _P)
    s.n_media, s.media = len(scene.media), C.cast(media, _P)
    if env is not None:
        s.env_h, s.env_w = env.shape[0], env.shape[1]
        s.env_rgba = p(env)
    uvs, tex, tkeep = scene.texture_arrays()
    s.uvs = None if uvs is None else p(uvs)
    if tex is not None:
        s.n_textures, s.textures = len(tex), C.cast(tex, _P)
    keep = (pos, nrm, mid, mats, lights, media, env, tkeep)
    return s, keep


def render(scene: SceneData, width, height, spp, max_depth=25, rr_start=0, filt=3, frame_base=0, threads=0,
           fp64=False, keep_paths=False, camera=None, color=None, albedo=None, normal=None, depth=None,
           with_depth=False):
    """(color, albedo, normal, paths), or (color, albedo, normal, depth, paths)
    with with_depth; the AOVs continue from the given buffers (progressive)."""
    s, keep = _scene_struct(scene)
    prm = OrcRenderParams(width, height, spp, max_depth, rr_start, filt, frame_base, threads, int(fp64), 0)
    color = np.zeros((height, width, 4), np.float32) if color is None else color.copy()
    albedo = np.zeros((height, width, 4), np.float32) if albedo is None else albedo.copy()
    normal = np.zeros((height, width, 4), np.float32) if normal is None else normal.copy()
    depth = np.zeros((height, width), np.float32) if depth is None else depth.copy()
    paths = np.zeros((spp, height, width, 3), np.float32) if keep_paths else None
    cam = camera or scene.camera
    rc = lib().orc_render(C.byref(s), C.byref(cam), C.byref(prm), p(color), p(albedo), p(normal), p(depth),
                          p(paths) if paths is not None else None)
    assert rc == 0
    if with_depth:
        return color, albedo, normal, depth, paths
    return color, albedo, normal, paths


def render_stats(scene: SceneData, width, height, spp, max_depth=25, rr_start=0, filt=3, frame_base=0, threads=0,
                 camera=None, row_step=1):
    """Ray and traversal counts of a render; with row_step > 1 only every
    row_step-th row is traced.  trace_s / paths: the time spent tracing (the
    oracle's LBVH build and environment cache excluded) and the paths traced."""
    s, keep = _scene_struct(scene)
    prm = OrcRenderParams(width, height, spp, max_depth, rr_start, filt, frame_base, threads, 0, row_step)
    out = np.zeros(8, np.uint64)
    cam = camera or scene.camera
    assert lib().orc_render_stats(C.byref(s), C.byref(cam), C.byref(prm), p(out)) == 0
    return dict(extension_rays=int(out[0]), shadow_rays=int(out[1]), node_fetches=int(out[2]), tri_tests=int(out[3]),
                ext_node_fetches=int(out[4]), ext_tri_tests=int(out[5]), trace_s=float(out[6]) * 1e-9,
                paths=int(out[7]))
