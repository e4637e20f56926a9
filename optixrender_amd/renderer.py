"""Python mirror of the reference's host interface, driven through the C ABI.

``Scene`` and ``Renderer`` keep the reference's method names and meaning
(core/include/Scene.h:195-214, core/include/Renderer.h:10-25):
``Renderer(scene)``, ``Resize``, ``SetCamera``, ``Render``,
``DownloadPixels`` and the public fields ``progressive`` / ``denoiserOn``
and ``launchParams.numPixelSamples``.  Errors surface as exceptions, as the
reference's CUDA_CHECK does (common/optix7.h:25-35).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import abi


@dataclass
class SceneData:
    """De-indexed triangle soup plus scene tables (one material slot per triangle)."""

    positions: np.ndarray            # (n, 3, 3) float32
    normals: np.ndarray              # (n, 3, 3) float32
    mat_ids: np.ndarray              # (n,) uint32
    materials: list = field(default_factory=list)
    lights: list = field(default_factory=list)
    media: list = field(default_factory=list)
    env: np.ndarray | None = None    # (h, w, 4) float32, row 0 at v ~ 0
    camera: abi.Camera | None = None
    uvs: np.ndarray | None = None    # (n, 3, 2) float32; None: (0,0), (1,0), (0,1) per triangle
    textures: list = field(default_factory=list)  # (h, w, 4) uint8 images, row 0 at v ~ 0

    @property
    def n_tris(self) -> int:
        return int(self.positions.shape[0])

    def bounds(self):
        p = self.positions.reshape(-1, 3)
        return p.min(axis=0), p.max(axis=0)

    def arrays(self):
        """Contiguous typed arrays in the layout the C ABI expects."""
        pos = np.ascontiguousarray(self.positions, dtype=np.float32).reshape(-1)
        nrm = np.ascontiguousarray(self.normals, dtype=np.float32).reshape(-1)
        mid = np.ascontiguousarray(self.mat_ids, dtype=np.uint32).reshape(-1)
        mats = (abi.Material * max(len(self.materials), 1))(*self.materials)
        lights = (abi.Light * max(len(self.lights), 1))(*self.lights)
        media = (abi.Medium * max(len(self.media), 1))(*self.media)
        env = None if self.env is None else np.ascontiguousarray(self.env, dtype=np.float32)
        return pos, nrm, mid, mats, lights, media, env

    def texture_arrays(self):
        """(uvs or None, a Texture array or None, arrays to keep alive)."""
        uvs = None if self.uvs is None else np.ascontiguousarray(self.uvs, dtype=np.float32).reshape(-1)
        if not self.textures:
            return uvs, None, (uvs,)
        imgs = [np.ascontiguousarray(t, dtype=np.uint8) for t in self.textures]
        for im in imgs:
            if im.ndim != 3 or im.shape[2] != 4:
                raise ValueError("a texture is an (h, w, 4) uint8 array")
        tex = (abi.Texture * len(imgs))(*[abi.Texture(im.shape[1], im.shape[0], abi.ptr(im)) for im in imgs])
        return uvs, tex, (uvs, imgs, tex)


class Context:
    """One device context (oxr_ctx)."""

    def __init__(self, device: int = 0):
        self._lib = abi.lib()
        h = C.c_void_p()
        rc = self._lib.oxr_create(device, C.byref(h))
        if rc != 0:
            raise abi.OxrError(f"oxr_create({device}) failed with status {rc}")
        self.h = h

    def check(self, rc: int, what: str):
        if rc != 0:
            msg = self._lib.oxr_last_error(self.h)
            raise abi.OxrError(f"{what}: status {rc}: {msg.decode() if msg else ''}")

    def close(self):
        if self.h:
            self._lib.oxr_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- scenes
    def upload(self, scene: SceneData) -> int:
        pos, nrm, mid, mats, lights, media, env = scene.arrays()
        uvs, tex, _keep = scene.texture_arrays()
        d = abi.SceneDesc()
        d.n_tris = scene.n_tris
        d.positions = abi.ptr(pos)
        d.normals = abi.ptr(nrm)
        d.uvs = None if uvs is None else abi.ptr(uvs)
        if tex is not None:
            d.n_textures = len(tex)
            d.textures = C.cast(tex, C.c_void_p)
        d.mat_ids = abi.ptr(mid)
        d.n_materials = len(scene.materials)
        d.materials = C.cast(mats, C.c_void_p)
        d.n_lights = len(scene.lights)
        d.lights = C.cast(lights, C.c_void_p)
        d.n_media = len(scene.media)
        d.media = C.cast(media, C.c_void_p)
        if env is not None:
            d.env_h, d.env_w = int(env.shape[0]), int(env.shape[1])
            d.env_rgba = abi.ptr(env)
        sid = C.c_uint32()
        self.check(self._lib.oxr_scene_upload(self.h, C.byref(d), C.byref(sid)), "oxr_scene_upload")
        return sid.value

    def build_bvh(self, sid: int, rounds: int = -1) -> float:
        """Device BVH build; rounds of treelet restructuring after the LBVH
        (0: the plain LBVH, -1: the library default)."""
        ms = C.c_float()
        self.check(self._lib.oxr_scene_build_bvh_rounds(self.h, sid, rounds, C.byref(ms)), "oxr_scene_build_bvh")
        return ms.value

    def select(self, sid: int):
        self.check(self._lib.oxr_scene_select(self.h, sid), "oxr_scene_select")

    def release(self, sid: int):
        self.check(self._lib.oxr_scene_release(self.h, sid), "oxr_scene_release")

    def download_bvh(self, sid: int, n: int):
        codes = np.zeros(n, np.uint32)
        order = np.zeros(n, np.uint32)
        ni = max(n - 1, 0)
        left = np.zeros(max(ni, 1), np.uint32)
        right = np.zeros(max(ni, 1), np.uint32)
        lo = np.zeros((max(2 * n - 1, 1), 3), np.float32)
        hi = np.zeros((max(2 * n - 1, 1), 3), np.float32)
        self.check(self._lib.oxr_scene_download_bvh(self.h, sid, abi.ptr(codes), abi.ptr(order), abi.ptr(left),
                                                    abi.ptr(right), abi.ptr(lo), abi.ptr(hi)), "download_bvh")
        return codes, order, left[:ni], right[:ni], lo[: 2 * n - 1], hi[: 2 * n - 1]

    def download_nodes(self, sid: int, n: int) -> np.ndarray:
        """The traversal's compressed node records: (n-1, 8) words."""
        out = np.zeros((max(n - 1, 1), 8), np.uint32)
        self.check(self._lib.oxr_scene_download_nodes(self.h, sid, abi.ptr(out)), "download_nodes")
        return out[: max(n - 1, 0)]

    def download_env_cache(self, sid: int, w: int, h: int) -> np.ndarray:
        out = np.zeros((h, w, 4), np.float32)
        self.check(self._lib.oxr_scene_download_env_cache(self.h, sid, abi.ptr(out)), "download_env_cache")
        return out

    def trace(self, rays: np.ndarray, any_hit: bool = False):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        hits = np.zeros((n, 4), np.float32)
        occ = np.zeros(n, np.uint8)
        counts = np.zeros((n, 2), np.uint32)
        self.check(self._lib.oxr_trace_rays(self.h, abi.ptr(rays), n, int(any_hit), abi.ptr(hits), abi.ptr(occ),
                                            abi.ptr(counts)), "oxr_trace_rays")
        return (occ.astype(bool) if any_hit else hits), counts

    # -- renderer
    def resize(self, w: int, h: int):
        self.check(self._lib.oxr_resize(self.h, w, h), "oxr_resize")

    def set_camera(self, cam: abi.Camera):
        self.check(self._lib.oxr_set_camera(self.h, C.byref(cam)), "oxr_set_camera")

    def render(self, spp=1, max_depth=25, rr_start=0, progressive=False, filter=abi.FILTER_GAUSSIAN, flags=0,
               batch_frames=0, denoise=False):
        p = abi.RenderParams(spp, max_depth, rr_start, int(progressive), filter, int(denoise), flags, batch_frames)
        self.check(self._lib.oxr_render(self.h, C.byref(p)), "oxr_render")
        return self.stats()

    def stats(self) -> abi.RenderStats:
        s = abi.RenderStats()
        self.check(self._lib.oxr_get_stats(self.h, C.byref(s)), "oxr_get_stats")
        return s

    def get_tuning(self) -> abi.Tuning:
        t = abi.Tuning()
        self.check(self._lib.oxr_get_tuning(self.h, C.byref(t)), "oxr_get_tuning")
        return t

    def set_tuning(self, t: abi.Tuning):
        self.check(self._lib.oxr_set_tuning(self.h, C.byref(t)), "oxr_set_tuning")

    def autotune(self, spp=4, max_depth=16, rr_start=0):
        p = abi.RenderParams(spp, max_depth, rr_start, 0, abi.FILTER_GAUSSIAN, 0, 0, 0)
        t = abi.Tuning()
        ms = C.c_double()
        self.check(self._lib.oxr_autotune(self.h, C.byref(p), C.byref(t), C.byref(ms)), "oxr_autotune")
        return t, ms.value

    def present(self, copy: bool = True):
        """(pixels (h, w) uint32, frame_id) of the newest render made with
        OXR_FLAG_PRESENT.  copy=False returns a view of the context's
        page-locked buffer, valid through the next render with the flag."""
        px = C.POINTER(C.c_uint32)()
        w, h, fid = C.c_int32(), C.c_int32(), C.c_int32()
        self.check(self._lib.oxr_present(self.h, C.byref(px), C.byref(w), C.byref(h), C.byref(fid)), "oxr_present")
        view = np.ctypeslib.as_array(px, shape=(h.value, w.value))
        return (view.copy() if copy else view), fid.value

    def download_rgba8(self, w: int, h: int) -> np.ndarray:
        out = np.zeros(w * h, np.uint32)
        self.check(self._lib.oxr_download_rgba8(self.h, abi.ptr(out), out.size), "oxr_download_rgba8")
        return out.reshape(h, w)

    def load_denoiser(self, flat_weights: np.ndarray):
        """oxr_denoiser_load: per layer w[co][ci][3][3] then b[co] (denoise.flat_weights)."""
        w = np.ascontiguousarray(flat_weights, np.float32)
        self.check(self._lib.oxr_denoiser_load(self.h, abi.ptr(w), w.size), "oxr_denoiser_load")

    def denoise(self, color, albedo, normal, depth) -> np.ndarray:
        """oxr_denoise_host on one image: (h, w, 4) AOVs and (h, w) depth -> (h, w, 4)."""
        h, w = depth.shape
        arrs = [np.ascontiguousarray(a, np.float32) for a in (color, albedo, normal, depth)]
        out = np.zeros((h, w, 4), np.float32)
        self.check(self._lib.oxr_denoise_host(self.h, w, h, *[abi.ptr(a) for a in arrs], abi.ptr(out)),
                   "oxr_denoise_host")
        return out

    def download_depth(self, w: int, h: int) -> np.ndarray:
        out = np.zeros((h, w), np.float32)
        self.check(self._lib.oxr_download_depth(self.h, abi.ptr(out), w * h), "oxr_download_depth")
        return out

    def download_aov(self, which: int, w: int, h: int) -> np.ndarray:
        out = np.zeros((h, w, 4), np.float32)
        self.check(self._lib.oxr_download_aov(self.h, which, abi.ptr(out), w * h), "oxr_download_aov")
        return out

    def download_path_radiance(self, spp: int, w: int, h: int) -> np.ndarray:
        out = np.zeros((spp, h, w, 3), np.float32)
        self.check(self._lib.oxr_download_path_radiance(self.h, abi.ptr(out), out.size), "download_path_radiance")
        return out

    def tonemap(self, rgba: np.ndarray) -> np.ndarray:
        rgba = np.ascontiguousarray(rgba, np.float32).reshape(-1, 4)
        out = np.zeros(rgba.shape[0], np.uint32)
        self.check(self._lib.oxr_tonemap_host(self.h, abi.ptr(rgba), rgba.shape[0], abi.ptr(out)), "oxr_tonemap_host")
        return out

    def material_batch(self, m: abi.Material, mode: int, inputs: np.ndarray) -> np.ndarray:
        inputs = np.ascontiguousarray(inputs, np.float32).reshape(-1, 12)
        out = np.zeros((inputs.shape[0], 8), np.float32)
        self.check(self._lib.oxr_material_batch(self.h, C.byref(m), mode, abi.ptr(inputs), inputs.shape[0],
                                                abi.ptr(out)), "oxr_material_batch")
        return out


class _FrameParams:
    def __init__(self):
        self.frameID = 0
        self.size = (0, 0)


class _LaunchParams:
    """Fields of LaunchParams (core/include/LaunchParams.h:28-69) the viewer touches."""

    def __init__(self):
        self.numPixelSamples = 1
        self.maxBounce = 25
        self.filterType = abi.FILTER_GAUSSIAN
        self.frame = _FrameParams()


class Renderer:
    """Renderer(scene) -> Resize / SetCamera / Render / DownloadPixels (Renderer.h:13-25)."""

    def __init__(self, scene: SceneData, device: int = 0, ctx: Context | None = None):
        self.ctx = ctx or Context(device)
        self.scene = scene
        self.scene_id = self.ctx.upload(scene)
        self.build_ms = self.ctx.build_bvh(self.scene_id)
        self.ctx.select(self.scene_id)
        self.denoiserOn = False
        self.progressive = False
        self.rrStart = 0
        self.launchParams = _LaunchParams()
        self._camera = scene.camera

    def Resize(self, size):
        w, h = int(size[0]), int(size[1])
        self.ctx.resize(w, h)
        self.launchParams.frame.size = (w, h)
        if self._camera is not None:
            self.SetCamera(self._camera)

    def SetCamera(self, cam: abi.Camera):
        self._camera = cam
        self.ctx.set_camera(cam)
        self.launchParams.frame.frameID = 0

    def Render(self):
        if self.launchParams.frame.size[0] == 0:  # Renderer.cpp:638-640
            return None
        st = self.ctx.render(spp=self.launchParams.numPixelSamples, max_depth=self.launchParams.maxBounce,
                             rr_start=self.rrStart, progressive=self.progressive,
                             filter=self.launchParams.filterType)
        self.launchParams.frame.frameID = st.frame_id
        return st

    def DownloadPixels(self, h_pixels: np.ndarray | None = None) -> np.ndarray:
        w, h = self.launchParams.frame.size
        px = self.ctx.download_rgba8(w, h).reshape(-1)
        if h_pixels is not None:
            h_pixels[:] = px
            return h_pixels
        return px


class Server:
    """Continuous-batching render server over a context's resident scenes (oxr_server_*)."""

    def __init__(self, ctx: Context, budget_bytes: int, max_active: int = 16, tick_paths: int = 1 << 24,
                 max_frame_pixels: int = 1920 * 1080, runners: int = 0):
        self.ctx = ctx
        self._lib = ctx._lib
        cfg = abi.ServerConfig(budget_bytes, max_active, runners, tick_paths, max_frame_pixels)
        h = C.c_void_p()
        ctx.check(self._lib.oxr_server_create(ctx.h, C.byref(cfg), C.byref(h)), "oxr_server_create")
        self.h = h
        self.sizes = {}

    def check(self, rc, what):
        if rc != 0:
            msg = self._lib.oxr_server_last_error(self.h)
            raise abi.OxrError(f"{what}: status {rc}: {msg.decode() if msg else ''}")

    def close(self):
        if self.h:
            self._lib.oxr_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, scene_id, width, height, spp, camera, max_depth=16, rr_start=0, filter=abi.FILTER_GAUSSIAN,
               denoise=False):
        r = abi.Request(scene_id, width, height, spp, max_depth, rr_start, filter, int(denoise), camera)
        rid = C.c_uint64()
        self.check(self._lib.oxr_submit(self.h, C.byref(r), C.byref(rid)), "oxr_submit")
        self.sizes[rid.value] = (width, height)
        return rid.value

    def poll(self, rid):
        st = C.c_int()
        self.check(self._lib.oxr_poll(self.h, rid, C.byref(st)), "oxr_poll")
        return st.value

    def wait(self, rid, timeout_ms=-1):
        return self._lib.oxr_wait(self.h, rid, timeout_ms)

    def cancel(self, rid):
        self.check(self._lib.oxr_cancel(self.h, rid), "oxr_cancel")

    def result_color(self, rid):
        w, h = self.sizes[rid]
        out = np.zeros((h, w, 4), np.float32)
        self.check(self._lib.oxr_result_color(self.h, rid, abi.ptr(out), w * h), "oxr_result_color")
        return out

    def result_rgba8(self, rid):
        w, h = self.sizes[rid]
        out = np.zeros((h, w), np.uint32)
        self.check(self._lib.oxr_result_rgba8(self.h, rid, abi.ptr(out), w * h), "oxr_result_rgba8")
        return out

    def release(self, rid):
        self.check(self._lib.oxr_release_request(self.h, rid), "oxr_release_request")
        self.sizes.pop(rid, None)

    def times(self, rid):
        """(queued ms, total ms) of a finished request."""
        q, t = C.c_double(), C.c_double()
        self.check(self._lib.oxr_request_times(self.h, rid, C.byref(q), C.byref(t)), "oxr_request_times")
        return q.value, t.value

    def stats(self) -> abi.ServerStats:
        s = abi.ServerStats()
        self.check(self._lib.oxr_server_get_stats(self.h, C.byref(s)), "oxr_server_get_stats")
        return s
