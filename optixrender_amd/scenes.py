"""Seeded synthetic scenes for the benchmark configurations.

Every scene is built in float64 with numpy and stored as float32, so the GPU
path and the oracle receive bit-identical inputs.  Random draws use numpy's
PCG64 seeded per scene.

  cornell_spheres (256x256, 16 spp, depth 8, RR from bounce 3):
      Cornell box of 6 diffuse quads (grey 0.7 / red / green), one quad light of
      radiance 15, a GGX conductor icosphere (alpha 0.2) and a GGX dielectric
      icosphere (IOR 1.5), both subdivision 3 (1280 triangles each).
  sphere_field (1920x1080, 64 spp, depth 16): 12 icospheres at subdivision 6
      (81,920 triangles each, 983,040 total), centres U[-10,10]^3, radii
      U[0.5,2], materials uniform over conductor/dielectric/plastic with
      roughness U[0.05,0.8], lit by a procedural 2048x1024 HDR sky with a
      Gaussian sun of peak 5e4.
  heightfield_medium (1920x1080, 32 spp, depth 32): fBm heightfield with
      interleaved per-triangle materials and a 4x4x4 homogeneous-medium box
      (sigma_t 0.5, albedo 0.9, HG g 0.3) around a dielectric sphere.
  sphere_field_5m: 61 icospheres at subdivision 6 (4,997,120 triangles), a
      resident scene of the serving benchmark.
  lit_sphere_field / cornell_two_lights: sphere and quad lights in view of
      the camera, alone or beside the sky.
  textured_cornell: the Cornell box with every material texture slot in use
      (albedo, roughness, anisotropy, specular, metallic, both coat
      roughnesses, normal maps) on RGBA8 images of odd sizes, texture
      coordinates beyond [0, 1] on the floor (wrap addressing).
"""
from __future__ import annotations

import functools
import math

import numpy as np

from . import abi
from .renderer import SceneData

_PHI = (1.0 + math.sqrt(5.0)) / 2.0


def icosphere(subdiv: int) -> np.ndarray:
    """Unit icosphere as a triangle soup (20*4^s, 3, 3) float64, CCW outward
    (a fresh copy; the subdivision is computed once per level)."""
    return _icosphere(subdiv).copy()


@functools.lru_cache(maxsize=8)
def _icosphere(subdiv: int) -> np.ndarray:
    v = np.array([[-1, _PHI, 0], [1, _PHI, 0], [-1, -_PHI, 0], [1, -_PHI, 0], [0, -1, _PHI], [0, 1, _PHI],
                  [0, -1, -_PHI], [0, 1, -_PHI], [_PHI, 0, -1], [_PHI, 0, 1], [-_PHI, 0, -1], [-_PHI, 0, 1]],
                 dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]])
    tris = v[f]
    for _ in range(subdiv):
        a, b, c = tris[:, 0], tris[:, 1], tris[:, 2]

        def mid(p, q):
            m = p + q  # symmetric in (p, q): shared edges give identical vertices
            return m / np.linalg.norm(m, axis=1, keepdims=True)

        ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
        tris = np.stack([np.stack([a, ab, ca], 1), np.stack([ab, b, bc], 1), np.stack([ca, bc, c], 1),
                         np.stack([ab, bc, ca], 1)], 1).reshape(-1, 3, 3)
    return tris


def flat_normals(tris: np.ndarray) -> np.ndarray:
    """Geometric normals as TriangleMesh::LoadMesh adds them (Scene.cpp:150-158)."""
    e1 = tris[:, 1] - tris[:, 0]
    e2 = tris[:, 2] - tris[:, 0]
    e1 /= np.linalg.norm(e1, axis=1, keepdims=True)
    e2 /= np.linalg.norm(e2, axis=1, keepdims=True)
    n = np.cross(e1, e2)
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    return np.repeat(n[:, None, :], 3, axis=1)


def quad(p0, p1, p2, p3) -> np.ndarray:
    """Two triangles (p0,p1,p2), (p0,p2,p3)."""
    p = np.array([p0, p1, p2, p3], dtype=np.float64)
    return np.stack([p[[0, 1, 2]], p[[0, 2, 3]]])


def _soup(parts):
    pos = np.concatenate([p for p, _, _ in parts]).astype(np.float32)
    nrm = np.concatenate([n for _, n, _ in parts]).astype(np.float32)
    mid = np.concatenate([np.full(len(p), m, np.uint32) for p, _, m in parts])
    return pos, nrm, mid


def cornell_spheres(subdiv: int = 3) -> SceneData:
    """Cornell box + conductor and dielectric icospheres; 2,572 triangles."""
    x0, x1, y0, y1, z0, z1 = -1.0, 1.0, 0.0, 2.0, -1.0, 3.2
    walls = [
        (quad((x0, y0, z0), (x0, y0, z1), (x1, y0, z1), (x1, y0, z0)), 0),  # floor, normal +y
        (quad((x0, y1, z0), (x1, y1, z0), (x1, y1, z1), (x0, y1, z1)), 0),  # ceiling, -y
        (quad((x0, y0, z0), (x1, y0, z0), (x1, y1, z0), (x0, y1, z0)), 0),  # back wall, +z
        (quad((x0, y0, z1), (x0, y1, z1), (x1, y1, z1), (x1, y0, z1)), 0),  # front wall, -z
        (quad((x0, y0, z0), (x0, y1, z0), (x0, y1, z1), (x0, y0, z1)), 1),  # left (red), +x
        (quad((x1, y0, z0), (x1, y0, z1), (x1, y1, z1), (x1, y1, z0)), 2),  # right (green), -x
    ]
    parts = [(q, flat_normals(q), m) for q, m in walls]
    ico = icosphere(subdiv)
    for centre, radius, m in (((-0.45, 0.42, -0.2), 0.42, 3), ((0.45, 0.42, 0.4), 0.42, 4)):
        t = ico * radius + np.asarray(centre)
        parts.append((t, flat_normals(t), m))
    pos, nrm, mid = _soup(parts)
    mats = [
        abi.material(abi.DIFFUSE, albedo=(0.7, 0.7, 0.7), roughness=0.0),
        abi.material(abi.DIFFUSE, albedo=(0.7, 0.1, 0.1), roughness=0.0),
        abi.material(abi.DIFFUSE, albedo=(0.1, 0.7, 0.1), roughness=0.0),
        abi.material(abi.CONDUCTOR, roughness=math.sqrt(0.2)),
        abi.material(abi.DIELECTRIC, int_ior=1.5, ext_ior=1.0),
    ]
    # Emitter just below the ceiling, u x v = +y.  Double-sided: with a
    # one-sided quad every point above its plane gets 0/0 from light NEE
    # (Light.h:63-66 then devicePrograms.cu:469), which drops the sample.
    lights = [abi.light(0, double_side=1, radiance=(15.0, 15.0, 15.0), position=(-0.3, 1.98, 0.3),
                        u=(0.6, 0.0, 0.0), v=(0.0, 0.0, -0.6))]
    cam = abi.camera((0.0, 1.0, 3.0), (0.0, 1.0, 0.0))
    return SceneData(pos, nrm, mid, mats, lights, [], None, cam)


def quad_uvs(scale: float = 1.0) -> np.ndarray:
    """Texture coordinates of quad()'s two triangles: p0 (0,0), p1 (s,0),
    p2 (s,s), p3 (0,s)."""
    c = np.array([[0, 0], [scale, 0], [scale, scale], [0, scale]], dtype=np.float64)
    return np.stack([c[[0, 1, 2]], c[[0, 2, 3]]])


def default_uvs(n: int) -> np.ndarray:
    """(0,0), (1,0), (0,1) per triangle, as LoadMesh gives an untextured face
    (Scene.cpp:160-165)."""
    return np.broadcast_to(np.array([[0, 0], [1, 0], [0, 1]], dtype=np.float64), (n, 3, 2)).copy()


def sphere_uvs(unit: np.ndarray) -> np.ndarray:
    """Latitude-longitude coordinates of points on the unit sphere."""
    u = np.arctan2(unit[..., 2], unit[..., 0]) / (2.0 * math.pi) + 0.5
    v = np.arccos(np.clip(unit[..., 1], -1.0, 1.0)) / math.pi
    return np.stack([u, v], axis=-1)


def rgba8(rgb: np.ndarray) -> np.ndarray:
    """(h, w, 3) values in [0, 1] -> (h, w, 4) uint8 with alpha 255."""
    h, w = rgb.shape[:2]
    out = np.full((h, w, 4), 255, np.uint8)
    out[..., :3] = np.clip(np.rint(rgb * 255.0), 0, 255).astype(np.uint8)
    return out


def normal_map(height: np.ndarray, strength: float) -> np.ndarray:
    """Tangent-space normal map of a periodic height image, encoded n*0.5+0.5."""
    dx = (np.roll(height, -1, axis=1) - np.roll(height, 1, axis=1)) * 0.5
    dy = (np.roll(height, -1, axis=0) - np.roll(height, 1, axis=0)) * 0.5
    n = np.stack([-strength * dx, -strength * dy, np.ones_like(height)], axis=-1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    return rgba8(n * 0.5 + 0.5)


def textured_cornell(subdiv: int = 3, seed: int = 3, with_uvs: bool = True) -> SceneData:
    """The Cornell box with textured materials (Scene::AddMesh's eight texture
    slots, Scene.cpp:22-57; lookups at devicePrograms.cu:86-115):

      floor       plastic: albedo checker, specular and roughness images,
                  coordinates 0..3 (each image repeats 3x3 times)
      back wall   diffuse grey with a bumpy normal map
      right wall  clear-coated conductor: coat roughness u and v images
      left sphere conductor: roughness and anisotropy images
      right sphere metal workflow: albedo, metallic and normal images

    with_uvs=False leaves the coordinates to the default (0,0), (1,0), (0,1)
    per triangle."""
    rng = np.random.default_rng(seed)
    sc = cornell_spheres(subdiv)
    n = sc.n_tris
    uvs = default_uvs(n)
    uvs[0:2] = quad_uvs(3.0)     # floor
    uvs[4:6] = quad_uvs(1.0)     # back wall
    uvs[10:12] = quad_uvs(1.0)   # right wall
    ico = icosphere(subdiv)
    m = len(ico)
    unit = ico.copy()
    unit /= np.linalg.norm(unit, axis=-1, keepdims=True)
    uvs[12:12 + m] = sphere_uvs(unit)
    uvs[12 + m:12 + 2 * m] = sphere_uvs(unit)
    # images (odd sizes: the wrap is not a power-of-two mask)
    yy, xx = np.mgrid[0:40, 0:48]
    checker = ((xx // 6 + yy // 5) % 2).astype(np.float64)
    albedo = np.stack([0.15 + 0.7 * checker, 0.2 + 0.5 * checker, 0.6 - 0.4 * checker], axis=-1)
    t_albedo = rgba8(albedo)
    t_rough = rgba8(np.repeat(rng.uniform(0.1, 0.8, (29, 37, 1)), 3, axis=-1))
    t_spec = rgba8(rng.uniform(0.2, 1.0, (17, 23, 3)))
    hy, hx = np.mgrid[0:32, 0:32] * (2.0 * math.pi / 32.0)
    t_bump = normal_map(np.sin(3 * hx) * np.cos(2 * hy), 1.5)
    t_coat_u = rgba8(np.repeat(rng.uniform(0.05, 0.6, (13, 11, 1)), 3, axis=-1))
    t_coat_v = rgba8(np.repeat(rng.uniform(0.05, 0.6, (11, 13, 1)), 3, axis=-1))
    t_aniso = rgba8(np.repeat(rng.uniform(0.0, 0.8, (21, 19, 1)), 3, axis=-1))
    t_metal = rgba8(np.repeat(rng.uniform(0.0, 1.0, (9, 15, 1)), 3, axis=-1))
    t_sph_alb = rgba8(rng.uniform(0.3, 0.95, (24, 48, 3)))
    sy, sx = np.mgrid[0:24, 0:48] * (2.0 * math.pi / np.array([24.0, 48.0]))[:, None, None]
    t_sph_bump = normal_map(np.sin(5 * sx) * np.sin(2 * sy), 1.0)
    textures = [t_albedo, t_rough, t_spec, t_bump, t_coat_u, t_coat_v, t_aniso, t_metal, t_sph_alb, t_sph_bump]

    def tex(m, **slots):
        order = ("albedo", "roughness", "anisotropy", "specular", "metallic", "coat_u", "coat_v", "normal")
        for k, v in slots.items():
            m.tex[order.index(k)] = v
        return m

    mats = list(sc.materials)
    mats.append(tex(abi.material(abi.PLASTIC, roughness=0.3), albedo=0, roughness=1, specular=2))       # 5
    mats.append(tex(abi.material(abi.DIFFUSE, albedo=(0.7, 0.7, 0.7), roughness=0.0), normal=3))        # 6
    mats.append(tex(abi.material(abi.CLEARCOAT_CONDUCTOR, roughness=0.3), coat_u=4, coat_v=5))          # 7
    mats.append(tex(abi.material(abi.CONDUCTOR, roughness=0.3, anisotropy=0.5), roughness=1, anisotropy=6))  # 8
    mats.append(tex(abi.material(abi.METAL_WORKFLOW, roughness=0.25), albedo=8, metallic=7, normal=9))  # 9
    mid = sc.mat_ids.copy()
    mid[0:2] = 5
    mid[4:6] = 6
    mid[10:12] = 7
    mid[12:12 + m] = 8
    mid[12 + m:12 + 2 * m] = 9
    sc.mat_ids = mid
    sc.materials = mats
    sc.textures = textures
    sc.uvs = uvs.astype(np.float32) if with_uvs else None
    return sc


def foggy_cornell(phase_type: int = 1, g: float = 0.3) -> SceneData:
    """The Cornell scene filled with a homogeneous medium (camera, walls and
    light all inside medium 0): exercises free-flight sampling, medium NEE
    and the phase functions (core/cuda/Medium.h, devicePrograms.cu:276-338)."""
    sc = cornell_spheres()
    med = abi.medium(phase_type=phase_type, g=(g, g, g), sigma_s=(0.35, 0.30, 0.25),
                     sigma_a=(0.03, 0.05, 0.08), scale=1.0)
    for m in sc.materials:
        m.in_medium = 0
        m.out_medium = 0
    for lt in sc.lights:
        lt.medium = 0
    sc.media = [med]
    sc.camera = abi.camera((0.0, 1.0, 3.0), (0.0, 1.0, 0.0), medium=0)
    return sc


def sky_env(w: int = 2048, h: int = 1024, sun_peak: float = 5e4, sun_sigma_deg: float = 0.5,
            sun_dir=(0.35, 0.65, -0.45)) -> np.ndarray:
    """Procedural equirect sky: gradient plus a Gaussian sun, (h, w, 4) float32.

    Texel (i, j) is centred at uv = ((j+0.5)/w, (i+0.5)/h) and the direction
    follows PlaneToSphere (core/cuda/Sampling.h:12-23).  A fresh copy of a
    cached image."""
    return _sky_env(w, h, sun_peak, sun_sigma_deg, tuple(sun_dir)).copy()


@functools.lru_cache(maxsize=4)
def _sky_env(w, h, sun_peak, sun_sigma_deg, sun_dir) -> np.ndarray:
    u = (np.arange(w) + 0.5) / w
    v = (np.arange(h) + 0.5) / h
    phi = 2.0 * math.pi * (u - 0.5)
    theta = math.pi * (v - 0.5)
    ct, st = np.cos(theta)[:, None], np.sin(theta)[:, None]
    d = np.stack([ct * np.cos(phi)[None, :], np.broadcast_to(st, (h, w)), ct * np.sin(phi)[None, :]], -1)
    y = d[..., 1:2]
    horizon = np.array([1.0, 0.85, 0.7])
    zenith = np.array([0.15, 0.35, 0.9])
    ground = np.array([0.25, 0.22, 0.2])
    t = np.clip(y, 0.0, 1.0) ** 0.5
    sky = horizon * (1 - t) + zenith * t
    rgb = np.where(y >= 0, sky, ground * (1.0 - 0.5 * np.clip(-y, 0, 1)))
    s = np.asarray(sun_dir, np.float64)
    s /= np.linalg.norm(s)
    ang = np.arccos(np.clip((d * s).sum(-1), -1.0, 1.0))
    sig = math.radians(sun_sigma_deg)
    rgb = rgb + sun_peak * np.exp(-0.5 * (ang / sig) ** 2)[..., None] * np.array([1.0, 0.95, 0.85])
    out = np.ones((h, w, 4), np.float32)
    out[..., :3] = rgb
    return out


def sphere_field(seed: int = 1, subdiv: int = 6, count: int = 12, env_w: int = 2048, env_h: int = 1024) -> SceneData:
    """12 icospheres (983,040 triangles at subdivision 6) under a procedural HDR sky."""
    rng = np.random.default_rng(seed)
    ico = icosphere(subdiv)
    nrm_unit = ico / np.linalg.norm(ico, axis=2, keepdims=True)
    parts, mats = [], []
    for i in range(count):
        c = rng.uniform(-10.0, 10.0, 3)
        r = rng.uniform(0.5, 2.0)
        kind = int(rng.integers(0, 3))
        rough = float(rng.uniform(0.05, 0.8))
        col = rng.uniform(0.2, 0.9, 3)
        if kind == 0:
            mats.append(abi.material(abi.CONDUCTOR, roughness=rough))
        elif kind == 1:
            mats.append(abi.material(abi.DIELECTRIC, roughness=rough, int_ior=1.5, ext_ior=1.0))
        else:
            mats.append(abi.material(abi.PLASTIC, roughness=rough, albedo=col, int_ior=1.5, ext_ior=1.0))
        parts.append((ico * r + c, nrm_unit, i))
    pos, nrm, mid = _soup(parts)
    env = sky_env(env_w, env_h)
    cam = abi.camera((0.0, 0.0, 30.0), (0.0, 0.0, 0.0))
    return SceneData(pos, nrm, mid, mats, [], [], env, cam)


def cornell_all_materials(subdiv: int = 2) -> SceneData:
    """The Cornell box with eight small icospheres, one per material type
    (Material.h: Diffuse, Conductor, Dielectric, Plastic, MetalWorkflow,
    ThinDielectric, ClearCoatedConductor, DiffuseTransmitter), in two rows of
    four on the floor."""
    sc = cornell_spheres(subdiv)
    n_walls = 12
    pos = [sc.positions[:n_walls].astype(np.float64)]
    mid = [sc.mat_ids[:n_walls]]
    mats = list(sc.materials[:3])
    kinds = [
        abi.material(abi.DIFFUSE, albedo=(0.8, 0.6, 0.3), roughness=0.5),
        abi.material(abi.CONDUCTOR, roughness=0.35, anisotropy=0.3),
        abi.material(abi.DIELECTRIC, roughness=0.2, int_ior=1.5, ext_ior=1.0),
        abi.material(abi.PLASTIC, roughness=0.3, albedo=(0.2, 0.5, 0.8)),
        abi.material(abi.METAL_WORKFLOW, roughness=0.4, albedo=(0.9, 0.7, 0.4), metallic=0.6),
        abi.material(abi.THIN_DIELECTRIC, roughness=0.25, int_ior=1.4, ext_ior=1.0),
        abi.material(abi.CLEARCOAT_CONDUCTOR, roughness=0.4, coat_weight=0.8, coat_roughness_u=0.15,
                     coat_roughness_v=0.3),
        abi.material(abi.DIFFUSE_TRANSMITTER, albedo=(0.6, 0.7, 0.5)),
    ]
    ico = icosphere(subdiv)
    for i, m in enumerate(kinds):
        r = 0.2
        c = (-0.75 + 0.5 * (i % 4), r, -0.3 + 0.8 * (i // 4))
        pos.append(ico * r + np.asarray(c))
        mid.append(np.full(len(ico), len(mats), np.uint32))
        mats.append(m)
    p = np.concatenate(pos).astype(np.float32)
    sc.positions = p
    sc.normals = flat_normals(np.concatenate(pos)).astype(np.float32)
    sc.mat_ids = np.concatenate(mid)
    sc.materials = mats
    return sc


def lit_sphere_field(seed: int = 1, subdiv: int = 2, env_w: int = 256, env_h: int = 128,
                     with_env: bool = True) -> SceneData:
    """The sphere field with two analytic lights in the camera's view: a
    sphere light (radius 2, in the sky beside the field) and a one-sided 10 x
    10 quad light at height 10 facing down (Light.h:5-141); with_env=False removes the
    sky, so a camera ray that misses the geometry meets a light or nothing."""
    sc = sphere_field(seed=seed, subdiv=subdiv, env_w=env_w, env_h=env_h)
    sc.lights = [
        abi.light(1, radiance=(6.0, 5.0, 3.0), position=(14.0, 7.0, -5.0), radius=2.0),
        abi.light(0, double_side=0, radiance=(2.0, 2.5, 3.0), position=(-15.0, 10.0, 0.0), u=(10.0, 0.0, 0.0),
                  v=(0.0, 0.0, -10.0)),
    ]
    if not with_env:
        sc.env = None
    return sc


def cornell_two_lights(subdiv: int = 3) -> SceneData:
    """The Cornell scene with a small sphere light (radius 0.12) beside the
    ceiling quad light: light selection over two lights, cone sampling, and
    the reference's light_pdf overwrite in lightTrace."""
    sc = cornell_spheres(subdiv)
    sc.lights = list(sc.lights) + [abi.light(1, radiance=(8.0, 6.0, 4.0), position=(0.5, 1.5, 1.0), radius=0.12)]
    return sc


def fbm(n: int, octaves: int = 6, seed: int = 1, base: int = 4) -> np.ndarray:
    """(n, n) fractional Brownian motion of smoothstep value noise in about
    [-1, 1]: octave o has a (base 2^o + 1)^2 lattice and amplitude 2^-o."""
    rng = np.random.default_rng(seed)
    u = np.linspace(0.0, 1.0, n, dtype=np.float64)
    h = np.zeros((n, n), np.float32)
    amp, freq, norm = 1.0, base, 0.0
    for _ in range(octaves):
        lat = rng.uniform(-1.0, 1.0, (freq + 1, freq + 1)).astype(np.float32)
        x = u * freq
        i = np.minimum(x.astype(np.int64), freq - 1)
        f = (x - i).astype(np.float32)
        f = f * f * (3.0 - 2.0 * f)
        a = lat[i][:, i] * (1 - f)[None, :] + lat[i][:, i + 1] * f[None, :]
        b = lat[i + 1][:, i] * (1 - f)[None, :] + lat[i + 1][:, i + 1] * f[None, :]
        h += amp * (a * (1 - f)[:, None] + b * f[:, None])
        norm += amp
        amp *= 0.5
        freq *= 2
    return h / norm


def box(lo, hi) -> np.ndarray:
    """12 outward-facing triangles of an axis-aligned box."""
    (x0, y0, z0), (x1, y1, z1) = lo, hi
    faces = [
        quad((x0, y0, z0), (x1, y0, z0), (x1, y0, z1), (x0, y0, z1)),  # -y
        quad((x0, y1, z0), (x0, y1, z1), (x1, y1, z1), (x1, y1, z0)),  # +y
        quad((x0, y0, z0), (x0, y1, z0), (x1, y1, z0), (x1, y0, z0)),  # -z
        quad((x0, y0, z1), (x1, y0, z1), (x1, y1, z1), (x0, y1, z1)),  # +z
        quad((x0, y0, z0), (x0, y0, z1), (x0, y1, z1), (x0, y1, z0)),  # -x
        quad((x1, y0, z0), (x1, y1, z0), (x1, y1, z1), (x1, y0, z1)),  # +x
    ]
    return np.concatenate(faces)


def heightfield_medium(verts: int = 3163, octaves: int = 6, seed: int = 1, extent: float = 20.0,
                       height: float = 3.0, sphere_subdiv: int = 5, env_w: int = 2048,
                       env_h: int = 1024) -> SceneData:
    """fBm heightfield (2 (verts-1)^2 triangles: 19,996,488 at 3163 vertices a
    side) with four materials interleaved per triangle, a 4x4x4 box of
    homogeneous medium (sigma_s 0.45, sigma_a 0.05, HG g 0.3) resting on it,
    bounded by an index-matched thin dielectric (the reference has no
    pass-through interface; ThinDielectric at IOR 1/1 transmits straight on),
    and a GGX dielectric sphere (IOR 1.5) inside the box, under the HDR sky."""
    n = verts
    h = fbm(n, octaves, seed) * height
    g = np.linspace(-extent, extent, n, dtype=np.float32)
    X, Z = np.meshgrid(g, g)                       # X varies along columns, Z along rows
    P = np.stack([X, h, Z], -1).reshape(-1, 3)
    dx = 2.0 * extent / (n - 1)
    gz, gx = np.gradient(h, dx)
    N = np.stack([-gx, np.ones_like(h), -gz], -1)
    N /= np.linalg.norm(N, axis=-1, keepdims=True)
    N = N.reshape(-1, 3).astype(np.float32)
    r = np.arange(n - 1)
    v00 = (r[:, None] * n + r[None, :]).reshape(-1)
    v01, v10, v11 = v00 + 1, v00 + n, v00 + n + 1
    idx = np.stack([np.stack([v00, v10, v11], 1), np.stack([v00, v11, v01], 1)], 1).reshape(-1, 3)
    del v00, v01, v10, v11
    terrain = P[idx].astype(np.float32)
    tnorm = N[idx]
    del idx, P, N
    tmat = (np.arange(len(terrain)) % 4).astype(np.uint32)
    # medium box resting on the terrain at the centre, sphere inside it
    c = n // 2
    y0 = float(h[c, c]) + 0.25
    bx = box((-2.0, y0, -2.0), (2.0, y0 + 4.0, 2.0))
    sph = icosphere(sphere_subdiv) * 1.2 + np.array([0.0, y0 + 2.0, 0.0])
    pos = np.concatenate([terrain, bx.astype(np.float32), sph.astype(np.float32)])
    nrm = np.concatenate([tnorm, flat_normals(bx).astype(np.float32),
                          (icosphere(sphere_subdiv)).astype(np.float32)])
    mid = np.concatenate([tmat, np.full(len(bx), 4, np.uint32), np.full(len(sph), 5, np.uint32)])
    del terrain, tnorm
    mats = [
        abi.material(abi.DIFFUSE, albedo=(0.45, 0.40, 0.33), roughness=0.5),
        abi.material(abi.DIFFUSE, albedo=(0.20, 0.45, 0.15), roughness=0.3),
        abi.material(abi.PLASTIC, albedo=(0.55, 0.50, 0.45), roughness=0.4, int_ior=1.5, ext_ior=1.0),
        abi.material(abi.CONDUCTOR, roughness=0.35),
        abi.material(abi.THIN_DIELECTRIC, albedo=(1.0, 1.0, 1.0), roughness=0.05, int_ior=1.0, ext_ior=1.0),
        abi.material(abi.DIELECTRIC, roughness=0.1, int_ior=1.5, ext_ior=1.0),
    ]
    mats[4].in_medium, mats[4].out_medium = 0, -1
    mats[5].in_medium, mats[5].out_medium = -1, 0
    med = abi.medium(phase_type=1, g=(0.3, 0.3, 0.3), sigma_s=(0.45, 0.45, 0.45), sigma_a=(0.05, 0.05, 0.05),
                     scale=1.0)
    env = sky_env(env_w, env_h)
    cam = abi.camera((7.0, y0 + 5.5, 12.0), (0.0, y0 + 2.0, 0.0))
    return SceneData(pos, nrm, mid, mats, [], [med], env, cam)


def sphere_field_5m(seed: int = 7) -> SceneData:
    """Random-sphere field of 61 icospheres at subdivision 6: 4,997,120 triangles."""
    return sphere_field(seed=seed, subdiv=6, count=61)


CONFIGS = {
    "cornell_spheres": dict(width=256, height=256, spp=16, max_depth=8, rr_start=3),
    "sphere_field": dict(width=1920, height=1080, spp=64, max_depth=16, rr_start=0),
    "heightfield_medium": dict(width=1920, height=1080, spp=32, max_depth=32, rr_start=0),
}
