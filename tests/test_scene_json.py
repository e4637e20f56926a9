"""The whole scene file (the models/test.json schema: camera, lights, shapes
with transforms, materials with constant or image parameters, media) read by
the C++ facade's SceneParser.  The headless driver writes what it parsed
(OXR_HEADLESS_DUMP) without touching a device, so these tests run on the CPU.

Pins: transformed vertices against rotation matrices built here, normals
against the transformed edges (perpendicular: the inverse-transpose rule),
BMP/PPM texels and row order, names resolved to indices, the smooth_* and
texture-type rules, the environment's scale, and the errors a bad file
raises."""
import json
import math
import os
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADLESS = ROOT / "optixrender_amd" / "headless"


def write_obj(path, verts, faces, uvs=None):
    lines = [f"v {x} {y} {z}" for x, y, z in verts]
    if uvs is not None:
        lines += [f"vt {u} {v}" for u, v in uvs]
        lines += ["f " + " ".join(f"{i}/{i}" for i in f) for f in faces]
    else:
        lines += ["f " + " ".join(str(i) for i in f) for f in faces]
    path.write_text("\n".join(lines) + "\n")


def write_bmp(path, rgb, top_down=False):
    h, w = rgb.shape[:2]
    stride = (w * 3 + 3) & ~3
    rows = []
    order = range(h) if top_down else range(h - 1, -1, -1)   # bottom-up files store the last image row first
    for y in order:
        row = rgb[y][:, ::-1].tobytes()                          # BGR
        rows.append(row + b"\0" * (stride - 3 * w))
    data = b"".join(rows)
    hh = -h if top_down else h
    header = b"BM" + struct.pack("<IHHI", 54 + len(data), 0, 0, 54)
    dib = struct.pack("<IiiHHIIiiII", 40, w, hh, 1, 24, 0, len(data), 2835, 2835, 0, 0)
    path.write_bytes(header + dib + data)


def write_ppm(path, rgb):
    h, w = rgb.shape[:2]
    path.write_bytes(f"P6\n{w} {h}\n255\n".encode() + rgb.astype(np.uint8).tobytes())


def write_pfm(path, rgb):
    h, w = rgb.shape[:2]
    path.write_bytes(f"PF\n{w} {h}\n-1.0\n".encode() + rgb.astype("<f4").tobytes())


def rot(deg):
    rx, ry, rz = np.radians(deg)
    X = np.array([[1, 0, 0], [0, math.cos(rx), -math.sin(rx)], [0, math.sin(rx), math.cos(rx)]])
    Y = np.array([[math.cos(ry), 0, math.sin(ry)], [0, 1, 0], [-math.sin(ry), 0, math.cos(ry)]])
    Z = np.array([[math.cos(rz), -math.sin(rz), 0], [math.sin(rz), math.cos(rz), 0], [0, 0, 1]])
    return Z @ Y @ X   # x first, then y, then z


def dump(tmp_path, scene):
    (tmp_path / "scene.json").write_text(json.dumps(scene))
    out = tmp_path / "dump.json"
    env = dict(os.environ, OXR_HEADLESS_DUMP=str(out))
    r = subprocess.run([str(HEADLESS), str(tmp_path / "scene.json"), str(tmp_path / "x.ppm"), "1"],
                       capture_output=True, text=True, timeout=60, env=env)
    return r, (json.loads(out.read_text()) if r.returncode == 0 else None)


def build_files(tmp_path):
    """A small scene in the layout models/test.json uses: model/, texture/, hdr/."""
    (tmp_path / "model").mkdir()
    (tmp_path / "texture").mkdir()
    (tmp_path / "hdr").mkdir()
    # a quad with texture coordinates, and a tetrahedron without them or normals
    write_obj(tmp_path / "model" / "quad.obj", [(-1, 0, -1), (1, 0, -1), (1, 0, 1), (-1, 0, 1)],
              [(1, 2, 3), (1, 3, 4)], uvs=[(0, 0), (1, 0), (1, 1), (0, 1)])
    write_obj(tmp_path / "model" / "tet.obj", [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)],
              [(1, 3, 2), (1, 2, 4), (1, 4, 3), (2, 3, 4)])
    rng = np.random.default_rng(5)
    albedo = rng.integers(0, 256, (3, 5, 3), dtype=np.uint8)
    rough = rng.integers(0, 256, (4, 2, 3), dtype=np.uint8)
    write_bmp(tmp_path / "texture" / "albedo.bmp", albedo)
    write_bmp(tmp_path / "texture" / "rough_td.bmp", rough, top_down=True)
    write_ppm(tmp_path / "texture" / "normal.ppm", np.full((2, 2, 3), (128, 128, 255), np.uint8))
    sky = rng.uniform(0.0, 2.0, (4, 8, 3)).astype(np.float32)
    write_pfm(tmp_path / "hdr" / "sky.pfm", sky)
    return albedo, rough, sky


def base_scene():
    return {
        "width": 48, "height": 32, "depth": 5, "filter": {"type": "tent"},
        "camera": {"look_from": [4, 3, 5], "up": [0, 1, 0], "medium": "fog"},
        "light": [
            {"type": "infinite", "hdr_path": "hdr/sky.pfm", "scale": 0.5},
            {"type": "sphere", "position": [0, 3, 0], "radius": 0.25, "radiance": [5, 4, 3], "medium": "fog"},
            {"type": "quad", "position": [-1, 2, 1], "u": [2, 0, 0], "v": [0, 0, -2], "radiance": 2,
             "double_side": True},
        ],
        "shape": [
            {"type": "mesh", "mesh_path": "model/quad.obj", "translate": [0, -1, 0], "rotate": [0, 0, 0],
             "scale": [3, 1, 3], "material": "floor"},
            {"type": "mesh", "mesh_path": "model/tet.obj", "translate": [0.5, 0.25, -0.5],
             "rotate": [30, 45, 60], "scale": [1, 2, 0.5], "material": "glass", "medium_int": "milk",
             "medium_ext": "fog"},
            {"type": "mesh", "mesh_path": "model/quad.obj", "material": "metal"},
        ],
        "material": [
            {"type": "smooth_diffuse", "name": "floor", "albedo_texture_type": "image",
             "albedo_path": "texture/albedo.bmp", "roughness_texture_type": "image",
             "roughness_path": "texture/rough_td.bmp"},
            {"type": "smooth_dielectric", "name": "glass", "albedo_texture_type": "constant", "albedo": [1, 1, 1],
             "int_ior": 1.33, "ext_ior": 1.0},
            {"type": "metal_workflow", "name": "metal", "albedo_texture_type": "image",
             "albedo_path": "texture/albedo.bmp", "metallic_texture_type": "constant", "metallic": 0.25,
             "roughness": 0.4, "normal_texture_type": "image", "normal_path": "texture/normal.ppm"},
        ],
        "medium": [
            {"type": "homogeneous", "name": "fog", "phase": "henyey_greenstein", "g": [0.3, 0.3, 0.3],
             "sigma_a": [0.1, 0.2, 0.3], "sigma_s": [0.4, 0.5, 0.6], "scale": 0.5},
            {"type": "homogeneous", "name": "milk", "phase": "isotropic", "sigma_a": [0.70, 1.22, 1.90],
             "sigma_s": [0.0014, 0.0025, 0.0142], "scale": 0.5},
        ],
    }


def test_headless_driver_built():
    assert HEADLESS.exists(), "run `python -m optixrender_amd.build`"


def test_whole_scene_is_read(tmp_path):
    albedo, rough, sky = build_files(tmp_path)
    r, d = dump(tmp_path, base_scene())
    assert r.returncode == 0, r.stdout + r.stderr
    assert (d["width"], d["height"], d["depth"], d["filter"]) == (48, 32, 5, 1)   # Box Tent Triangle Gaussian
    floor, tet, metal = d["meshes"]
    # --- transforms: p' = t + Rz Ry Rx (s * p)
    R = rot([30, 45, 60])
    src = np.array([(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1)], float)
    faces = np.array([(1, 3, 2), (1, 2, 4), (1, 4, 3), (2, 3, 4)]) - 1
    want = (src[faces.reshape(-1)] * [1, 2, 0.5]) @ R.T + [0.5, 0.25, -0.5]
    got = np.array(tet["vertex"])
    assert np.allclose(got, want, atol=1e-5)
    # normals (flat, from the untransformed OBJ) stay perpendicular to the
    # transformed edges and keep their side: the inverse-transpose rule
    nrm = np.array(tet["normal"]).reshape(-1, 3, 3)
    tri = got.reshape(-1, 3, 3)
    for t in range(4):
        e1, e2 = tri[t, 1] - tri[t, 0], tri[t, 2] - tri[t, 0]
        n = nrm[t, 0]
        assert abs(np.linalg.norm(n) - 1) < 1e-5
        assert abs(n @ e1) < 1e-5 and abs(n @ e2) < 1e-5
        assert n @ np.cross(e1, e2) > 0
    assert np.allclose(np.array(floor["vertex"])[:3], [(-3, -1, -3), (3, -1, -3), (3, -1, 3)])
    assert np.allclose(np.array(floor["texcoord"])[:3], [(0, 0), (1, 0), (1, 1)])
    assert np.allclose(np.array(tet["texcoord"])[:3], [(0, 0), (1, 0), (0, 1)])   # no vt: Scene.cpp:160-165
    # --- materials: types, smooth rules, constants, one image per file
    fm, gm, mm = floor["material"], tet["material"], metal["material"]
    assert (fm["type"], fm["roughness"]) == (0, 0.0)          # smooth_diffuse: Lambert
    assert gm["type"] == 2 and gm["roughness"] == pytest.approx(0.02)
    assert gm["int_ior"] == pytest.approx(1.33)
    assert mm["type"] == 4 and mm["metallic"] == 0.25 and mm["roughness"] == pytest.approx(0.4)
    assert fm["tex"][0] == mm["tex"][0] == 0                   # albedo.bmp shared
    assert fm["tex"][1] == 1 and mm["tex"][7] == 2 and gm["tex"] == [-1] * 8
    # --- textures: BMP rows bottom first like stbi's flipped load, BGR -> RGB
    t0, t1, t2 = d["textures"]
    a = np.array(t0["rgba"], np.uint8).reshape(3, 5, 4)
    assert np.array_equal(a[..., :3], albedo[::-1]) and np.all(a[..., 3] == 255)
    b = np.array(t1["rgba"], np.uint8).reshape(4, 2, 4)
    assert np.array_equal(b[..., :3], rough[::-1])             # top-down file, same result
    assert (t2["width"], t2["height"]) == (2, 2)
    # --- media and names
    fog, milk = d["media"]
    assert fog["phase"] == 1 and np.allclose(fog["g"], 0.3) and np.allclose(fog["sigma_s"], [0.4, 0.5, 0.6])
    assert milk["phase"] == 0 and milk["scale"] == 0.5
    assert (tet["in_medium"], tet["out_medium"]) == (1, 0)
    assert (floor["in_medium"], floor["out_medium"]) == (-1, -1)
    # --- lights and environment
    sph, quad = d["lights"]
    assert sph["type"] == 1 and sph["radius"] == 0.25 and sph["medium"] == 0
    assert quad["type"] == 0 and quad["double_side"] == 1 and np.allclose(quad["radiance"], 2)
    assert (d["env"]["width"], d["env"]["height"]) == (8, 4)
    assert d["env"]["rgb_sum"] == pytest.approx(0.5 * float(sky.astype(np.float64).sum()), rel=1e-6)
    # --- camera: look_at defaults to the centre of the shapes' bounds
    cam = d["camera"]
    allv = np.concatenate([np.array(m["vertex"]) for m in d["meshes"]])
    centre = (allv.min(0) + allv.max(0)) / 2
    assert np.allclose(cam["from"], [4, 3, 5]) and np.allclose(cam["at"], centre, atol=1e-5)
    assert cam["medium"] == 0


@pytest.mark.parametrize("edit,msg", [
    (lambda s: s["shape"][0].update(material="nope"), "unknown material"),
    (lambda s: s["material"][0].update(type="velvet"), "unknown type"),
    (lambda s: s["shape"][1].update(medium_int="soup"), "unknown medium"),
    (lambda s: s["material"][0].update(albedo_path="texture/missing.bmp"), "Could not read texture"),
    (lambda s: s["light"].append({"type": "spot"}), "unknown light type"),
    (lambda s: s["shape"][0].update(scale=[1, 0, 1]), "zero scale"),
])
def test_bad_scene_files_are_reported(tmp_path, edit, msg):
    build_files(tmp_path)
    sc = base_scene()
    edit(sc)
    r, _ = dump(tmp_path, sc)
    assert r.returncode == 1 and msg in r.stdout, r.stdout


def scene_from_dump(d, env_rgb):
    """The dumped scene as the Python mirror's SceneData (one material slot
    per mesh, as the facade's Renderer uploads it)."""
    from optixrender_amd import abi
    from optixrender_amd.renderer import SceneData

    pos, nrm, uvs, mids, mats = [], [], [], [], []
    for i, m in enumerate(d["meshes"]):
        v = np.array(m["vertex"], np.float32).reshape(-1, 3, 3)
        pos.append(v)
        nrm.append(np.array(m["normal"], np.float32).reshape(-1, 3, 3))
        uvs.append(np.array(m["texcoord"], np.float32).reshape(-1, 3, 2))
        mids.append(np.full(len(v), i, np.uint32))
        a = m["material"]
        mat = abi.material(a["type"], albedo=a["albedo"], roughness=a["roughness"], anisotropy=a["anisotropy"],
                           eta=a["eta"], k=a["k"], int_ior=a["int_ior"], ext_ior=a["ext_ior"],
                           nonlinear=a["nonlinear"], specular=a["specular"], metallic=a["metallic"],
                           coat_weight=a["coat_weight"], coat_roughness_u=a["coat_roughness_u"],
                           coat_roughness_v=a["coat_roughness_v"], in_medium=m["in_medium"],
                           out_medium=m["out_medium"])
        mat.tex[:] = a["tex"]
        mats.append(mat)
    lights = [abi.light(l["type"], double_side=l["double_side"], radiance=l["radiance"], position=l["position"],
                        u=l["u"], v=l["v"], radius=l["radius"], medium=l["medium"]) for l in d["lights"]]
    media = [abi.medium(phase_type=m["phase"], g=m["g"], sigma_a=m["sigma_a"], sigma_s=m["sigma_s"], scale=m["scale"])
             for m in d["media"]]
    env = np.ones(env_rgb.shape[:2] + (4,), np.float32)
    env[..., :3] = env_rgb
    c = d["camera"]
    cam = abi.camera(c["from"], c["at"], c["up"], medium=c["medium"])
    tex = [np.array(t["rgba"], np.uint8).reshape(t["height"], t["width"], 4) for t in d["textures"]]
    return SceneData(np.concatenate(pos), np.concatenate(nrm), np.concatenate(mids), mats, lights, media, env, cam,
                     np.concatenate(uvs), tex)


@pytest.mark.gpu
def test_scene_file_renders_like_the_python_mirror(tmp_path, ctx):
    # the facade's upload of a parsed scene (materials, textures, texture
    # coordinates, lights, media, environment, camera medium) against the
    # Python mirror's upload of what the parser produced: the same RGBA8
    _, _, sky = build_files(tmp_path)
    r, d = dump(tmp_path, base_scene())
    assert r.returncode == 0, r.stdout
    frames = 3
    out = tmp_path / "img.ppm"
    r = subprocess.run([str(HEADLESS), str(tmp_path / "scene.json"), str(out), str(frames)],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = out.read_bytes().split(b"\n", 3)[3]
    img = np.frombuffer(raw, np.uint8).reshape(32, 48, 3)[::-1]   # the PPM is written top row first
    sc = scene_from_dump(d, (sky * np.float32(0.5)).astype(np.float32))
    sid = ctx.upload(sc)
    try:
        ctx.build_bvh(sid)
        ctx.select(sid)
        ctx.resize(48, 32)
        ctx.set_camera(sc.camera)
        ctx.render(spp=frames, max_depth=25, rr_start=0, filter=d["filter"])
        px = ctx.download_rgba8(48, 32)
    finally:
        ctx.release(sid)
    rgb = np.stack([px & 255, (px >> 8) & 255, (px >> 16) & 255], -1).astype(np.uint8)
    print(f"[scene-json] mean {img.mean():.1f}, differing px {np.count_nonzero((rgb != img).any(-1))}", flush=True)
    assert img.mean() > 5
    assert np.array_equal(rgb, img)


# ---- BSON scene files (SceneParser.h:38-39: a name ending in "bson" is read
# with nlohmann::json::from_bson) ----

def bson_encode(doc, int64=False):
    """A BSON 1.1 encoder (bsonspec.org) for the JSON subset scene files use."""
    def elem(key, v):
        k = key.encode() + b"\0"
        if isinstance(v, bool):
            return b"\x08" + k + (b"\x01" if v else b"\x00")
        if isinstance(v, int):
            if not int64 and -2 ** 31 <= v < 2 ** 31:
                return b"\x10" + k + struct.pack("<i", v)
            return b"\x12" + k + struct.pack("<q", v)
        if isinstance(v, float):
            return b"\x01" + k + struct.pack("<d", v)
        if isinstance(v, str):
            s = v.encode() + b"\0"
            return b"\x02" + k + struct.pack("<i", len(s)) + s
        if v is None:
            return b"\x0a" + k
        if isinstance(v, dict):
            return b"\x03" + k + document(v.items())
        if isinstance(v, (list, tuple)):
            return b"\x04" + k + document((str(i), x) for i, x in enumerate(v))
        raise TypeError(type(v))

    def document(items):
        body = b"".join(elem(k, v) for k, v in items)
        return struct.pack("<i", len(body) + 5) + body + b"\0"

    return document(doc.items())


def test_bson_encoder_matches_the_specification_examples():
    # the two example documents of bsonspec.org's FAQ, byte for byte
    assert bson_encode({"hello": "world"}) == b"\x16\x00\x00\x00\x02hello\x00\x06\x00\x00\x00world\x00\x00"
    assert bson_encode({"BSON": ["awesome", 5.05, 1986]}) == (
        b"\x31\x00\x00\x00\x04BSON\x00\x26\x00\x00\x00\x020\x00\x08\x00\x00\x00awesome\x00"
        b"\x011\x00\x33\x33\x33\x33\x33\x33\x14\x40\x102\x00\xc2\x07\x00\x00\x00\x00")


def dump_bson(tmp_path, data, name="scene.bson"):
    (tmp_path / name).write_bytes(data)
    out = tmp_path / "dump.json"
    if out.exists():
        out.unlink()
    env = dict(os.environ, OXR_HEADLESS_DUMP=str(out))
    r = subprocess.run([str(HEADLESS), str(tmp_path / name), str(tmp_path / "x.ppm"), "1"],
                       capture_output=True, text=True, timeout=60, env=env)
    return r, (json.loads(out.read_text()) if r.returncode == 0 else None)


@pytest.mark.parametrize("int64", [False, True])
def test_bson_scene_reads_like_its_json(tmp_path, int64):
    build_files(tmp_path)
    sc = base_scene()
    sc["comment"] = None          # null, and an element nothing reads
    r, want = dump(tmp_path, sc)
    assert r.returncode == 0, r.stdout + r.stderr
    r, got = dump_bson(tmp_path, bson_encode(sc, int64=int64))
    assert r.returncode == 0, r.stdout + r.stderr
    assert got == want


@pytest.mark.parametrize("mutate,msg", [
    (lambda b: b[:-1], "bad document size"),                                # truncated file
    (lambda b: b + b"\0", "bytes after the document"),
    (lambda b: struct.pack("<i", len(b) + 4) + b[4:], "bad document size"),
    (lambda b: b[:4] + b"\x05" + b[5:], "unsupported element type 0x05"),   # first element as binary
    (lambda b: b[:-8] + struct.pack("<i", 100) + b[-4:], "unexpected end"),  # string longer than the file
])
def test_bad_bson_is_reported(tmp_path, mutate, msg):
    build_files(tmp_path)
    data = bson_encode({"width": 48, "height": 32, "s": "xy"})   # ends with the string element
    r, _ = dump_bson(tmp_path, mutate(data))
    assert r.returncode == 1 and "bson: " + msg in r.stdout, r.stdout + r.stderr
