"""The C++ facade inside the reference's own source tree (CPU only).

The facade replaces core/include/{Renderer,Scene,SceneParser}.h and
core/src/*.cpp; main.cpp and the viewer (common/glfWindow) stay as they are.
Built with the reference's include paths (CMakeLists.txt:13-24, 37, 64-68),
the facade takes gdt's vector types and the reference's vendored stb_image,
so that:

  * main.cpp compiles unchanged against the facade headers (its GLFWindow.h
    pulls in gdt/math/AffineSpace.h, main.cpp:4 and GLFWindow.h:7);
  * the facade sources compile against gdt's types;
  * main.cpp's own scene (head, body, base and plane meshes with their PNG /
    JPG texture maps, main.cpp:165-246) loads through Scene::AddMesh with
    stb decoding, test.json giving only the settings as LoadFromJson does in
    the reference (SceneParser.cpp:11-23).

Skipped where the reference tree is absent.  Nothing here touches a GPU: the
headless driver's dump mode exits before any device call.
"""
import json
import os
import shutil
import struct
import subprocess
import zlib
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
REF = Path("/root/reference")
FACADE = ROOT / "optixrender_amd" / "cpp"
LIB = ROOT / "optixrender_amd"

pytestmark = pytest.mark.skipif(not (REF / "main.cpp").exists() or shutil.which("g++") is None,
                                reason="the reference tree is not present")


def ref_includes():
    # facade first, then the paths the reference's CMakeLists.txt gives main.cpp
    return ["-I", str(FACADE), "-I", str(REF / "common" / "gdt"), "-I", str(REF / "common"),
            "-I", str(REF / "common" / "3rdParty" / "nlohmann_json" / "include"),
            "-I", str(REF / "common" / "3rdParty" / "glfw" / "include")]


def test_main_cpp_compiles_against_the_facade():
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", *ref_includes(), str(REF / "main.cpp")],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]


@pytest.mark.parametrize("src", ["Scene.cpp", "SceneParser.cpp", "Renderer.cpp", "headless_main.cpp"])
def test_facade_compiles_with_gdt_types(src):
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wno-comment", *ref_includes(),
                        str(FACADE / src)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-4000:]
    own = [ln for ln in r.stderr.splitlines() if ln.startswith(str(FACADE)) and "warning:" in ln]
    assert not own, "\n".join(own)      # warnings of the vendored headers themselves are theirs


@pytest.fixture(scope="module")
def headless_in_tree(tmp_path_factory):
    """The headless driver built as it would be inside the reference's tree."""
    if not (LIB / "liboxr.so").exists():
        pytest.skip("liboxr.so not built")
    out = tmp_path_factory.mktemp("intree") / "headless"
    srcs = [str(FACADE / f) for f in ("Scene.cpp", "SceneParser.cpp", "Renderer.cpp", "headless_main.cpp")]
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-Wno-comment", *ref_includes(), *srcs, "-L", str(LIB), "-loxr",
                        f"-Wl,-rpath,{LIB}", "-o", str(out)], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    return out


def png_size(path):
    head = Path(path).read_bytes()[:24]
    assert head[:8] == b"\x89PNG\r\n\x1a\n"
    return struct.unpack(">II", head[16:24])


def jpg_size(path):
    data = Path(path).read_bytes()
    i = 2
    while i < len(data):
        marker, length = data[i + 1], struct.unpack(">H", data[i + 2:i + 4])[0]
        if marker in (0xC0, 0xC1, 0xC2):
            h, w = struct.unpack(">HH", data[i + 5:i + 9])
            return w, h
        i += 2 + length
    raise AssertionError("no SOF marker")


def run_dump(exe, tmp_path, env_extra, args=("", "x.ppm", "1")):
    out = tmp_path / "dump.json"
    env = dict(os.environ, OXR_HEADLESS_DUMP=str(out), OXR_HEADLESS_DUMP_BRIEF="1", **env_extra)
    a = list(args)
    a[1] = str(tmp_path / a[1])
    r = subprocess.run([str(exe), *a], capture_output=True, text=True, timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    return json.loads(out.read_text()), r


def test_main_scene_loads_with_its_texture_maps(headless_in_tree, tmp_path):
    models = REF / "models"
    d, r = run_dump(headless_in_tree, tmp_path, {"OXR_HEADLESS_MAIN_SCENE": str(models)})
    tris = [m["triangles"] for m in d["meshes"]]
    assert tris == [40370, 37712, 1000, 200]          # head, body, base, plane: 79,282 in all
    # texture ids in AddMesh's order (albedo, roughness, metallic, normal for the
    # head and base; the plane keeps the base's metallic file, which main.cpp
    # never clears), images decoded by stb
    files = ["01_Head_Base_Color.png", "01_Head_Roughness.png", "01_Head_Metallic.png", "01_Head_Normal_DirectX.png",
             "03_Base_Base_Color.png", "03_Base_Roughness.png", "03_Base_Metallic.png", "03_Base_Normal_DirectX.png",
             "grid.jpg", "rusty_metal_sheet_rough_1k.png", "03_Base_Metallic.png"]
    assert len(d["textures"]) == len(files)
    for f, t in zip(files, d["textures"]):
        w, h = jpg_size(models / f) if f.endswith(".jpg") else png_size(models / f)
        assert (t["width"], t["height"]) == (w, h), f
        assert t["bytes"] == 4 * w * h and t["byte_sum"] > 0
    head, body, base, plane = [m["material"] for m in d["meshes"]]
    assert head["tex"][0] == 0 and head["tex"][1] == 1 and head["tex"][4] == 2 and head["tex"][7] == 3
    assert body["type"] == 2 and body["tex"] == [-1] * 8 and d["meshes"][1]["in_medium"] == 0
    assert plane["type"] == 0 and plane["tex"][0] == 8 and plane["tex"][1] == 9 and plane["tex"][4] == 10
    assert plane["tex"][7] == -1
    # test.json gives the settings only (800 x 600, depth 5, Gaussian filter);
    # its shapes are not added, as in the reference
    assert (d["width"], d["height"], d["depth"], d["filter"]) == (800, 600, 5, 3)
    # the HDR main.cpp names is not in models/: no environment, as the reference's empty one
    assert d["env"] is None or d["env"]["width"] == 0
    assert "no environment" in r.stderr


def write_png(path, rgba):
    """An 8-bit RGBA PNG with filter type 0 on every row."""
    h, w = rgba.shape[:2]
    raw = b"".join(b"\0" + rgba[y].tobytes() for y in range(h))

    def chunk(tag, body):
        return struct.pack(">I", len(body)) + tag + body + struct.pack(">I", zlib.crc32(tag + body) & 0xFFFFFFFF)

    path.write_bytes(b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)) +
                     chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b""))


def test_png_texels_and_row_order(headless_in_tree, tmp_path):
    # a PNG written here, decoded by the facade through stb: texels exact and
    # rows flipped (row 0 of the texture is the image's bottom row)
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, (7, 11, 4), dtype=np.uint8)
    write_png(tmp_path / "t.png", img)
    (tmp_path / "tri.obj").write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    out = tmp_path / "full.json"
    env = dict(os.environ, OXR_HEADLESS_DUMP=str(out), OXR_HEADLESS_ALBEDO=str(tmp_path / "t.png"))
    r = subprocess.run([str(headless_in_tree), "", str(tmp_path / "x.ppm"), "1", str(tmp_path / "tri.obj")],
                       capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    d = json.loads(out.read_text())
    t = d["textures"][0]
    got = np.array(t["rgba"], np.uint8).reshape(t["height"], t["width"], 4)
    assert np.array_equal(got, img[::-1])
