"""The C++ host facade (Renderer / Scene / SceneParser / CUDABuffer over the C
ABI) builds and drives the library like the reference's main.cpp does."""
import json
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
HEADLESS = ROOT / "optixrender_amd" / "headless"


def test_headless_driver_built():
    assert HEADLESS.exists(), "run `python -m optixrender_amd.build`"


def read_ppm(path):
    data = path.read_bytes()
    head = data.split(b"\n", 3)
    w, h = map(int, head[1].split())
    return np.frombuffer(head[3], np.uint8).reshape(h, w, 3)


@pytest.mark.gpu
def test_headless_renders_settings_json(tmp_path):
    out = tmp_path / "img.ppm"
    r = subprocess.run([str(HEADLESS), str(ROOT / "tests" / "data" / "settings.json"), str(out), "4"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    info = json.loads(r.stdout.strip().splitlines()[-1])
    assert info["width"] == 64 and info["height"] == 48
    assert info["frame_id"] == 4          # progressive: frameID advances once per Render()
    img = read_ppm(out)
    assert img.shape == (48, 64, 3)
    assert img.mean() > 10                # the lit box is visible


def test_missing_obj_is_reported(tmp_path):
    # LoadMesh throws on a missing OBJ; main catches std::runtime_error and exits 1
    r = subprocess.run([str(HEADLESS), "", str(tmp_path / "x.ppm"), "1", str(tmp_path / "missing.obj")],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "FATAL ERROR" in r.stdout


@pytest.mark.gpu
def test_headless_denoised_frames(tmp_path):
    # denoiserOn through the facade: weights from denoiser_kpcnn.f32 next to the binary
    import os
    out = tmp_path / "dn.ppm"
    env = dict(os.environ, OXR_HEADLESS_DENOISE="1")
    r = subprocess.run([str(HEADLESS), str(ROOT / "tests" / "data" / "settings.json"), str(out), "2"],
                       capture_output=True, text=True, timeout=120, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    img = read_ppm(out)
    assert img.shape == (48, 64, 3) and img.mean() > 10


@pytest.mark.gpu
def test_headless_textured_obj(tmp_path):
    # main.cpp's path for textured meshes: AddMesh with a TextureFile, an HDR
    # environment, Renderer uploading Scene::textures and the OBJ's texture
    # coordinates.  A quad facing the camera whose albedo image is red on its
    # left half (u < 0.5) and blue on its right half, under a white sky.
    obj = tmp_path / "quad.obj"
    obj.write_text("v -1 -1 -2\nv 1 -1 -2\nv 1 1 -2\nv -1 1 -2\n"
                   "vt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\n"
                   "vn 0 0 1\nf 1/1/1 2/2/1 3/3/1\nf 1/1/1 3/3/1 4/4/1\n")
    tex = np.zeros((8, 8, 3), np.uint8)
    tex[:, :4] = (230, 20, 20)
    tex[:, 4:] = (20, 20, 230)
    (tmp_path / "albedo.ppm").write_bytes(b"P6\n8 8\n255\n" + tex.tobytes())
    env = np.ones((4, 8, 3), np.float32)
    (tmp_path / "sky.pfm").write_bytes(b"PF\n8 4\n-1.0\n" + env.astype("<f4").tobytes())
    import os
    run_env = dict(os.environ, OXR_HEADLESS_ALBEDO=str(tmp_path / "albedo.ppm"),
                   OXR_HEADLESS_ENV=str(tmp_path / "sky.pfm"))
    out = tmp_path / "tex.ppm"
    r = subprocess.run([str(HEADLESS), str(ROOT / "tests" / "data" / "settings.json"), str(out), "4", str(obj)],
                       capture_output=True, text=True, timeout=120, env=run_env)
    assert r.returncode == 0, r.stdout + r.stderr
    img = read_ppm(out).astype(np.float64)
    h, w = img.shape[:2]
    mid = img[h // 3: 2 * h // 3]
    # the quad spans about the middle half of the image's width
    left, right = mid[:, int(0.28 * w): int(0.42 * w)], mid[:, int(0.58 * w): int(0.72 * w)]
    assert left[..., 0].mean() > 2 * left[..., 2].mean()     # red half on the left
    assert right[..., 2].mean() > 2 * right[..., 0].mean()   # blue half on the right


@pytest.mark.gpu
def test_headless_present_readback_is_the_same_image(tmp_path):
    # presentOn: the frame comes back through the streamed page-locked copy
    import os
    imgs = []
    for present in ("0", "1"):
        out = tmp_path / f"p{present}.ppm"
        env = dict(os.environ, OXR_HEADLESS_PRESENT=present)
        r = subprocess.run([str(HEADLESS), str(ROOT / "tests" / "data" / "settings.json"), str(out), "3"],
                           capture_output=True, text=True, timeout=120, env=env)
        assert r.returncode == 0, r.stdout + r.stderr
        imgs.append(read_ppm(out))
    assert imgs[0].mean() > 10 and np.array_equal(imgs[0], imgs[1])
