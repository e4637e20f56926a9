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
