"""Build recipe for the HIP library (liboxr.so, gfx950) and the test oracle.

Run ``python -m optixrender_amd.build`` or ``__graft_entry__.build()``.  The
shared object is written in-tree next to this file so that it travels with
the repository snapshot.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "liboxr.so"

# -ffp-contract=off: no FMA contraction, so device arithmetic rounds as the
# reference's expressions are written.  The packed-fp32-ops feature is
# switched off (hipcc would otherwise emit v_pk_* FP32 ops on its own).
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero",
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]

SOURCES = ["api.cpp", "serve.cpp", "wavefront.cpp", "bvh.hip", "render.hip", "denoise.hip"]
STAMP = PKG / "liboxr.stamp"


def source_hash() -> str:
    """Hash of everything liboxr.so is built from; stored next to the library."""
    import hashlib

    h = hashlib.sha256()
    files = sorted(CSRC.glob("*")) + [ROOT / "include" / "oxr.h"]
    for f in files:
        if f.is_file():
            h.update(f.name.encode())
            h.update(f.read_bytes())
    return h.hexdigest()


def _run(cmd, cwd=None):
    proc = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError(f"build step failed ({proc.returncode}): {' '.join(map(str, cmd))}")
    return proc.stdout


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build_library(verbose: bool = False) -> Path:
    tmp = LIB.with_suffix(".so.tmp")
    objs = []
    objdir = PKG / "_build"
    objdir.mkdir(exist_ok=True)
    for src in SOURCES:
        obj = objdir / (src + ".o")
        flags = [f for f in HIP_FLAGS if f != "-shared"]
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        _run([hipcc(), *flags, *lang, "-c", str(CSRC / src), "-o", str(obj)])
        objs.append(str(obj))
    digest = source_hash()
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", str(tmp)])
    tmp.replace(LIB)
    STAMP.write_text(digest + "\n")
    if verbose:
        print(f"built {LIB}")
    return LIB


FACADE = ["Scene.cpp", "SceneParser.cpp", "Renderer.cpp", "headless_main.cpp"]
HEADLESS = PKG / "headless"


def build_facade(verbose: bool = False) -> Path:
    """The C++ host facade (Renderer/Scene/SceneParser/CUDABuffer) and the
    headless driver, linked against liboxr.so."""
    cpp = PKG / "cpp"
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-I", str(cpp), *[str(cpp / f) for f in FACADE], "-L", str(PKG),
          "-loxr", "-Wl,-rpath,$ORIGIN", "-o", str(HEADLESS)])
    if verbose:
        print(f"built {HEADLESS}")
    return HEADLESS


def build_oracle(verbose: bool = False) -> None:
    _run(["make", "-s", "-C", str(ROOT / "oracle")])
    if Path("/root/reference/common/gdt/gdt/random/random.h").exists():
        _run(["make", "-s", "-C", str(ROOT / "oracle"), "ref"])
    if verbose:
        print("built oracle")


def build_all(verbose: bool = False) -> None:
    build_library(verbose)
    build_facade(verbose)
    build_oracle(verbose)


if __name__ == "__main__":
    build_all(verbose=True)
