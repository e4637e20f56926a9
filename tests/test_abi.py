"""The C-ABI library loads and exports every entry point include/oxr.h declares;
the oracle's structs have the C ABI's layout.  CPU only (no compute calls)."""
import ctypes as C
import re
from pathlib import Path

from optixrender_amd import abi

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "oxr.h").read_text()
    return sorted(set(re.findall(r"\b(oxr_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(abi.LIB_PATH))
    names = declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert sorted(abi.SIGNATURES) == declared_symbols()


def test_struct_sizes_match_header_layout():
    # every field is 4 bytes (or a pointer / int64 in the scene descriptor)
    assert C.sizeof(abi.Material) == 4 * 32
    assert C.sizeof(abi.Light) == 4 * 16
    assert C.sizeof(abi.Medium) == 4 * 12
    assert C.sizeof(abi.Camera) == 4 * 11
    assert C.sizeof(abi.RenderParams) == 4 * 8


def test_oracle_structs_share_layout():
    import oracle_lib

    text = (ROOT / "oracle" / "oracle.h").read_text()
    body = text[text.index("typedef struct orc_material"):text.index("} orc_material;")]
    fields = re.findall(r"(int32_t|float)\s+([^;]+);", body)
    n = sum(len(names.split(",")) + sum(int(x) - 1 for x in re.findall(r"\[(\d+)\]", names)) for _, names in fields)
    assert n * 4 == C.sizeof(abi.Material)
    assert C.sizeof(oracle_lib.OrcRenderParams) == 40
