"""Pin the oracle's RNG (gdt::LCG<16>, random.h:55-89).

Two independent references: the known answers produced by the reference's own
random.h (tests/golden/gdt_kat.json, written by tests/golden/make_gdt_kat.py
from oracle/_ref), and TEA-16 + the Numerical Recipes LCG restated here from
the published algorithms (Wheeler & Needham 1994; Press et al.)."""
import json
import struct
from pathlib import Path

import oracle_lib

KAT = json.loads((Path(__file__).parent / "golden" / "gdt_kat.json").read_text())
M = 0xFFFFFFFF


def tea16(v0, v1):
    s = 0
    for _ in range(16):
        s = (s + 0x9E3779B9) & M
        v0 = (v0 + ((((v1 << 4) & M) + 0xA341316C) ^ ((v1 + s) & M) ^ ((v1 >> 5) + 0xC8013EA4))) & M
        v1 = (v1 + ((((v0 << 4) & M) + 0xAD90777D) ^ ((v0 + s) & M) ^ ((v0 >> 5) + 0x7E95761E))) & M
    return v0


def bits(f):
    return struct.unpack("<I", struct.pack("<f", f))[0]


def test_seed_matches_reference_headers():
    for case in KAT["lcg"]:
        assert oracle_lib.lcg_init(case["v0"], case["v1"]) == case["state0"]
        assert tea16(case["v0"], case["v1"]) == case["state0"]


def test_draws_match_reference_headers():
    for case in KAT["lcg"]:
        got = oracle_lib.lcg_draws(case["state0"], len(case["draws"]))
        for (f, st), (fb, sb) in zip(got, case["draws"]):
            assert st == sb
            assert bits(f) == fb


def test_lcg_closed_form():
    s = 12345
    for f, st in oracle_lib.lcg_draws(s, 100):
        s = (1664525 * s + 1013904223) & M
        assert st == s
        assert f == (s & 0xFFFFFF) / 16777216.0
        assert 0.0 <= f < 1.0
