"""The spec tables in av1_tables.h equal the reference's src/tables.rs
(hashes recorded by tests/golden/make_table_hashes.py)."""
import json
import os
import re

from golden.make_table_hashes import NAMES, values_hash

HERE = os.path.dirname(os.path.abspath(__file__))
HDR = os.path.join(os.path.dirname(HERE), "rivet-av1_amd", "csrc", "av1_tables.h")


def header_tables():
    src = open(HDR).read()
    out = {}
    for m in re.finditer(r"AV1_TABLE \w+ AV1_(\w+)((?:\[\d+\])+) = (\{.*?\});\n", src, re.S):
        vals = [int(v) for v in re.findall(r"-?\d+", m.group(3))]
        out[m.group(1)] = vals
    return out


def test_tables_match_reference():
    want = json.load(open(os.path.join(HERE, "golden", "table_hashes.json")))
    got = header_tables()
    for n in NAMES:
        assert n in got, n
        assert len(got[n]) == want[n]["count"], n
        assert values_hash(got[n]) == want[n]["sha256"], n


RANGES = {"int8_t": (-128, 127), "uint8_t": (0, 255), "int16_t": (-32768, 32767), "uint16_t": (0, 65535),
          "int32_t": (-2 ** 31, 2 ** 31 - 1)}


def test_table_types_hold_their_values():
    src = open(HDR).read()
    for m in re.finditer(r"AV1_TABLE (\w+) AV1_(\w+)((?:\[\d+\])+) = (\{.*?\});\n", src, re.S):
        lo, hi = RANGES[m.group(1)]
        vals = [int(v) for v in re.findall(r"-?\d+", m.group(4))]
        assert all(lo <= v <= hi for v in vals), m.group(2)


def test_derived_tables():
    t = header_tables()
    # Div_Table is 840 / k (7.15.2)
    assert t["DIV_TABLE"] == [0] + [840 // k for k in range(1, 9)]
    # Gaussian sequence: 2048 entries, multiples of 4 within 12 bits
    g = t["GAUSSIAN_SEQUENCE"]
    assert len(g) == 2048 and all(v % 4 == 0 and -2048 <= v < 2048 for v in g)
    # every super-resolution filter phase sums to 128 (7.16)
    u = t["UPSCALE_FILTER"]
    assert len(u) == 512 and all(sum(u[8 * i:8 * i + 8]) == 128 for i in range(64))
