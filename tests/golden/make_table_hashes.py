"""Writes tests/golden/table_hashes.json: SHA-256 of each normative AV1
table the post-filter chain uses, taken from the reference's own
src/tables.rs (read as text in this container; the reference itself
never travels with the repository). The test compares them with the
values transcribed into rivet-av1_amd/csrc/av1_tables.h."""
import hashlib
import json
import os
import re
import sys

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference/src/tables.rs"
NAMES = ["CDEF_DIRECTIONS", "CDEF_PRI_TAPS", "CDEF_SEC_TAPS", "CDEF_UV_DIR", "DIV_TABLE",
         "NUM_4X4_BLOCKS_WIDE", "NUM_4X4_BLOCKS_HIGH", "SUBSAMPLED_SIZE", "TX_WIDTH", "TX_HEIGHT",
         "SGR_PARAMS", "WIENER_TAPS_MIN", "WIENER_TAPS_MAX", "WIENER_TAPS_MID", "SGRPROJ_XQD_MIN",
         "SGRPROJ_XQD_MAX", "SGRPROJ_XQD_MID", "AC_QLOOKUP", "GAUSSIAN_SEQUENCE", "UPSCALE_FILTER"]


def values_hash(vals):
    return hashlib.sha256(",".join(str(v) for v in vals).encode()).hexdigest()


def main():
    src = open(REF).read()
    out = {}
    for n in NAMES:
        m = re.search(r"static %s: [^=]+= (\[.*?\]);\n" % n, src, re.S)
        vals = [int(v) for v in re.findall(r"-?\d+", m.group(1))]
        out[n] = {"count": len(vals), "sha256": values_hash(vals)}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "table_hashes.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", path)


if __name__ == "__main__":
    main()
