"""Loop restoration on the GPU against the oracle with the unit types forced:
every unit self-guided, its parameter set running through all 16 sets (both
passes, r = 1 alone, r = 2 alone), or every unit Wiener, at 8, 10 and 12
bits in 4:2:0 (chroma units of 32 samples), 4:2:2, 4:4:4 and monochrome.
The inputs are the synthetic frames, and frames driven to the sample
extremes (mostly the maximum, with zeros sprinkled in), where the
self-guided sums and products are largest (postfilter.rs:781-991)."""
import sys

import numpy as np
import pytest

import common
import pyoracle
from rvpf import api

pytestmark = pytest.mark.gpu

CASES = [
    ("352x288_8b_420", dict(width=352, height=288, bit_depth=8, qindex=128, lr=2, seed=41)),
    ("354x290_10b_420", dict(width=354, height=290, bit_depth=10, qindex=160, lr=2, seed=42)),
    ("320x256_12b_mono", dict(width=320, height=256, bit_depth=12, mono=True, qindex=120, lr=2, seed=43)),
    ("330x262_12b_444", dict(width=330, height=262, bit_depth=12, ssx=0, ssy=0, qindex=64, lr=2, seed=44)),
    ("300x200_10b_422", dict(width=300, height=200, bit_depth=10, ssx=1, ssy=0, qindex=180, lr=2, seed=45)),
]

XQD0, XQD1 = (-96, 31), (-32, 95)
WIENER_MIN, WIENER_MAX = (-5, -23, -17), (10, 8, 46)
SGR_R = [(2, 1)] * 10 + [(0, 1)] * 4 + [(2, 0)] * 2  # radii of the 16 sets


@pytest.fixture(scope="module")
def ctx():
    c = api.Context()
    yield c
    c.close()


def force_units(f, mode, rng):
    for p in range(f.num_planes):
        u = f.lr_units[p].reshape(-1, 12)
        for k in range(u.shape[0]):
            u[k, :] = 0
            if mode == "sgr":
                s = (k + 5 * p) % 16
                x0 = int(rng.integers(XQD0[0], XQD0[1] + 1)) if SGR_R[s][0] else 0
                x1 = int(rng.integers(XQD1[0], XQD1[1] + 1)) if SGR_R[s][1] else min(max(128 - x0, XQD1[0]), XQD1[1])
                u[k, 0] = 2
                u[k, 1] = s
                u[k, 2:4] = np.array([x0, x1], dtype=np.int8).view(np.uint8)
            else:
                taps = [[int(rng.integers(WIENER_MIN[i], WIENER_MAX[i] + 1)) for i in range(3)] for _ in range(2)]
                if p:
                    taps[0][0] = taps[1][0] = 0
                u[k, 0] = 1
                u[k, 4:10] = np.array(taps[0] + taps[1], dtype=np.int8).view(np.uint8)


def extreme(a, maxv, rng):
    out = np.full_like(a, maxv)
    out[rng.random(a.shape) < 0.08] = 0
    return out


@pytest.mark.parametrize("data", ["synthetic", "extreme"])
@pytest.mark.parametrize("mode", ["sgr", "wiener"])
@pytest.mark.parametrize("name,kw", CASES, ids=[c[0] for c in CASES])
def test_lr_forced(ctx, name, kw, mode, data):
    print("case", name, mode, data, flush=True)
    sys.stdout.flush()
    rng = np.random.default_rng(kw["seed"])
    f = common.make_frame(**kw)
    force_units(f, mode, rng)
    maxv = (1 << kw["bit_depth"]) - 1
    if data == "synthetic":
        deb = [None if a is None else a.copy() for a in f.rec]
        cdef = pyoracle.cdef(f, deb)
    else:
        deb = [None if a is None else extreme(a, maxv, rng) for a in f.rec]
        cdef = [None if a is None else extreme(a, maxv, rng) for a in f.rec]
    want = pyoracle.lr(f, deb, cdef)
    df = api.DeviceFrame(f)
    deb_dev, cdef_dev = df.upload_planes(deb), df.upload_planes(cdef)
    out = df.alloc_like_recon()
    api.lr(ctx, df, deb_dev, cdef_dev, out)
    ctx.sync()
    common.assert_planes(api.to_host(out), want, f, "mi", "%s %s %s lr" % (name, mode, data))
