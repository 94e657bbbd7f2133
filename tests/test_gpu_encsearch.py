"""GPU parity of the encoder filter searches (rvpf_lf_search,
rvpf_cdef_search) against the oracle's search_lf and cdef::search: the same
levels, the same CDEF parameters and the same per-block indices."""
import sys

import numpy as np
import pytest

import pyoracle
from rvpf import api, workload

pytestmark = pytest.mark.gpu

CASES = [
    ("200x136_10b_420", dict(width=200, height=136, bit_depth=10, qindex=140, lr=0, seed=31), 2),
    ("130x72_8b_444", dict(width=130, height=72, bit_depth=8, ssx=0, ssy=0, qindex=200, lr=0, seed=32), 2),
    ("96x80_12b_mono", dict(width=96, height=80, bit_depth=12, mono=True, qindex=90, lr=0, seed=33), 2),
    ("150x100_8b_422_sb128", dict(width=150, height=100, bit_depth=8, ssx=1, ssy=0, qindex=60, lr=0, seed=34,
                                  sb_size=128), 2),
    ("264x136_8b_420_u8", dict(width=264, height=136, bit_depth=8, qindex=120, lr=0, seed=35, lf_extras=True), 1),
    ("330x200_10b_420_extras", dict(width=330, height=200, bit_depth=10, qindex=180, lr=0, seed=36,
                                    lf_extras=True, frame_index=2, pan_px=1.5), 2),
    # frame sizes of the reference's odd_sizes test, whose encoder runs these searches (tests/encode.rs:119)
    ("1x1_8b_420_u8", dict(width=1, height=1, bit_depth=8, qindex=60, lr=0, seed=37), 1),
    ("7x5_10b_444", dict(width=7, height=5, bit_depth=10, ssx=0, ssy=0, qindex=60, lr=0, seed=38), 2),
    ("16x200_12b_422", dict(width=16, height=200, bit_depth=12, ssx=1, ssy=0, qindex=60, lr=0, seed=39), 2),
    ("130x34_8b_mono", dict(width=130, height=34, bit_depth=8, mono=True, qindex=60, lr=0, seed=40), 2),
]
IDS = [c[0] for c in CASES]


@pytest.fixture(scope="module")
def ctx():
    c = api.Context()
    yield c
    c.close()


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name,kw,bps", CASES, ids=IDS)
@pytest.mark.parametrize("thorough,lam", [(False, 0.0), (False, 2500.0), (True, 400.0), (False, 1e12)])
def test_cdef_search(ctx, name, kw, bps, thorough, lam):
    _say("cdef search", name, thorough, lam)
    f = workload.Frame(**kw)
    deb = pyoracle.loop_filter(f, [None if a is None else a.copy() for a in f.rec])
    want, want_table, _, _ = pyoracle.cdef_search(f, deb, f.src, kw["qindex"], lam, thorough, threads=4)
    df = api.DeviceFrame(f, bytes_per_sample=bps)
    got, table = api.cdef_search(ctx, df, df.upload_planes(deb), df.upload_planes(f.src), kw["qindex"], lam,
                                 thorough)
    assert got.as_tuple() == want.as_tuple(), (got.as_tuple(), want.as_tuple())
    g = f.geom
    n = ((g.mi_rows + 15) // 16) * g.cdef_stride
    assert np.array_equal(table[:n], want_table[:n])


@pytest.mark.parametrize("name,kw,bps", CASES, ids=IDS)
def test_lf_search(ctx, name, kw, bps):
    _say("lf search", name)
    f = workload.Frame(**kw)
    want, _ = pyoracle.lf_search(f, [None if a is None else a.copy() for a in f.rec], f.src, threads=4)
    df = api.DeviceFrame(f, bytes_per_sample=bps)
    rec = df.upload_planes(f.rec)
    got = api.lf_search(ctx, df, rec, df.upload_planes(f.src))
    assert got == want
    for p, t in enumerate(api.to_host(rec)):
        if t is not None:
            assert np.array_equal(t, f.rec[p]), "recon was modified"


def test_cdef_search_without_coded_blocks(ctx):
    name, kw, bps = CASES[0]
    f = workload.Frame(**kw)
    f.cdef_idx[:] = -1
    df = api.DeviceFrame(f)
    got, table = api.cdef_search(ctx, df, df.upload_planes(f.rec), df.upload_planes(f.src), 100, 10.0)
    assert got.as_tuple() == (0, 0, [(0, 0)], [(0, 0)])
    assert np.all(table == -1)
