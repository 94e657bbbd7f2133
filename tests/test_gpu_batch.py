"""GPU parity of the batched path (rvpf_batch_*): one graph launch over a
mixed set of frames (8/10/12 bits, 8- and 16-bit sample planes, 4:2:0,
4:2:2, 4:4:4, mono, stages switched off, frames not shown) equals the
oracle chain frame by frame, on every replay of the graph, for every tile
shape and grid the autotuner can choose, and after autotuning."""
import pytest

import common
import pyoracle
from rvpf import api

pytestmark = pytest.mark.gpu

CASES = dict(common.SMALL_CASES)
# (frame kwargs, bytes per sample, shown)
BATCH = [
    (CASES["c1_352x288_8b_420_wiener"], 1, True),
    (CASES["66x66_8b_420_all"], 1, True),
    (CASES["226x226_10b_420_all"], 2, True),
    (CASES["200x120_12b_420_all"], 2, False),
    (CASES["170x90_10b_444_all"], 2, True),
    (CASES["150x100_8b_422_all"], 2, True),
    (CASES["130x70_12b_mono_all"], 2, True),
    (CASES["257x129_10b_420_inter"], 2, True),
    (CASES["96x64_8b_420_sharp"], 1, True),
    (CASES["64x40_10b_444_lag1"], 2, True),
    (dict(width=120, height=80, bit_depth=8, qindex=90, lf=False, cdef=False, lr=0, seed=21), 1, True),
    (dict(width=100, height=60, bit_depth=10, qindex=150, cdef=False, lr=2, grain=True, seed=22), 2, True),
    (dict(width=90, height=50, bit_depth=10, qindex=70, lr=0, grain=False, seed=23), 2, True),
]


def _say(*a):
    print(*a, flush=True)


class Entry:
    def __init__(self, kw, bps, shown):
        self.f = common.make_frame(**kw)
        self.want = pyoracle.postfilter_frame(self.f, threads=4)
        self.df = api.DeviceFrame(self.f, bytes_per_sample=bps)
        self.pristine = self.df.upload_planes(self.f.rec)
        self.rec = [None if t is None else t.clone() for t in self.pristine]
        self.ref = self.df.alloc_like_recon()
        self.disp = self.df.alloc_display() if shown else None
        self.name = "%dx%d %d-bit bps %d" % (self.f.params.frame_width, self.f.params.frame_height,
                                            self.f.bit_depth, bps)

    def refill(self):
        for t, s in zip(self.rec, self.pristine):
            if t is not None:
                t.copy_(s)
        for t in self.ref + (self.disp or []):
            if t is not None:
                t.fill_(0)

    def check(self, what):
        deb, ref, disp = self.want
        common.assert_planes(api.to_host(self.rec), deb, self.f, "full", "%s %s deblocked" % (what, self.name))
        common.assert_planes(api.to_host(self.ref), ref, self.f, "mi", "%s %s reference" % (what, self.name))
        if self.disp is not None:
            common.assert_planes(api.to_host(self.disp), disp, self.f, "crop", "%s %s display" % (what, self.name))


@pytest.fixture(scope="module")
def ctx():
    c = api.Context()
    yield c
    c.close()


@pytest.fixture(scope="module")
def entries():
    return [Entry(kw, bps, shown) for kw, bps, shown in BATCH]


def _run_and_check(ctx, batch, entries, what):
    import torch
    for e in entries:
        e.refill()
    batch.run()
    ctx.sync()
    torch.cuda.synchronize()
    for e in entries:
        e.check(what)


def test_batch_replays(ctx, entries):
    _say("batch of", len(entries), "frames, default tiles, 3 replays")
    b = api.Batch(ctx, [(e.df, e.rec, e.ref, e.disp) for e in entries])
    try:
        for r in range(3):
            _say("replay", r)
            _run_and_check(ctx, b, entries, "replay %d" % r)
    finally:
        b.close()


TILES = [
    {"deblock": (1, 4, 1), "grain_apply": (0, 4, 1)},
    {"deblock": (2, 2, 4), "grain_apply": (0, 2, 4), "copy": (0, 4, 4)},
    {"deblock": (0, 0, 1), "cdef": (0, 0, 1), "lr": (0, 0, 1), "grain_prepare": (0, 0, 1),
     "grain_apply": (0, 0, 1), "copy": (0, 0, 1)},
    {"cdef": (0, 1, 2), "lr": (0, 1, 2), "grain_prepare": (0, 2, 1)},
]


@pytest.mark.parametrize("tiles", TILES, ids=["t%d" % i for i in range(len(TILES))])
def test_batch_tile_choices(ctx, entries, tiles):
    _say("tiles", tiles)
    b = api.Batch(ctx, [(e.df, e.rec, e.ref, e.disp) for e in entries])
    try:
        b.set_tiles(tiles)
        got = b.get_tiles()
        for k, v in tiles.items():
            assert got[k][:3] == v
        _run_and_check(ctx, b, entries, "tiles %s" % tiles)
        _run_and_check(ctx, b, entries, "tiles %s, replay" % tiles)
    finally:
        b.close()


def test_batch_autotune_and_cache(ctx, entries):
    _say("autotune")
    b = api.Batch(ctx, [(e.df, e.rec, e.ref, e.disp) for e in entries])
    try:
        b.autotune(force=True)
        tiles = b.get_tiles()
        for stage in ("deblock", "cdef", "lr", "grain_prepare", "grain_apply", "copy"):
            assert tiles[stage][3] > 0, stage  # every stage has work here and was timed
        _run_and_check(ctx, b, entries, "autotuned")
    finally:
        b.close()
    # a batch of the same shape takes the remembered tiles without timing
    b2 = api.Batch(ctx, [(e.df, e.rec, e.ref, e.disp) for e in entries])
    try:
        assert {k: v[:3] for k, v in b2.get_tiles().items()} == {k: v[:3] for k, v in tiles.items()}
        b2.autotune(force=False)
        assert {k: v[:3] for k, v in b2.get_tiles().items()} == {k: v[:3] for k, v in tiles.items()}
        _run_and_check(ctx, b2, entries, "cached tiles")
    finally:
        b2.close()


def test_batch_odd_sizes(ctx):
    """The reference's odd frame sizes (1x1 to 130x34, four layouts, 8-bit
    ones in 8-bit planes) in one batch: every frame against the oracle, on
    two replays, with the default tiles and after autotuning."""
    odd = [Entry(kw, 1 if kw["bit_depth"] == 8 else 2, True) for _, kw in common.ODD_CASES]
    _say("batch of", len(odd), "odd-size frames")
    b = api.Batch(ctx, [(e.df, e.rec, e.ref, e.disp) for e in odd])
    try:
        for rep in range(2):
            _run_and_check(ctx, b, odd, "odd replay %d" % rep)
        b.autotune(force=True)
        _run_and_check(ctx, b, odd, "odd autotuned")
    finally:
        b.close()
