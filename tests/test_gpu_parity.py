"""GPU parity: the HIP chain (through the C ABI) against the CPU oracle,
sample for sample, per stage and for the whole chain."""
import sys

import numpy as np
import pytest

import common
import pyoracle
from rvpf import api

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = api.Context()
    yield c
    c.close()


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name,kw", common.SMALL_CASES + common.ODD_CASES,
                         ids=[c[0] for c in common.SMALL_CASES + common.ODD_CASES])
def test_stages(ctx, name, kw):
    _say("case", name)
    f = common.make_frame(**kw)
    df = api.DeviceFrame(f)
    # deblock: in place, whole buffer
    want_deb = pyoracle.loop_filter(f, [None if a is None else a.copy() for a in f.rec]) \
        if (f.params.loop_filter_level[0] or f.params.loop_filter_level[1]) else [a for a in f.rec]
    planes = df.upload_planes(f.rec)
    api.deblock(ctx, df, planes)
    ctx.sync()
    common.assert_planes(api.to_host(planes), want_deb, f, "full", name + " deblock")
    # CDEF from the oracle's deblocked frame
    want_cdef = pyoracle.cdef(f, want_deb)
    deb_dev = df.upload_planes(want_deb)
    out = df.alloc_like_recon()
    api.cdef(ctx, df, deb_dev, out)
    ctx.sync()
    common.assert_planes(api.to_host(out), want_cdef, f, "mi", name + " cdef")
    # LR from the oracle's deblocked + CDEF frames
    want_lr = pyoracle.lr(f, want_deb, want_cdef)
    cdef_dev = df.upload_planes(want_cdef)
    out2 = df.alloc_like_recon()
    api.lr(ctx, df, deb_dev, cdef_dev, out2)
    ctx.sync()
    common.assert_planes(api.to_host(out2), want_lr, f, "mi", name + " lr")
    # display (crop + grain) from the oracle's LR frame
    w, h = f.params.upscaled_width, f.params.frame_height
    # a copy: a crop of one row (or 1x1) is contiguous, and the in-place
    # grain would otherwise write into want_lr, the GPU stage's input
    want_disp = [None if p >= f.num_planes else
                 want_lr[p][:f.crop_shape(p)[0], :f.crop_shape(p)[1]].copy() for p in range(3)]
    if f.params.film_grain_params_present and f.params.grain.apply_grain:
        pyoracle.grain_apply(f.params, w, h, want_disp)
    lr_dev = df.upload_planes(want_lr)
    disp = df.alloc_display()
    api.grain(ctx, df, lr_dev, disp)
    ctx.sync()
    common.assert_planes(api.to_host(disp), want_disp, f, "crop", name + " display")


@pytest.mark.parametrize("name,kw", common.SMALL_CASES, ids=[c[0] for c in common.SMALL_CASES])
def test_chain(ctx, name, kw):
    _say("case", name)
    f = common.make_frame(**kw)
    deb, ref, disp = pyoracle.postfilter_frame(f, threads=4)
    df = api.DeviceFrame(f)
    rec = df.upload_planes(f.rec)
    ref_dev = df.alloc_like_recon()
    disp_dev = df.alloc_display()
    api.postfilter_frame(ctx, df, rec, ref_dev, disp_dev)
    ctx.sync()
    common.assert_recon_kept(api.to_host(rec), f, name)
    common.assert_planes(api.to_host(ref_dev), ref, f, "mi", name + " reference")
    common.assert_planes(api.to_host(disp_dev), disp, f, "crop", name + " display")


@pytest.mark.parametrize("name,kw", common.ODD_CASES + common.EXTREME_CASES,
                         ids=[c[0] for c in common.ODD_CASES + common.EXTREME_CASES])
def test_chain_odd_sizes(ctx, name, kw):
    """The reference's odd frame sizes (down to 1x1) and the widest and
    tallest synthetic frames through the whole chain, in 16-bit planes and,
    at 8 bits, in 8-bit planes."""
    _say("case odd", name)
    f = common.make_frame(**kw)
    deb, ref, disp = pyoracle.postfilter_frame(f, threads=4)
    for bps in ((2, 1) if kw["bit_depth"] == 8 else (2,)):
        df = api.DeviceFrame(f, bytes_per_sample=bps)
        rec = df.upload_planes(f.rec)
        ref_dev = df.alloc_like_recon()
        disp_dev = df.alloc_display()
        api.postfilter_frame(ctx, df, rec, ref_dev, disp_dev)
        ctx.sync()
        what = "%s bps %d" % (name, bps)
        common.assert_recon_kept(api.to_host(rec), f, what)
        common.assert_planes(api.to_host(ref_dev), ref, f, "mi", what + " reference")
        common.assert_planes(api.to_host(disp_dev), disp, f, "crop", what + " display")


U8_CASES = [c for c in common.SMALL_CASES if c[1]["bit_depth"] == 8]


@pytest.mark.parametrize("name,kw", U8_CASES, ids=[c[0] for c in U8_CASES])
def test_chain_u8_planes(ctx, name, kw):
    """8-bit frames in 8-bit sample planes (bytes_per_sample 1)."""
    _say("case u8", name)
    f = common.make_frame(**kw)
    deb, ref, disp = pyoracle.postfilter_frame(f, threads=4)
    df = api.DeviceFrame(f, bytes_per_sample=1)
    rec = df.upload_planes(f.rec)
    ref_dev = df.alloc_like_recon()
    disp_dev = df.alloc_display()
    api.postfilter_frame(ctx, df, rec, ref_dev, disp_dev)
    ctx.sync()
    common.assert_recon_kept(api.to_host(rec), f, name)
    common.assert_planes(api.to_host(ref_dev), ref, f, "mi", name + " reference")
    common.assert_planes(api.to_host(disp_dev), disp, f, "crop", name + " display")


GRAIN_LAYOUTS = {"420": dict(ssx=1, ssy=1), "422": dict(ssx=1, ssy=0), "444": dict(ssx=0, ssy=0),
                 "mono": dict(mono=True)}


@pytest.mark.parametrize("lag", [0, 1, 2, 3])
@pytest.mark.parametrize("layout", sorted(GRAIN_LAYOUTS))
@pytest.mark.parametrize("bd", [8, 10, 12])
@pytest.mark.parametrize("rows", ["w98", "w128", "w98pad"])
def test_grain_lags(ctx, lag, layout, bd, rows):
    """Film grain alone (templates, LUTs, offsets, blend) at every AR lag,
    chroma layout and bit depth, on random planes: rows of 98 samples
    (sample-by-sample kernel), of 128 (8-sample segments) and of 98 inside
    64-aligned rows (segments, the last one partial)."""
    _say("grain lag", lag, layout, bd, rows)
    width = 128 if rows == "w128" else 98
    f = common.make_frame(width=width, height=70, bit_depth=bd, qindex=120, lr=0, grain=True, grain_lag=lag,
                          seed=100 + 10 * lag + bd, **GRAIN_LAYOUTS[layout])
    g = f.params.grain
    if f.num_planes > 1 and not g.chroma_scaling_from_luma:
        # every template in use: cr takes cb's scaling points (or four fixed ones)
        if g.num_cb_points == 0:
            g.num_cb_points = 4
            for i, (x, y) in enumerate([(16, 40), (80, 200), (150, 90), (240, 160)]):
                g.point_cb_value[i], g.point_cb_scaling[i] = x, y
        g.num_cr_points = g.num_cb_points
        for i in range(g.num_cb_points):
            g.point_cr_value[i], g.point_cr_scaling[i] = g.point_cb_value[i], 255 - g.point_cb_scaling[i]
    rng = np.random.default_rng(lag * 100 + bd)
    planes = [None if p >= f.num_planes else
              rng.integers(0, 1 << bd, size=f.crop_shape(p), dtype=np.uint16) for p in range(3)]
    want = [None if a is None else a.copy() for a in planes]
    pyoracle.grain_apply(f.params, f.params.upscaled_width, f.params.frame_height, want)
    df = api.DeviceFrame(f)
    src = df.upload_planes(planes)
    disp = df.alloc_display()
    if rows == "w98pad":
        import torch
        def padded(t):
            if t is None:
                return None
            big = torch.zeros((t.shape[0], (t.shape[1] + 63) // 64 * 64), dtype=t.dtype, device=t.device)
            big[:, :t.shape[1]] = t
            return big[:, :t.shape[1]]
        src, disp = [padded(t) for t in src], [padded(t) for t in disp]
    api.grain(ctx, df, src, disp)
    ctx.sync()
    common.assert_planes(api.to_host(disp), want, f, "crop", "grain lag %d %s %d-bit %s" % (lag, layout, bd, rows))
