"""GPU parity with super-resolution: the upscaling stage, loop restoration
on upscaled frames and the whole chain (through the C ABI) against the CPU
oracle, sample for sample."""
import sys

import numpy as np
import pytest

import common
import pyoracle
from rvpf import api

pytestmark = pytest.mark.gpu

SR_CASES = [
    ("sr9_226x120_10b_420", dict(width=226, height=120, bit_depth=10, qindex=160, lr=2, grain=True,
                                 lf_extras=True, seed=21, superres_denom=9)),
    ("sr16_200x96_8b_420", dict(width=200, height=96, bit_depth=8, qindex=120, lr=2, grain=True, seed=22,
                                superres_denom=16)),
    ("sr12_170x90_12b_444", dict(width=170, height=90, bit_depth=12, ssx=0, ssy=0, qindex=64, lr=2,
                                 grain=True, seed=23, superres_denom=12)),
    ("sr13_150x100_10b_422", dict(width=150, height=100, bit_depth=10, ssx=1, ssy=0, qindex=180, lr=1,
                                  seed=24, superres_denom=13)),
    ("sr11_130x70_8b_mono", dict(width=130, height=70, bit_depth=8, mono=True, qindex=120, lr=2, seed=25,
                                 superres_denom=11)),
    ("sr14_300x64_10b_420_nolr", dict(width=300, height=64, bit_depth=10, qindex=90, lr=0, grain=True,
                                      seed=26, superres_denom=14)),
    ("sr10_180x80_10b_420_nocdef", dict(width=180, height=80, bit_depth=10, qindex=150, lr=2, cdef=False,
                                        seed=27, superres_denom=10)),
    ("sr15_520x136_10b_420_u128", dict(width=520, height=136, bit_depth=10, qindex=140, lr=2, seed=28,
                                       lr_unit_size=128, sb_size=128, superres_denom=15)),
]
IDS = [c[0] for c in SR_CASES]


@pytest.fixture(scope="module")
def ctx():
    c = api.Context()
    yield c
    c.close()


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name,kw", SR_CASES, ids=IDS)
def test_superres_stages(ctx, name, kw):
    """rvpf_upscale of the oracle's CDEF output, then rvpf_lr on the
    oracle's upscaled deblocked and CDEF frames."""
    _say("case", name)
    f = common.make_frame(**kw)
    assert f.params.use_superres and f.params.frame_width < f.params.upscaled_width
    want_deb = pyoracle.loop_filter(f, [None if a is None else a.copy() for a in f.rec])
    want_cdef = pyoracle.cdef(f, want_deb)
    up_cdef = pyoracle.upscale(f, want_cdef)
    up_deb = pyoracle.upscale(f, want_deb)
    df = api.DeviceFrame(f)
    out = df.alloc_ref()
    api.upscale(ctx, df, df.upload_planes(want_cdef), out)
    ctx.sync()
    common.assert_planes(api.to_host(out), up_cdef, f, "crop", name + " upscale")
    want_lr = pyoracle.lr(f, up_deb, up_cdef)
    out2 = df.alloc_ref()
    api.lr(ctx, df, df.upload_planes(up_deb), df.upload_planes(up_cdef), out2)
    ctx.sync()
    common.assert_planes(api.to_host(out2), want_lr, f, "ref", name + " lr")


@pytest.mark.parametrize("name,kw", SR_CASES, ids=IDS)
def test_superres_chain(ctx, name, kw):
    _say("case", name)
    f = common.make_frame(**kw)
    deb, ref, disp = pyoracle.postfilter_frame(f, threads=4)
    df = api.DeviceFrame(f)
    rec = df.upload_planes(f.rec)
    ref_dev = df.alloc_ref()
    disp_dev = df.alloc_display()
    api.postfilter_frame(ctx, df, rec, ref_dev, disp_dev)
    ctx.sync()
    common.assert_recon_kept(api.to_host(rec), f, name)
    common.assert_planes(api.to_host(ref_dev), ref, f, "ref", name + " reference")
    common.assert_planes(api.to_host(disp_dev), disp, f, "crop", name + " display")


def test_superres_chain_u8_planes(ctx):
    """An 8-bit super-resolution frame in 8-bit sample planes, twice on one
    context (the second call reuses the scratch planes)."""
    name, kw = SR_CASES[1]
    _say("case u8", name)
    f = common.make_frame(**kw)
    deb, ref, disp = pyoracle.postfilter_frame(f, threads=4)
    df = api.DeviceFrame(f, bytes_per_sample=1)
    for rep in range(2):
        rec = df.upload_planes(f.rec)
        ref_dev = df.alloc_ref()
        disp_dev = df.alloc_display()
        api.postfilter_frame(ctx, df, rec, ref_dev, disp_dev)
        ctx.sync()
        common.assert_planes(api.to_host(ref_dev), ref, f, "ref", "%s reference (call %d)" % (name, rep))
        common.assert_planes(api.to_host(disp_dev), disp, f, "crop", "%s display (call %d)" % (name, rep))


def test_superres_rejections(ctx):
    """Planes too small for the upscaled frame, rvpf_upscale without
    use_superres, and super-resolution frames in a batch."""
    f = common.make_frame(**SR_CASES[1][1])
    df = api.DeviceFrame(f)
    rec = df.upload_planes(f.rec)
    small = df.alloc_like_recon()  # coded width: narrower than the upscaled frame
    assert small[0].shape[1] < f.params.upscaled_width
    with pytest.raises(api.RvpfError) as e:
        api.postfilter_frame(ctx, df, rec, small, None)
    assert e.value.code == -1
    g = common.make_frame(width=96, height=64, bit_depth=10, seed=3)
    dg = api.DeviceFrame(g)
    with pytest.raises(api.RvpfError) as e:
        api.upscale(ctx, dg, dg.upload_planes(g.rec), dg.alloc_display())
    assert e.value.code == -1
    with pytest.raises(api.RvpfError) as e:
        api.Batch(ctx, [(df, rec, df.alloc_ref(), df.alloc_display())])
    assert e.value.code == -2
    ctx.sync()
