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


@pytest.mark.parametrize("name,kw", common.SMALL_CASES, ids=[c[0] for c in common.SMALL_CASES])
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
    want_disp = [None if p >= f.num_planes else
                 np.ascontiguousarray(want_lr[p][:f.crop_shape(p)[0], :f.crop_shape(p)[1]]) for p in range(3)]
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
    common.assert_planes(api.to_host(rec), deb, f, "full", name + " deblocked")
    common.assert_planes(api.to_host(ref_dev), ref, f, "mi", name + " reference")
    common.assert_planes(api.to_host(disp_dev), disp, f, "crop", name + " display")


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
    common.assert_planes(api.to_host(rec), deb, f, "full", name + " deblocked")
    common.assert_planes(api.to_host(ref_dev), ref, f, "mi", name + " reference")
    common.assert_planes(api.to_host(disp_dev), disp, f, "crop", name + " display")
