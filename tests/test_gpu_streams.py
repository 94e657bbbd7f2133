"""One context, several streams: rvpf_postfilter_frame calls enqueued on two
streams in turn share the context's scratch planes (CDEF output for the
restored planes, the upscaled frames with super-resolution) and its film
grain slots. Every frame must still equal the oracle's chain, as the
reference's post-filters do, which are pure functions over their own
buffers (src/decoder/postfilter.rs:273, :608)."""
import sys

import pytest

import common
import pyoracle
from rvpf import api

pytestmark = pytest.mark.gpu

# frames of different sizes so the scratch layouts differ between calls
FRAMES = [
    dict(width=640, height=360, bit_depth=10, qindex=120, lr=2, grain=True, seed=71, sb_size=128),
    dict(width=352, height=288, bit_depth=8, qindex=160, lr=2, grain=True, seed=72),
    dict(width=480, height=272, bit_depth=10, ssx=0, ssy=0, qindex=90, lr=2, grain=True, seed=73),
    dict(width=640, height=360, bit_depth=12, qindex=200, lr=2, grain=True, seed=74, frame_index=2, pan_px=1.5),
    dict(width=400, height=240, bit_depth=10, qindex=140, lr=2, grain=True, seed=75, superres_denom=12),
    dict(width=320, height=200, bit_depth=8, qindex=60, lr=2, grain=True, seed=76, superres_denom=16),
]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def test_two_streams_one_context():
    import torch
    ctx = api.Context()
    try:
        streams = [torch.cuda.Stream(), torch.cuda.Stream()]
        jobs = []
        for i, kw in enumerate(FRAMES):
            f = common.make_frame(**kw)
            df = api.DeviceFrame(f)
            rec = df.upload_planes(f.rec)
            ref = df.alloc_ref()
            disp = df.alloc_display()
            want = pyoracle.postfilter_frame(f, threads=4)
            jobs.append((kw, f, df, rec, ref, disp, want))
        torch.cuda.synchronize()
        # two rounds over the frames, alternating streams, nothing waited for in between
        for rnd in range(2):
            _say("round", rnd)
            for i, (kw, f, df, rec, ref, disp, _) in enumerate(jobs):
                st = streams[(i + rnd) % 2]
                if rnd:
                    with torch.cuda.stream(st):
                        for t, a in zip(rec, df.upload_planes(f.rec)):
                            if t is not None:
                                t.copy_(a)
                api.postfilter_frame(ctx, df, rec, ref, disp, st)
            for st in streams:
                st.synchronize()
            for kw, f, df, rec, ref, disp, (deb_w, ref_w, disp_w) in jobs:
                name = "%dx%d %d-bit round %d" % (kw["width"], kw["height"], kw["bit_depth"], rnd)
                _say("check", name)
                common.assert_recon_kept(api.to_host(rec), f, name)
                common.assert_planes(api.to_host(ref), ref_w, f, "ref", name + " reference")
                common.assert_planes(api.to_host(disp), disp_w, f, "crop", name + " display")
    finally:
        ctx.close()
