"""rvpf_grain_noalloc (the output stage with no device allocation: its
scratch is the library's static device array) against rvpf_grain, the
entry point it is a variant of, and against the oracle: two calls in a
row on one GPU, the second on a smaller frame, so the second runs on the
scratch the first left behind. Frames past the scratch's maximum are
rejected before any work is enqueued."""
import sys

import numpy as np
import pytest

import common
import pyoracle
from rvpf import abi, api

pytestmark = pytest.mark.gpu

# the first larger than the second; every template and layout path in use
CASES = [
    dict(width=1920, height=1080, bit_depth=10, qindex=120, lr=0, grain=True, grain_lag=3, seed=91),
    dict(width=646, height=362, bit_depth=8, ssx=0, ssy=0, qindex=80, lr=0, grain=True, grain_lag=1, seed=92),
]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _random_ref(f, seed):
    rng = np.random.default_rng(seed)
    return [None if p >= f.num_planes else
            rng.integers(0, 1 << f.bit_depth, size=f.crop_shape(p), dtype=np.uint16) for p in range(3)]


def test_noalloc_equals_grain_on_two_calls():
    ctx = api.Context()
    try:
        for i, kw in enumerate(CASES):
            _say("noalloc call", i, "%dx%d" % (kw["width"], kw["height"]))
            f = common.make_frame(**kw)
            assert f.params.film_grain_params_present and f.params.grain.apply_grain
            df = api.DeviceFrame(f)
            host_ref = _random_ref(f, 10 + i)
            ref = df.upload_planes(host_ref)
            want_dev, got_dev = df.alloc_display(), df.alloc_display()
            api.grain(ctx, df, ref, want_dev)
            api.grain_noalloc(df, ref, got_dev)
            ctx.sync()
            want, got = api.to_host(want_dev), api.to_host(got_dev)
            common.assert_planes(got, want, f, "crop", "noalloc call %d vs rvpf_grain" % i)
            oracle = [None if a is None else a.copy() for a in host_ref]
            pyoracle.grain_apply(f.params, f.params.upscaled_width, f.params.frame_height, oracle)
            common.assert_planes(got, oracle, f, "crop", "noalloc call %d vs oracle" % i)
    finally:
        ctx.close()


def test_noalloc_rejects_frames_past_its_scratch():
    import torch
    f = common.make_frame(width=64, height=64, bit_depth=10, lr=0, grain=True, seed=93)
    df = api.DeviceFrame(f)
    p = f.params
    p.frame_width = p.upscaled_width = 8200
    p.mi_cols = 2 * ((8200 + 7) >> 3)
    # planes large enough for the frame, so only the scratch's maximum rejects it
    shapes = [(64, 8200), (32, 4100), (32, 4100)]
    ref = [torch.zeros(s_, dtype=torch.int16, device="cuda") for s_ in shapes]
    disp = [torch.zeros(s_, dtype=torch.int16, device="cuda") for s_ in shapes]
    with pytest.raises(api.RvpfError) as e:
        api.grain_noalloc(df, ref, disp)
    assert e.value.code == abi.RVPF_E_INVALID
    assert "maximum" in str(e.value), str(e.value)
