"""GPU parity of the packed output picture (rvpf_output_packed: crop, film
grain and packing of Decoder::output) and of byte display planes from
16-bit planes, against the oracle's display planes packed on the host, byte
for byte and by MD5."""
import sys

import numpy as np
import pytest

import common
import pyoracle
from rvpf import api
from test_output_pack import pack_numpy

pytestmark = pytest.mark.gpu

CASES = [
    ("66x66_8b_420_grain", dict(width=66, height=66, bit_depth=8, qindex=100, lr=2, grain=True, seed=3), 2),
    ("66x66_8b_420_grain_u8", dict(width=66, height=66, bit_depth=8, qindex=100, lr=2, grain=True, seed=3), 1),
    ("150x100_8b_422_nograin", dict(width=150, height=100, bit_depth=8, ssx=1, ssy=0, qindex=180, lr=2, seed=7), 2),
    ("226x130_10b_420_grain", dict(width=226, height=130, bit_depth=10, qindex=160, lr=2, grain=True, seed=4), 2),
    ("170x90_10b_444_grain", dict(width=170, height=90, bit_depth=10, ssx=0, ssy=0, qindex=64, lr=2, grain=True,
                                  seed=6), 2),
    ("130x70_12b_mono_grain", dict(width=130, height=70, bit_depth=12, mono=True, qindex=120, lr=2, grain=True,
                                   seed=8), 2),
    ("256x64_10b_420_aligned", dict(width=256, height=64, bit_depth=10, qindex=90, lr=2, grain=True, seed=13), 2),
    ("256x64_8b_420_aligned", dict(width=256, height=64, bit_depth=8, qindex=90, lr=2, grain=True, seed=14), 2),
    ("200x96_8b_420_superres", dict(width=200, height=96, bit_depth=8, qindex=120, lr=2, grain=True, seed=22,
                                    superres_denom=16), 2),
]


@pytest.fixture(scope="module")
def ctx():
    c = api.Context()
    yield c
    c.close()


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name,kw,bps", CASES, ids=[c[0] for c in CASES])
def test_output_packed(ctx, name, kw, bps):
    _say("case", name)
    f = common.make_frame(**kw)
    deb, ref, disp = pyoracle.postfilter_frame(f, threads=4)
    want = pack_numpy(disp, f.params)
    df = api.DeviceFrame(f, bytes_per_sample=bps)
    rec = df.upload_planes(f.rec)
    ref_dev = df.alloc_ref()
    api.postfilter_frame(ctx, df, rec, ref_dev, None)
    packed = api.alloc_packed(df)
    assert packed.numel() == len(want)
    api.output_packed(ctx, df, ref_dev, packed)
    ctx.sync()
    got = packed.cpu().numpy().tobytes()
    if got != want:
        a, b = np.frombuffer(got, np.uint8), np.frombuffer(want, np.uint8)
        i = int(np.nonzero(a != b)[0][0])
        raise AssertionError("%s: first differing byte %d of %d (got %d, want %d)" % (name, i, len(want), a[i], b[i]))
    assert api.frame_md5(got, f.params) == api.frame_md5(want, f.params)


def test_byte_display_from_16bit_planes(ctx):
    """An 8-bit frame in 16-bit planes with byte display planes, through
    rvpf_grain and the chain."""
    name, kw, _ = CASES[0]
    _say("case byte display", name)
    f = common.make_frame(**kw)
    deb, ref, disp = pyoracle.postfilter_frame(f, threads=4)
    import torch
    df = api.DeviceFrame(f)
    rec = df.upload_planes(f.rec)
    ref_dev = df.alloc_ref()
    disp8 = [torch.zeros(f.crop_shape(p), dtype=torch.uint8, device=df.device) for p in range(f.num_planes)]
    api.postfilter_frame(ctx, df, rec, ref_dev, disp8, display_bps=1)
    ctx.sync()
    common.assert_planes(api.to_host(disp8), disp, f, "crop", name + " chain display")
    disp8b = [torch.zeros(f.crop_shape(p), dtype=torch.uint8, device=df.device) for p in range(f.num_planes)]
    api.grain(ctx, df, ref_dev, disp8b, display_bps=1)
    ctx.sync()
    common.assert_planes(api.to_host(disp8b), disp, f, "crop", name + " grain display")


def test_output_packed_rejections(ctx):
    f = common.make_frame(width=96, height=64, bit_depth=10, qindex=100, lr=0, seed=2)
    df = api.DeviceFrame(f)
    ref_dev = df.upload_planes(f.rec)
    packed = api.alloc_packed(df)
    with pytest.raises(api.RvpfError) as e:
        api.output_packed(ctx, df, ref_dev, packed[:-2])
    assert e.value.code == -1
    with pytest.raises(api.RvpfError) as e:
        api.output_packed(ctx, df, ref_dev, packed[1:-1])  # odd address for 16-bit samples
    assert e.value.code == -1
    ctx.sync()
