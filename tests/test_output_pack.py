"""The packed picture of Decoder::output (layout of rivet-av1's Frame,
src/frame.rs:131-190) and its MD5 as the AOM vectors' .md5 files hash it
(tests/common/mod.rs:101-119 of the reference): host-side checks."""
import hashlib

import numpy as np
import pytest

from rvpf import abi, api, workload


def pack_numpy(planes, params):
    """The packed bytes of cropped planes: Y, U, V one after another,
    tightly packed, u8 at 8 bits else little-endian u16."""
    dt = np.uint8 if params.bit_depth == 8 else np.dtype("<u2")
    return b"".join(np.ascontiguousarray(a).astype(dt).tobytes() for a in planes if a is not None)


@pytest.mark.parametrize("kw", [
    dict(width=66, height=66, bit_depth=8),
    dict(width=226, height=130, bit_depth=10, ssx=1, ssy=0),
    dict(width=97, height=41, bit_depth=12, ssx=0, ssy=0),
    dict(width=130, height=70, bit_depth=12, mono=True),
    dict(width=200, height=64, bit_depth=10, superres_denom=13),
])
def test_layout_is_frame_new(kw):
    f = workload.Frame(**dict(dict(qindex=100, lr=0, seed=1, with_source=False), **kw))
    lay = api.packed_layout(f.params)
    bps = 1 if kw["bit_depth"] == 8 else 2
    assert lay.bytes_per_sample == bps and lay.num_planes == f.num_planes
    off = 0
    w, h = kw["width"], kw["height"]
    for p in range(f.num_planes):
        sx = (1 if kw.get("ssx", 1) else 0) if p else 0
        sy = (1 if kw.get("ssy", 1) else 0) if p else 0
        pw, ph = (w + sx) >> sx, (h + sy) >> sy
        assert (lay.width[p], lay.height[p], lay.offset[p]) == (pw, ph, off)
        off += pw * ph * bps
    assert lay.bytes == off


def test_layout_rejects_bad_params():
    f = workload.Frame(96, 64, bit_depth=10, qindex=100, lr=0, seed=1, with_source=False)
    f.params.bit_depth = 9
    with pytest.raises(api.RvpfError):
        api.packed_layout(f.params)


def test_md5_known_answers():
    """RFC 1321's test suite pins the digest itself."""
    p = abi.RvpfFrameParams()
    p.mono_chrome = 0
    assert api.frame_md5(b"", p) == "d41d8cd98f00b204e9800998ecf8427e"
    assert api.frame_md5(b"abc", p) == "900150983cd24fb0d6963f7d28e17f72"
    assert api.frame_md5(b"message digest", p) == "f96b697d7cb7938d525a2f31aaf161d0"


@pytest.mark.parametrize("bd", [8, 10])
def test_md5_of_monochrome_appends_neutral_chroma(bd):
    f = workload.Frame(67, 33, bit_depth=bd, mono=True, qindex=100, lr=0, seed=1, with_source=False)
    rng = np.random.default_rng(bd)
    y = rng.integers(0, 1 << bd, size=(33, 67), dtype=np.uint16)
    packed = pack_numpy([y], f.params)
    mid = np.full((17, 34), 1 << (bd - 1), dtype=np.uint8 if bd == 8 else np.dtype("<u2"))
    want = hashlib.md5(packed + mid.tobytes() + mid.tobytes()).hexdigest()
    assert api.frame_md5(packed, f.params) == want
    f.params.mono_chrome = 0
    assert api.frame_md5(packed, f.params) == hashlib.md5(packed).hexdigest()
