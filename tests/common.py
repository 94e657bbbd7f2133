"""Shared helpers of the test suite: synthetic frames, the oracle chain,
the GPU chain, and parity comparisons."""
import numpy as np

from rvpf import workload

# (name, Frame kwargs) -- small cases the oracle runs in well under a second
SMALL_CASES = [
    ("c1_352x288_8b_420_wiener", dict(width=352, height=288, bit_depth=8, qindex=128, lr=1, seed=11)),
    ("66x66_8b_420_all", dict(width=66, height=66, bit_depth=8, qindex=100, lr=2, grain=True, seed=3)),
    ("226x226_10b_420_all", dict(width=226, height=226, bit_depth=10, qindex=160, lr=2, grain=True,
                                 lf_extras=True, seed=4)),
    ("200x120_12b_420_all", dict(width=200, height=120, bit_depth=12, qindex=200, lr=2, grain=True,
                                 lf_extras=True, seed=5, sb_size=128)),
    ("170x90_10b_444_all", dict(width=170, height=90, bit_depth=10, ssx=0, ssy=0, qindex=64, lr=2,
                                grain=True, seed=6, lr_unit_size=64)),
    ("150x100_8b_422_all", dict(width=150, height=100, bit_depth=8, ssx=1, ssy=0, qindex=180, lr=2,
                                grain=True, seed=7, lf_extras=True)),
    ("130x70_12b_mono_all", dict(width=130, height=70, bit_depth=12, mono=True, qindex=120, lr=2,
                                 grain=True, seed=8)),
    ("257x129_10b_420_inter", dict(width=257, height=129, bit_depth=10, qindex=90, lr=2, grain=True,
                                   frame_index=3, pan_px=1.5, seed=9, lr_unit_size=128, sb_size=128)),
    ("96x64_8b_420_sharp", dict(width=96, height=64, bit_depth=8, qindex=250, lr=2, lf_extras=True,
                                grain=True, grain_lag=0, seed=10, lr_unit_size=256)),
    ("64x40_10b_444_lag1", dict(width=64, height=40, bit_depth=10, ssx=0, ssy=0, qindex=30, lr=2,
                                grain=True, grain_lag=1, seed=12, mc_identity=True)),
]

# The frame sizes of the reference's odd_sizes test (tests/encode.rs:119-126:
# 1x1, 7x5, 66x66, 130x34, 16x200, decoded with every post-filter on; 66x66
# is in SMALL_CASES), plus 2x3 and 9x1: a frame smaller than one 8x8 block, a
# stripe one row high, one or two samples per chroma row.
ODD_SIZES = [(1, 1), (7, 5), (130, 34), (16, 200), (2, 3), (9, 1)]
ODD_LAYOUTS = [("8b_420", dict(bit_depth=8)), ("10b_444", dict(bit_depth=10, ssx=0, ssy=0)),
               ("12b_422", dict(bit_depth=12, ssx=1, ssy=0)), ("8b_mono", dict(bit_depth=8, mono=True))]
ODD_CASES = [("%dx%d_%s" % (w, h, name), dict(width=w, height=h, qindex=60, lr=2, grain=True, lf_extras=True,
                                              seed=7 * w + h, **kw))
             for (w, h) in ODD_SIZES for name, kw in ODD_LAYOUTS]

# The widest and tallest frames the synthesiser makes (16384 samples): rows
# of 16384 samples (512 noise blocks per row of film grain blocks), 257
# restoration stripes, and an 8200-wide 4:2:2 frame with 256-sample units.
EXTREME_CASES = [
    ("16384x64_10b_420", dict(width=16384, height=64, bit_depth=10, qindex=120, lr=2, grain=True, lf_extras=True,
                              seed=50)),
    ("64x16384_8b_444_sb128", dict(width=64, height=16384, bit_depth=8, ssx=0, ssy=0, qindex=140, lr=2, grain=True,
                                   seed=51, sb_size=128)),
    ("8200x136_12b_422_u256", dict(width=8200, height=136, bit_depth=12, ssx=1, ssy=0, qindex=90, lr=2, grain=True,
                                   seed=52, lr_unit_size=256)),
]


def make_frame(**kw):
    return workload.Frame(**kw)


def region_equal(a, b, rows, cols):
    return np.array_equal(a[:rows, :cols], b[:rows, :cols])


def first_diff(a, b, rows, cols):
    d = np.argwhere(a[:rows, :cols] != b[:rows, :cols])
    if len(d) == 0:
        return None
    y, x = d[0]
    return "n=%d first at (y=%d, x=%d): got %d want %d" % (len(d), y, x, a[y, x], b[y, x])


def assert_recon_kept(got, frame, what):
    """The whole chain deblocks into context scratch: `recon` keeps its samples."""
    for p in range(frame.num_planes):
        msg = first_diff(got[p], frame.rec[p], *frame.rec[p].shape)
        assert msg is None, "%s plane %d: recon modified: %s" % (what, p, msg)


def assert_planes(got, want, frame, region, what):
    for p in range(frame.num_planes):
        if region == "full":
            rows, cols = want[p].shape
        elif region == "mi":
            rows, cols = frame.mi_area(p)
        elif region == "ref":
            rows, cols = frame.ref_area(p)
        else:
            rows, cols = frame.crop_shape(p)
        msg = first_diff(got[p], want[p], rows, cols)
        assert msg is None, "%s plane %d differs: %s" % (what, p, msg)
