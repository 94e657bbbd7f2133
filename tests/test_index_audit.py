"""Row and index arithmetic of the kernels, replayed on the host.

Row offsets in the kernels are signed 64-bit products (row * stride). One
site uses row deltas that are negative by design: the horizontal deblocking
pass writes the rows a filter modifies relative to the edge row,
(k - H) * stride for k - H in [-R, R - 1] (lf.hip, lf_class: R = 2 for the
4- and 6-sample filters, 3 for the 8-sample one, 6 for the 16-sample one).
An unsigned 24-bit product of the masked delta (what v_mul_u32_u24 computes)
turns each negative delta into a positive offset of about 2^24 rows, which
is far outside any plane: a GPU memory fault on the first horizontal pass.

The replay below checks, for every edge position the kernel filters, that
the rows and columns a filter of each size writes stay inside the plane
(the edge position is at least 4 — the frame's top and left borders are not
filtered — and a 16-sample filter closer than 8 samples to the start is
reduced to 8 samples), and that the signed delta arithmetic, unlike the
masked 24-bit one, lands on those rows. It also checks the division
shortcuts that were tried with the 24-bit change and taken out with it:
they are exact over their domains, so they did not cause the fault.
"""
import numpy as np

# rows (or columns) a filter of each size modifies on either side of the edge
REACH = {4: 2, 6: 2, 8: 3, 16: 6}


def kernel_filter_size(size, pos):
    """The size edge_decision keeps at edge position pos (lf.hip)."""
    return 8 if size == 16 and pos < 8 else size


def mul_u24(a, b):
    """v_mul_u32_u24 / v_mul_hi_u32_u24: the 48-bit product of the low 24 bits."""
    return (a & 0xFFFFFF) * (b & 0xFFFFFF)


def test_deblock_write_footprint_stays_in_the_plane():
    margin = 8  # every deblocked plane has 8 rows and columns past the mode-info area
    for area in (8, 12, 64, 68, 1080, 2160):
        for pos in range(4, area, 4):  # filtered edges: inside the area, not at 0
            for size in REACH:
                r = REACH[kernel_filter_size(size, pos)]
                lo, hi = pos - r, pos + r - 1
                assert lo >= 0, (area, pos, size)
                assert hi < area + margin, (area, pos, size)


def test_signed_deltas_and_the_masked_24_bit_product():
    stride = 3872  # the reference's padded 4K luma stride
    for delta in range(-6, 6):
        assert delta * stride == np.int64(delta) * np.int64(stride)
        masked = mul_u24(delta & 0xFFFFFFFF, stride)
        if delta >= 0:
            assert masked == delta * stride
        else:
            # a negative delta becomes (2^24 + delta) rows
            assert masked == ((1 << 24) + delta) * stride
            assert masked // stride > 16_000_000


def test_division_shortcuts_of_the_withdrawn_change_are_exact():
    # loop restoration's wide copy: it / nch through a float reciprocal
    for nch in range(1, 33):
        it = np.arange(0, nch * 300, dtype=np.int64)
        rc = np.float32(1.0) / np.float32(nch)
        assert np.array_equal(np.floor(it.astype(np.float32) * rc).astype(np.int64), it // nch), nch
    # film grain with 64-column tiles: r / segs, t / tw, segs = 8 >> ssx, tw = 64 >> ssx
    r = np.arange(0, 1 << 20)
    for ssx in (0, 1):
        assert np.array_equal(r >> (3 - ssx), r // (8 >> ssx))
        assert np.array_equal(r >> (6 - ssx), r // (64 >> ssx))
