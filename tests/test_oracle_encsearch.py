"""The oracle's encoder filter searches (orc_cdef_search, orc_lf_search:
src/encoder/cdef.rs:52-253, src/encoder/mod.rs:1460-1490) pinned through the
decoder filters, which are pinned on their own: every per-block error of the
CDEF search equals the error of the decoder's CDEF with that candidate as
the frame's strength, summed over the 8x8 blocks CDEF filters; every error
of the loop filter search equals the error of the decoder's loop filter at
the scaled levels; and the choices are checked against their definitions
by exhaustive evaluation (first best pair, greedy additions, each block's
pick, the cost over the signalled bits)."""
import numpy as np
import pytest

import pyoracle
from rvpf import workload

CASES = [
    dict(width=200, height=136, bit_depth=10, qindex=140, lr=0, seed=31),
    dict(width=130, height=72, bit_depth=8, ssx=0, ssy=0, qindex=200, lr=0, seed=32),
    dict(width=96, height=80, bit_depth=12, mono=True, qindex=90, lr=0, seed=33),
    dict(width=150, height=100, bit_depth=8, ssx=1, ssy=0, qindex=60, lr=0, seed=34, sb_size=128),
]


def frame(kw):
    return workload.Frame(**kw)


def sec_strength(s):
    return 4 if s == 3 else s


def skip_map(f):
    """skip flag per 4x4 (rows, cols)."""
    g = f.geom
    mi = f.mi.reshape(g.mi_alloc_rows, g.mi_stride, 8)
    return mi[:g.mi_rows, :g.mi_cols, 3] & 1


def per_block_sse(f, out, src, plane):
    """Squared error of `out` against `src` inside the frame, per 64x64
    block, counting only the 8x8 blocks CDEF filters (not all four 4x4
    skipped)."""
    g = f.geom
    sx = (f.params.ssx if plane else 0)
    sy = (f.params.ssy if plane else 0)
    fw, fh = (f.params.frame_width + sx) >> sx, (f.params.frame_height + sy) >> sy
    d = out[plane][:fh, :fw].astype(np.int64) - src[plane][:fh, :fw].astype(np.int64)
    sq = d * d
    sk = skip_map(f)
    allskip = sk[0::2, 0::2] & sk[1::2, 0::2] & sk[0::2, 1::2] & sk[1::2, 1::2]  # per 8x8 luma block
    fbr, fbc = (g.mi_rows + 15) // 16, (g.mi_cols + 15) // 16
    res = np.zeros((fbr, fbc), dtype=np.int64)
    bw, bh = 8 >> sx, 8 >> sy
    for r8 in range(allskip.shape[0]):
        for c8 in range(allskip.shape[1]):
            if allskip[r8, c8]:
                continue
            y0, x0 = r8 * bh, c8 * bw
            res[r8 // 8, c8 // 8] += sq[y0:y0 + bh, x0:x0 + bw].sum()
    return res


@pytest.mark.parametrize("kw", CASES)
def test_cdef_tables_are_the_decoder_filter_errors(kw):
    f = frame(kw)
    deb = pyoracle.loop_filter(f, [None if a is None else a.copy() for a in f.rec])
    q = kw["qindex"]
    res, table, ys, uvs = pyoracle.cdef_search(f, deb, f.src, q, 1000.0, False, threads=4)
    ycand, uvcand = pyoracle.cdef_candidates(False)
    g = f.geom
    coded = [(r, c) for r in range((g.mi_rows + 15) // 16) for c in range((g.mi_cols + 15) // 16)
             if f.cdef_idx[r * g.cdef_stride + c] >= 0]
    assert ys.shape == (len(coded), len(ycand))
    damping = min(q >> 6, 3) + 3
    # every luma candidate k, and chroma candidate k (cycled), as strength 0 of the frame
    p = f.params
    p.enable_cdef, p.cdef_damping = 1, damping
    idx0 = f.cdef_idx.copy()
    f.cdef_idx[:] = np.where(idx0 >= 0, 0, -1)
    for k in range(len(ycand)):
        p.cdef_y_pri_strength[0], p.cdef_y_sec_strength[0] = ycand[k][0], sec_strength(ycand[k][1])
        ku = k % len(uvcand)
        p.cdef_uv_pri_strength[0], p.cdef_uv_sec_strength[0] = uvcand[ku][0], sec_strength(uvcand[ku][1])
        out = pyoracle.cdef(f, deb)
        yb = per_block_sse(f, out, f.src, 0)
        assert [int(yb[r, c]) for r, c in coded] == [int(v) for v in ys[:, k]], k
        if f.num_planes > 1:
            ub = per_block_sse(f, out, f.src, 1) + per_block_sse(f, out, f.src, 2)
            assert [int(ub[r, c]) for r, c in coded] == [int(v) for v in uvs[:, ku]], k
    f.cdef_idx[:] = idx0


@pytest.mark.parametrize("kw", CASES)
@pytest.mark.parametrize("lam", [0.0, 3e3, 1e9])
@pytest.mark.parametrize("thorough", [False, True])
def test_cdef_choice_follows_its_definition(kw, lam, thorough):
    f = frame(kw)
    deb = pyoracle.loop_filter(f, [None if a is None else a.copy() for a in f.rec])
    res, table, ys, uvs = pyoracle.cdef_search(f, deb, f.src, kw["qindex"], lam, thorough, threads=4)
    ycand, uvcand = pyoracle.cdef_candidates(thorough)
    assert len(ycand) == (44 if thorough else 28) and len(uvcand) == (24 if thorough else 15)
    assert res.damping_minus_3 == min(kw["qindex"] >> 6, 3)
    nb = ys.shape[0]
    n = 1 << res.bits
    set_ = [(ycand.index((res.y_pri[i], res.y_sec[i])), uvcand.index((res.uv_pri[i], res.uv_sec[i])))
            for i in range(n)]
    pair = ys[:, :, None].astype(np.int64) + uvs[:, None, :].astype(np.int64)  # [nb, ny, nuv]
    # greedy: each pair of the set is, given the ones before it, the first with the least total
    cur = np.full(nb, np.iinfo(np.int64).max, dtype=np.int64)
    for i, (yk, uk) in enumerate(set_):
        tot = np.minimum(cur[:, None, None], pair).sum(axis=0)
        for (a, b) in set_[:i]:
            tot[a, b] = np.iinfo(np.int64).max
        best = np.unravel_index(np.argmin(tot), tot.shape)  # first in (luma, chroma) order
        assert (yk, uk) == (int(best[0]), int(best[1])), i
        cur = np.minimum(cur, pair[:, yk, uk])
    # each block: the first pair of the set with its least error
    g = f.geom
    coded = [(r, c) for r in range((g.mi_rows + 15) // 16) for c in range((g.mi_cols + 15) // 16)
             if f.cdef_idx[r * g.cdef_stride + c] >= 0]
    errs = np.stack([pair[:, a, b] for a, b in set_], axis=1)
    assert [int(table[r * g.cdef_stride + c]) for r, c in coded] == [int(v) for v in np.argmin(errs, axis=1)]
    # the signalled bits: the least total + lambda * bits * blocks over the prefixes of 1, 2, 4, 8 pairs
    total = float(errs.min(axis=1).sum())
    for b in range(res.bits):
        sub = errs[:, :1 << b].min(axis=1).sum()
        assert total + lam * res.bits * nb < float(sub) + lam * b * nb + 1e-6 * max(1.0, float(sub))
    untouched = np.ones_like(table, dtype=bool)
    for r, c in coded:
        untouched[r * g.cdef_stride + c] = False
    assert np.array_equal(table[untouched], f.cdef_idx[untouched])


def test_cdef_search_without_coded_blocks_is_off():
    f = frame(CASES[0])
    f.cdef_idx[:] = -1
    deb = [None if a is None else a.copy() for a in f.rec]
    res, table, ys, uvs = pyoracle.cdef_search(f, deb, f.src, 100, 10.0, False)
    assert res.as_tuple() == (0, 0, [(0, 0)], [(0, 0)])
    assert ys.shape[0] == 0 and np.all(table == -1)


@pytest.mark.parametrize("kw", CASES)
def test_lf_search_errors_are_the_decoder_filter_errors(kw):
    f = frame(dict(kw, lf_extras=True))
    levels, sse = pyoracle.lf_search(f, [None if a is None else a.copy() for a in f.rec], f.src, threads=4)
    start = list(f.params.loop_filter_level)
    assert any(start)
    best = [None, None, None]
    for k, scale in enumerate([0.0, 0.5, 0.75, 1.0, 1.25, 1.5]):
        lv = [min(int(np.floor(l * scale + 0.5)), 63) for l in start]  # half away from zero for l >= 0
        for i in range(4):
            f.params.loop_filter_level[i] = lv[i]
        planes = [None if a is None else a.copy() for a in f.rec]
        if lv[0] or lv[1]:
            planes = pyoracle.loop_filter(f, planes)
        for p in range(f.num_planes):
            sx = f.params.ssx if p else 0
            sy = f.params.ssy if p else 0
            fw, fh = (f.params.frame_width + sx) >> sx, (f.params.frame_height + sy) >> sy
            d = planes[p][:fh, :fw].astype(np.int64) - f.src[p][:fh, :fw].astype(np.int64)
            e = int((d * d).sum())
            assert int(sse[k, p]) == e, (k, p)
            if best[p] is None or e < best[p][0]:
                best[p] = (e, lv[0] if p == 0 else lv[1 + p])
    for i in range(4):
        f.params.loop_filter_level[i] = start[i]
    want = [0, 0, 0, 0] if best[0][1] == 0 else [best[0][1], best[0][1],
                                                 best[1][1] if f.num_planes > 1 else 0,
                                                 best[2][1] if f.num_planes > 1 else 0]
    assert levels == want


def test_lf_search_from_zero_levels_is_off():
    f = frame(CASES[0])
    for i in range(4):
        f.params.loop_filter_level[i] = 0
    levels, sse = pyoracle.lf_search(f, [None if a is None else a.copy() for a in f.rec], f.src)
    assert levels == [0, 0, 0, 0]
    assert np.all(sse == sse[0])
