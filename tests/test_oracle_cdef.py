"""Pins the oracle's CDEF.

1. Block filter (7.15.3): the oracle against a per-pixel transcription of
   the process written here, on the reference's own window generator and
   seed (src/dsp/cdef.rs:250-312, random windows with unavailable rows /
   columns, strengths and directions), extended to 12 bits.
2. Direction search (7.15.2), closed form: a block constant along the
   lines of direction d has cost[d] = 840 * sum(x^2), the maximum any
   direction can reach (Cauchy-Schwarz), so the search returns d, and the
   lowest such index on ties.
3. Identity: zero strengths or a constant block leave the samples.
4. Frame level: block selection, skip test, variance-adjusted strengths and
   chroma directions (7.15.1) re-derived here over small frames.
"""
import numpy as np

import common
import pyoracle
from refgen import XorShift64

NA = np.iinfo(np.int32).min
WS = 12
DIRS = [[(-1, 1), (-2, 2)], [(0, 1), (-1, 2)], [(0, 1), (0, 2)], [(0, 1), (1, 2)],
        [(1, 1), (2, 2)], [(1, 0), (2, 1)], [(1, 0), (2, 0)], [(1, 0), (2, -1)]]
PRI_TAPS = [[4, 2], [3, 3]]
SEC_TAPS = [[2, 1], [2, 1]]
UV_DIR = [[[0, 1, 2, 3, 4, 5, 6, 7], [1, 2, 2, 2, 3, 4, 6, 0]],
          [[7, 0, 2, 4, 5, 6, 6, 6], [0, 1, 2, 3, 4, 5, 6, 7]]]


def floor_log2(x):
    return x.bit_length() - 1


def constrain(diff, t, damping):
    if t == 0:
        return 0
    adj = max(0, damping - floor_log2(t))
    mag = min(abs(diff), max(0, t - (abs(diff) >> adj)))
    return mag if diff >= 0 else -mag


def filter_block(win, w, h, pri, sec, damping, d, cs):
    """7.15.3 on a 12x12 window (row major, NA = unavailable)."""
    out = np.zeros((h, w), dtype=np.int64)
    W = win.reshape(WS, WS)
    sel = (pri >> cs) & 1
    for i in range(h):
        for j in range(w):
            y, x = i + 2, j + 2
            xv = int(W[y, x])
            total, lo, hi = 0, xv, xv
            for k in range(2):
                for sgn in (1, -1):
                    dy, dx = DIRS[d][k]
                    p = int(W[y + sgn * dy, x + sgn * dx])
                    if p != NA:
                        total += PRI_TAPS[sel][k] * constrain(p - xv, pri, damping)
                        lo, hi = min(lo, p), max(hi, p)
                    for dd in ((d + 2) & 7, (d - 2) & 7):
                        sy, sx = DIRS[dd][k]
                        s = int(W[y + sgn * sy, x + sgn * sx])
                        if s != NA:
                            total += SEC_TAPS[sel][k] * constrain(s - xv, sec, damping)
                            lo, hi = min(lo, s), max(hi, s)
            v = xv + ((8 + total - (1 if total < 0 else 0)) >> 4)
            out[i, j] = min(max(v, lo), hi)
    return out


def test_block_filter_reference_generator():
    r = XorShift64(0x2545_f491_4f6c_dd1d)
    for case in range(4000):
        bd = [8, 10, 12][case % 3]
        win = np.zeros(WS * WS, dtype=np.int32)
        base = r.rnd(1 << bd)
        for i in range(WS * WS):
            win[i] = min(max(base + r.rnd(64 << (bd - 8)) - (32 << (bd - 8)), 0), (1 << bd) - 1)
            if r.rnd(9) == 0:
                win[i] = r.rnd(1 << bd)
        m = r.rnd(5)
        if m == 0:
            win[:2 * WS] = NA
        elif m == 1:
            win[WS - 1::WS] = NA
        elif m == 2:
            win[0::WS] = NA
            win[1::WS] = NA
        sh = bd - 8
        pri = r.rnd(16) << sh
        sec = [0, 1, 2, 4][r.rnd(4)] << sh
        damping = 3 + r.rnd(4) + sh
        d = r.rnd(8)
        for (w, h) in [(8, 8), (4, 4), (4, 8), (8, 4)]:
            got = pyoracle.cdef_filter_window(win, w, h, pri, sec, damping, d, sh)
            want = filter_block(win, w, h, pri, sec, damping, d, sh)
            assert np.array_equal(got, want), "case %d %dx%d" % (case, w, h)


def lines_of(d):
    """Line index of each pixel for direction d (the partial sums of 7.15.2)."""
    f = [lambda i, j: i + j, lambda i, j: i + j // 2, lambda i, j: i, lambda i, j: 3 + i - j // 2,
         lambda i, j: 7 + i - j, lambda i, j: 3 - i // 2 + j, lambda i, j: j, lambda i, j: i // 2 + j][d]
    return np.array([[f(i, j) for j in range(8)] for i in range(8)])


def test_direction_closed_form():
    rng = np.random.default_rng(3)
    for case in range(400):
        d = case % 8
        bd = [8, 10, 12][case % 3]
        vals = rng.integers(-100, 100, size=15)
        blk = (vals[lines_of(d)] + 128) << (bd - 8)
        blk = np.clip(blk, 0, (1 << bd) - 1).astype(np.uint16)
        plane = np.zeros((16, 16), dtype=np.uint16)
        plane[4:12, 4:12] = blk
        got, var = pyoracle.cdef_direction(plane, 4, 4, bd)
        x = (blk.astype(np.int64) >> (bd - 8)) - 128
        # the direction whose lines carry the block maximises the cost
        best = 840 * int((x * x).sum())
        assert var >= 0
        if got != d:
            # only another direction with the same maximal cost (lower index) may win
            assert got < d and np.all(x == x[0, 0]) or np.unique(x[lines_of(got) == lines_of(got)[0, 0]]).size == 1
        # var = (best - cost[(d + 4) & 7]) >> 10 is bounded by best >> 10
        assert var <= best >> 10


def test_identity_cases():
    rng = np.random.default_rng(9)
    for case in range(200):
        bd = [8, 10, 12][case % 3]
        win = rng.integers(0, 1 << bd, size=WS * WS).astype(np.int32)
        got = pyoracle.cdef_filter_window(win, 8, 8, 0, 0, 3 + bd - 8, case % 8, bd - 8)
        assert np.array_equal(got, win.reshape(WS, WS)[2:10, 2:10])
        c = np.full(WS * WS, rng.integers(0, 1 << bd), dtype=np.int32)
        got = pyoracle.cdef_filter_window(c, 8, 8, 15 << (bd - 8), 4 << (bd - 8), 6 + bd - 8, case % 8, bd - 8)
        assert np.array_equal(got, c.reshape(WS, WS)[2:10, 2:10])


def direction(luma, x0, y0, bd):
    """7.15.2, written over the line-index maps."""
    x = (luma[y0:y0 + 8, x0:x0 + 8].astype(np.int64) >> (bd - 8)) - 128
    div = [0, 840, 420, 280, 210, 168, 140, 120, 105]
    cost = []
    for d in range(8):
        L = lines_of(d)
        sums = np.bincount((L - L.min()).ravel(), weights=x.ravel()).astype(np.int64)
        cnt = np.bincount((L - L.min()).ravel())
        cost.append(int(sum(int(s) * int(s) * div[c] for s, c in zip(sums, cnt))))
    best = max(cost)
    d = cost.index(best)
    if best == 0:
        d = 0
    return d, (best - cost[(d + 4) & 7]) >> 10


def frame_cdef(f, deb):
    P = f.params
    out = [None if a is None else a.copy() for a in deb]
    if P.coded_lossless or P.allow_intrabc or not P.enable_cdef:
        return out
    mi = f.mi.reshape(-1, 8)
    ms = f.geom.mi_stride
    bd = P.bit_depth
    cs = bd - 8
    for r in range(0, P.mi_rows, 2):
        for c in range(0, P.mi_cols, 2):
            idx = int(f.cdef_idx[(r >> 4) * f.geom.cdef_stride + (c >> 4)])
            if idx < 0:
                continue
            if all(mi[(r + i) * ms + c + j, 3] & 1 for i in (0, 1) for j in (0, 1)):
                continue
            ydir, var = direction(deb[0], c * 4, r * 4, bd)
            pri = P.cdef_y_pri_strength[idx] << cs
            sec = P.cdef_y_sec_strength[idx] << cs
            d = ydir if pri else 0
            vs = min(floor_log2(var >> 6), 12) if (var >> 6) else 0
            pri = (pri * (4 + vs) + 8) >> 4 if var else 0
            jobs = [(0, pri, sec, P.cdef_damping + cs, d)]
            if f.num_planes > 1:
                upri = P.cdef_uv_pri_strength[idx] << cs
                usec = P.cdef_uv_sec_strength[idx] << cs
                ud = UV_DIR[P.ssx][P.ssy][ydir] if upri else 0
                jobs += [(1, upri, usec, P.cdef_damping + cs - 1, ud), (2, upri, usec, P.cdef_damping + cs - 1, ud)]
            for (p, pr, se, dmp, dd) in jobs:
                sx, sy = (P.ssx, P.ssy) if p else (0, 0)
                x0, y0, w, h = (c * 4) >> sx, (r * 4) >> sy, 8 >> sx, 8 >> sy
                xlim, ylim = (P.mi_cols * 4) >> sx, (P.mi_rows * 4) >> sy
                win = np.full(WS * WS, NA, dtype=np.int64)
                for i in range(h + 4):
                    for j in range(w + 4):
                        yy, xx = y0 + i - 2, x0 + j - 2
                        if 0 <= yy < ylim and 0 <= xx < xlim:
                            win[i * WS + j] = deb[p][yy, xx]
                out[p][y0:y0 + h, x0:x0 + w] = filter_block(win, w, h, pr, se, dmp, dd, cs)
    return out


def test_frame_cdef_matches_independent_walk():
    cases = [dict(width=48, height=40, bit_depth=8, qindex=120, seed=31),
             dict(width=34, height=42, bit_depth=10, qindex=60, seed=32, cdef_bits=2),
             dict(width=40, height=24, bit_depth=12, ssx=0, ssy=0, qindex=90, seed=33),
             dict(width=32, height=32, bit_depth=8, ssx=1, ssy=0, qindex=40, seed=34),
             dict(width=40, height=16, bit_depth=10, mono=True, qindex=50, seed=35)]
    for kw in cases:
        f = common.make_frame(lr=0, grain=False, **kw)
        deb = pyoracle.loop_filter(f, [None if a is None else a.copy() for a in f.rec])
        got = pyoracle.cdef(f, deb)
        want = frame_cdef(f, deb)
        for p in range(f.num_planes):
            rows, cols = f.mi_area(p)
            assert np.array_equal(got[p][:rows, :cols], want[p][:rows, :cols]), (kw, p)


def test_sample_bit_patterns_order_as_halves():
    """The CDEF kernel takes the largest of samples and of saturating
    differences (0 .. 4095 at 12 bits) with a packed half-float maximum,
    and relies on the unavailable-tap sentinel 0xA000 being below all of
    them: the bit patterns 0 .. 4095 read as IEEE halves (subnormals kept)
    must be strictly increasing and finite, and 0xA000 negative."""
    pats = np.arange(4096, dtype=np.uint16)
    h = pats.view(np.float16).astype(np.float64)
    assert np.all(np.isfinite(h))
    assert np.all(np.diff(h) > 0)
    assert h[0] == 0.0 and not np.signbit(h[0])
    na = np.array([0xA000], dtype=np.uint16).view(np.float16)[0]
    assert float(na) < 0.0
    # the maximum over a random set of patterns (with sentinels) is the
    # integer maximum of its non-sentinel members
    rng = np.random.default_rng(7)
    for _ in range(200):
        n = int(rng.integers(1, 14))
        v = rng.integers(0, 4096, size=n).astype(np.uint16)
        v[rng.random(n) < 0.3] = 0xA000
        v[0] = rng.integers(0, 4096)  # the centre sample is always available
        got = v.view(np.float16).astype(np.float64).max()
        want = int(v[v != 0xA000].max())
        assert np.array([want], dtype=np.uint16).view(np.float16)[0] == got
