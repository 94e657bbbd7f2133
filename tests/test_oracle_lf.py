"""Pins the oracle's loop filter.

1. Sample filter (7.14.6): the oracle's four-lane form against a
   one-sample transcription of the process written here in the
   specification's structure (mask -> narrow or wide filter), on the
   reference's own generator and seed (src/dsp/lf.rs:309-353, 20 000
   random edges at 8/10/12 bits), as the reference tests its AVX2 lanes.
2. Invariants: a constant edge is never changed; a masked edge is left
   as it is; outputs stay in [0, 2^bd - 1].
3. Frame level: edge decisions, filter sizes and strengths (7.14.2-7.14.5)
   re-derived here by an independent per-edge walker over small frames.
"""
import numpy as np

import common
import pyoracle
from refgen import XorShift64


def lf_one(v, limit, blimit, thresh, fs, chroma, bd):
    """7.14.6 for one sample position across the edge (v[7] = q0)."""
    v = list(v)
    sh = bd - 8
    p3, p2, p1, p0, q0, q1, q2, q3 = v[3:11]
    lim, blim, thr = limit << sh, blimit << sh, thresh << sh
    hev = abs(p1 - p0) > thr or abs(q1 - q0) > thr
    flen = 4 if fs == 4 else (6 if chroma else (8 if fs == 8 else 16))
    bad = abs(p1 - p0) > lim or abs(q1 - q0) > lim or abs(p0 - q0) * 2 + abs(p1 - q1) // 2 > blim
    if flen >= 6:
        bad = bad or abs(p2 - p1) > lim or abs(q2 - q1) > lim
    if flen >= 8:
        bad = bad or abs(p3 - p2) > lim or abs(q3 - q2) > lim
    if bad:
        return v
    t = 1 << sh
    flat = False
    if fs >= 8:
        taps = [p1, q1, p2, q2] + ([p3, q3] if flen >= 8 else [])
        flat = all(abs(x - (p0 if i % 2 == 0 else q0)) <= t for i, x in enumerate(taps))
    flat2 = False
    if fs >= 16:
        flat2 = all(abs(v[k] - p0) <= t for k in (0, 1, 2)) and all(abs(v[k] - q0) <= t for k in (11, 12, 13))
    if fs == 4 or not flat:
        lo, hi = -(1 << (bd - 1)), (1 << (bd - 1)) - 1

        def c(x):
            return min(max(x, lo), hi)

        off = 0x80 << sh
        ps1, ps0, qs0, qs1 = p1 - off, p0 - off, q0 - off, q1 - off
        f = c(ps1 - qs1) if hev else 0
        f = c(f + 3 * (qs0 - ps0))
        f1 = c(f + 4) >> 3
        f2 = c(f + 3) >> 3
        v[7] = c(qs0 - f1) + off
        v[6] = c(ps0 + f2) + off
        if not hev:
            f = (f1 + 1) >> 1
            v[8] = c(qs1 - f) + off
            v[5] = c(ps1 + f) + off
        return v
    log2 = 3 if (fs == 8 or not flat2) else 4
    n = 6 if log2 == 4 else (2 if chroma else 3)
    n2 = 0 if (log2 == 3 and not chroma) else 1
    orig = list(v)
    for i in range(-n, n):
        acc = 0
        for j in range(-n, n + 1):
            pos = min(max(i + j, -(n + 1)), n)
            acc += orig[7 + pos] * (2 if abs(j) <= n2 else 1)
        v[7 + i] = (acc + (1 << (log2 - 1))) >> log2
    return v


def test_lanes_match_one_sample_process_reference_generator():
    r = XorShift64(0x0bad_5eed_1234_5678)
    for case in range(20000):
        bd = [8, 10, 12][case % 3]
        limit = 1 + r.rnd(9)
        blimit = r.rnd(130)
        thresh = r.rnd(4)
        fs = min([4, 8, 16, 6][r.rnd(4)], 8 if case % 5 == 0 else 16)
        chroma = case % 5 == 0
        if chroma:
            fs = min(fs, 8)
        base = r.rnd(1 << bd)
        spread = [1, 4, 16, 64][r.rnd(4)] << (bd - 8)
        lanes = np.zeros((14, 4), dtype=np.int32)
        for k in range(14):
            for l in range(4):
                lanes[k, l] = min(max(base + r.rnd(spread) - spread // 2, 0), (1 << bd) - 1)
        want = np.array([lf_one(lanes[:, l], limit, blimit, thresh, fs, chroma, bd) for l in range(4)]).T
        got = pyoracle.lf_filter_lanes(lanes.copy(), limit, blimit, thresh, fs, chroma, bd)
        assert np.array_equal(got, want), "case %d" % case


def test_lane_invariants():
    rng = np.random.default_rng(5)
    for case in range(3000):
        bd = [8, 10, 12][case % 3]
        fs = [4, 8, 16][case % 3] if case % 4 else 8
        chroma = case % 4 == 0
        # constant edge: never changed
        v = np.full((14, 4), rng.integers(0, 1 << bd), dtype=np.int32)
        got = pyoracle.lf_filter_lanes(v.copy(), 1 + case % 9, case % 130, case % 4, fs, chroma, bd)
        assert np.array_equal(got, v)
        # random edge: range kept; samples beyond the filter's reach untouched
        v = rng.integers(0, 1 << bd, size=(14, 4)).astype(np.int32)
        got = pyoracle.lf_filter_lanes(v.copy(), 1 + case % 9, case % 130, case % 4, fs, chroma, bd)
        assert got.min() >= 0 and got.max() < (1 << bd)
        reach = 6 if fs == 16 else (2 if chroma and fs == 8 else (3 if fs == 8 else 2))
        keep = [k for k in range(14) if not (7 - reach <= k < 7 + reach)]
        assert np.array_equal(got[keep], v[keep])
        # blimit 0, limit 1 at a large step: masked, unchanged
        v = np.zeros((14, 4), dtype=np.int32)
        v[7:] = 200 << (bd - 8)
        got = pyoracle.lf_filter_lanes(v.copy(), 1, 0, 0, fs, chroma, bd)
        assert np.array_equal(got, v)


# ---------------------------------------------------------------- frame level
TXW = [4, 8, 16, 32, 64, 4, 8, 8, 16, 16, 32, 32, 64, 4, 16, 8, 32, 16, 64]
TXH = [4, 8, 16, 32, 64, 8, 4, 16, 8, 32, 16, 64, 32, 16, 4, 32, 8, 64, 16]
BW4 = [1, 1, 2, 2, 2, 4, 4, 4, 8, 8, 8, 16, 16, 16, 32, 32, 1, 4, 2, 8, 4, 16]
BH4 = [1, 2, 1, 2, 4, 2, 4, 8, 4, 8, 16, 8, 16, 32, 16, 32, 4, 1, 8, 2, 16, 4]


def block_of(w4, h4):
    for b in range(22):
        if BW4[b] == w4 and BH4[b] == h4:
            return b
    return 22


def mi_fields(f):
    a = f.mi.reshape(-1, 8)
    return dict(size=a[:, 0].astype(int), mode=a[:, 1].astype(int), ref=a[:, 2].view(np.int8).astype(int),
                skip=(a[:, 3] & 1).astype(int), seg=((a[:, 3] >> 1) & 7).astype(int),
                dlf=a[:, 4:8].view(np.int8).astype(int))


def strength(f, mi, idx, plane, pas):
    """7.14.4 / 7.14.5 for the 4x4 at flat index idx."""
    P = f.params
    i = pas if plane == 0 else plane + 1
    d = mi["dlf"][idx][i] if P.delta_lf_multi else mi["dlf"][idx][0]
    lvl = min(max(d + P.loop_filter_level[i], 0), 63)
    seg = mi["seg"][idx]
    if P.segmentation_enabled and P.seg_feature_enabled[seg][1 + i]:
        lvl = min(max(lvl + P.seg_feature_data[seg][1 + i], 0), 63)
    if P.loop_filter_delta_enabled:
        sh = lvl >> 5
        rf = mi["ref"][idx]
        mode = mi["mode"][idx]
        if rf == 0:
            lvl += P.loop_filter_ref_deltas[0] << sh
        elif rf > 0:
            mt = 1 if (mode >= 14 and mode not in (16, 24)) else 0
            lvl += (P.loop_filter_ref_deltas[rf] << sh) + (P.loop_filter_mode_deltas[mt] << sh)
        lvl = min(max(lvl, 0), 63)
    sharp = P.loop_filter_sharpness
    shift = 2 if sharp > 4 else (1 if sharp > 0 else 0)
    limit = min(max(lvl >> shift, 1), 9 - sharp) if sharp > 0 else max(1, lvl >> shift)
    return lvl, limit, 2 * (lvl + 2) + limit, lvl >> 4


def walk_edges(f, planes):
    """7.14.1-7.14.3 over every edge, one sample row at a time."""
    P = f.params
    mi = mi_fields(f)
    ms = f.geom.mi_stride
    SUB = None
    for plane in range(f.num_planes):
        if plane > 0 and P.loop_filter_level[1 + plane] == 0:
            continue
        sx, sy = (P.ssx, P.ssy) if plane else (0, 0)
        buf = planes[plane]
        tx = f.tx[plane]
        for pas in range(2):
            for row in range(0, P.mi_rows, 1 << sy):
                for col in range(0, P.mi_cols, 1 << sx):
                    x, y = col * 4, row * 4
                    if x >= P.frame_width or y >= P.frame_height:
                        continue
                    if (pas == 0 and x == 0) or (pas == 1 and y == 0):
                        continue
                    r, c = row | sy, col | sx
                    idx = r * ms + c
                    pidx = (r - (sy << sy) * 0 - ((1 << sy) if pas else 0)) * ms + (c - ((1 << sx) if pas == 0 else 0))
                    tsz = tx[(r >> sy) * ms + (c >> sx)]
                    ptsz = tx[((r - ((1 << sy) if pas else 0)) >> sy) * ms + ((c - ((1 << sx) if pas == 0 else 0)) >> sx)]
                    psize = SUB_SIZE[mi["size"][idx]][sx][sy]
                    xp, yp = x >> sx, y >> sy
                    if psize == 22:
                        block_edge = True
                    else:
                        block_edge = (xp % (4 * BW4[psize]) == 0) if pas == 0 else (yp % (4 * BH4[psize]) == 0)
                    tx_edge = (xp % TXW[tsz] == 0) if pas == 0 else (yp % TXH[tsz] == 0)
                    if not (tx_edge and (block_edge or not mi["skip"][idx] or mi["ref"][idx] <= 0)):
                        continue
                    base = min(TXW[ptsz], TXW[tsz]) if pas == 0 else min(TXH[ptsz], TXH[tsz])
                    fs = min(16 if plane == 0 else 8, base)
                    st = strength(f, mi, idx, plane, pas)
                    if st[0] == 0:
                        st = strength(f, mi, pidx, plane, pas)
                    if st[0] == 0:
                        continue
                    reach = 7 if fs >= 16 else 4
                    span = 6 if fs >= 16 else 3
                    for l in range(4):
                        line = [0] * 14
                        for k in range(-reach, reach):
                            line[7 + k] = int(buf[yp + l, xp + k] if pas == 0 else buf[yp + k, xp + l])
                        out = lf_one(line, st[1], st[2], st[3], fs, plane > 0, P.bit_depth)
                        for k in range(-span, span):
                            if pas == 0:
                                buf[yp + l, xp + k] = out[7 + k]
                            else:
                                buf[yp + k, xp + l] = out[7 + k]
    return planes


SUB_SIZE = None


def _load_sub_size():
    global SUB_SIZE
    import re
    import os
    src = open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                            "rivet-av1_amd", "csrc", "av1_tables.h")).read()
    m = re.search(r"AV1_SUBSAMPLED_SIZE\[22\]\[2\]\[2\] = (\{.*?\});", src, re.S)
    vals = [int(v) for v in re.findall(r"\d+", m.group(1))]
    SUB_SIZE = np.array(vals).reshape(22, 2, 2)


def test_frame_loop_filter_matches_edge_walker():
    _load_sub_size()
    cases = [dict(width=72, height=40, bit_depth=8, qindex=150, seed=21),
             dict(width=66, height=34, bit_depth=10, qindex=90, lf_extras=True, seed=22),
             dict(width=48, height=56, bit_depth=12, ssx=0, ssy=0, qindex=220, lf_extras=True, seed=23),
             dict(width=40, height=40, bit_depth=8, ssx=1, ssy=0, qindex=200, seed=24, sb_size=128),
             dict(width=64, height=24, bit_depth=10, mono=True, qindex=60, lf_extras=True, seed=25)]
    changed = 0
    for kw in cases:
        f = common.make_frame(lr=0, grain=False, **kw)
        want = walk_edges(f, [None if a is None else a.copy() for a in f.rec])
        got = pyoracle.loop_filter(f, [None if a is None else a.copy() for a in f.rec])
        for p in range(f.num_planes):
            assert np.array_equal(got[p], want[p]), (kw, p)
        changed += sum(int((got[p] != f.rec[p]).sum()) for p in range(f.num_planes))
    assert changed > 1000
