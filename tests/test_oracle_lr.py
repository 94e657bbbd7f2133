"""Pins the oracle's loop restoration.

1. Closed form: Wiener taps (0, 0, 0) in both directions are the
   identity at 8, 10 and 12 bits; a self-guided set with r1 = 0 and
   w0 = 0 projects to 128 * u, the identity.
2. Frame level: an independent numpy formulation: the stripe and unit of
   every sample from the per-4x4 formulas of the process (stripe =
   (lumaY + 8) / 64, unit row = ((lumaY + 8) >> subY) / unitSize, as the
   specification evaluates them block by block), the source window from
   the stripe rule, Wiener as two vectorised passes, self-guided box sums
   from integral images (the oracle sums each box directly).
"""
import numpy as np

import common
import pyoracle
from rvpf import abi

SGR = [[2, 12, 1, 4], [2, 15, 1, 6], [2, 18, 1, 8], [2, 21, 1, 9], [2, 24, 1, 10], [2, 29, 1, 11],
       [2, 36, 1, 12], [2, 45, 1, 13], [2, 56, 1, 14], [2, 68, 1, 15], [0, 0, 1, 5], [0, 0, 1, 8],
       [0, 0, 1, 11], [0, 0, 1, 14], [2, 30, 0, 0], [2, 75, 0, 0]]


def r2(x, n):
    return x if n == 0 else (x + (1 << (n - 1))) >> n


def count_units(unit, size):
    return max(1, (size + unit // 2) // unit)


def units_of(f, p):
    a = f.lr_units[p].reshape(-1, 12)
    return a


def sgr_pass(win, r, eps, bd, h, w, y0, pass_):
    """A/B on (h+2)x(w+2) positions from integral images, then the weighted
    3x3 filter; returns flt (h x w) as int64."""
    W = win.astype(np.int64)
    S = np.zeros((W.shape[0] + 1, W.shape[1] + 1), dtype=np.int64)
    Q = np.zeros_like(S)
    S[1:, 1:] = W.cumsum(0).cumsum(1)
    Q[1:, 1:] = (W * W).cumsum(0).cumsum(1)
    # position (i, j) of the A/B grid is window (i + 2, j + 2); box rows i+2-r .. i+2+r
    ia = np.arange(h + 2)[:, None] + 2 - r
    ib = ia + 2 * r + 1
    ja = np.arange(w + 2)[None, :] + 2 - r
    jb = ja + 2 * r + 1

    def bs(T):
        return T[ib, jb] - T[ia, jb] - T[ib, ja] + T[ia, ja]

    bsum, bsq = bs(S), bs(Q)
    n = (2 * r + 1) ** 2
    s = ((1 << 20) + n * n * eps // 2) // (n * n * eps)
    one_over_n = ((1 << 12) + n // 2) // n
    a = (bsq + ((1 << (2 * (bd - 8))) >> 1)) >> (2 * (bd - 8))
    d = (bsum + ((1 << (bd - 8)) >> 1)) >> (bd - 8)
    p = np.maximum(a * n - d * d, 0)
    z = np.minimum((p * s + (1 << 19)) >> 20, 255)
    A2 = np.where(z >= 255, 256, np.where(z == 0, 1, ((z << 8) + z // 2) // (z + 1)))
    B2 = ((256 - A2) * bsum * one_over_n + 2048) >> 12
    u = W[3:3 + h, 3:3 + w]
    yabs = y0 + np.arange(h)[:, None]
    def sh(T, dy, dx):
        return T[1 + dy:1 + dy + h, 1 + dx:1 + dx + w]
    if pass_ == 0:
        odd = (yabs & 1) == 1
        a_odd = 6 * sh(A2, 0, 0) + 5 * (sh(A2, 0, -1) + sh(A2, 0, 1))
        b_odd = 6 * sh(B2, 0, 0) + 5 * (sh(B2, 0, -1) + sh(B2, 0, 1))
        a_even = 6 * (sh(A2, -1, 0) + sh(A2, 1, 0)) + 5 * (sh(A2, -1, -1) + sh(A2, -1, 1) + sh(A2, 1, -1) + sh(A2, 1, 1))
        b_even = 6 * (sh(B2, -1, 0) + sh(B2, 1, 0)) + 5 * (sh(B2, -1, -1) + sh(B2, -1, 1) + sh(B2, 1, -1) + sh(B2, 1, 1))
        av = np.where(odd, a_odd, a_even)
        bv = np.where(odd, b_odd, b_even)
        shift = np.where(odd, 4, 5)
    else:
        cross = sh(A2, 0, 0) + sh(A2, 0, -1) + sh(A2, 0, 1) + sh(A2, -1, 0) + sh(A2, 1, 0)
        diag = sh(A2, -1, -1) + sh(A2, -1, 1) + sh(A2, 1, -1) + sh(A2, 1, 1)
        av = 4 * cross + 3 * diag
        crossb = sh(B2, 0, 0) + sh(B2, 0, -1) + sh(B2, 0, 1) + sh(B2, -1, 0) + sh(B2, 1, 0)
        diagb = sh(B2, -1, -1) + sh(B2, -1, 1) + sh(B2, 1, -1) + sh(B2, 1, 1)
        bv = 4 * crossb + 3 * diagb
        shift = np.full((h, w), 5)
    v = av * u + bv
    k = 4 + shift
    return (v + (1 << (k - 1))) >> k


def lr_frame(f, deb, cdef):
    P = f.params
    bd = P.bit_depth
    maxv = (1 << bd) - 1
    out = [None if a is None else a.copy() for a in cdef]
    for p in range(f.num_planes):
        if P.frame_restoration_type[p] == 0:
            continue
        sx, sy = (P.ssx, P.ssy) if p else (0, 0)
        unit = P.loop_restoration_size[p]
        pex, pey = r2(P.upscaled_width, sx) - 1, r2(P.frame_height, sy) - 1
        urows, ucols = count_units(unit, pey + 1), count_units(unit, pex + 1)
        assert (urows, ucols) == (P.lr_unit_rows[p], P.lr_unit_cols[p])
        U = units_of(f, p)
        ys, xs = np.arange(pey + 1), np.arange(pex + 1)
        luma_y = ((ys << sy) >> 2) * 4                     # lumaY of the sample's 4x4 block row
        stripe = (luma_y + 8) // 64
        urow = np.minimum(urows - 1, ((luma_y + 8) >> sy) // unit)
        ucol = np.minimum(ucols - 1, xs // unit)
        for s in np.unique(stripe):
            rows = ys[stripe == s]
            y0, y1 = rows[0], rows[-1]
            assert np.all(urow[stripe == s] == urow[y0])
            sst = (-8 + int(s) * 64) >> sy
            sen = sst + (64 >> sy) - 1
            for uc in range(ucols):
                cols = xs[ucol == uc]
                if cols.size == 0:
                    continue
                x0, x1 = cols[0], cols[-1]
                u = U[urow[y0] * ucols + uc]
                typ = int(u[0])
                if typ not in (1, 2):
                    continue
                h, w = y1 - y0 + 1, x1 - x0 + 1
                wy = np.clip(np.arange(y0 - 3, y1 + 4), 0, pey)
                wx = np.clip(np.arange(x0 - 3, x1 + 4), 0, pex)
                above = wy < sst
                below = wy > sen
                ry = np.where(above, np.maximum(sst - 2, wy), np.where(below, np.minimum(sen + 2, wy), wy))
                win = np.where((above | below)[:, None], deb[p][ry][:, wx], cdef[p][ry][:, wx]).astype(np.int64)
                if typ == 1:
                    c = u[4:10].view(np.int8).astype(np.int64)
                    vt = [c[0], c[1], c[2], 128 - 2 * (c[0] + c[1] + c[2]), c[2], c[1], c[0]]
                    ht = [c[3], c[4], c[5], 128 - 2 * (c[3] + c[4] + c[5]), c[5], c[4], c[3]]
                    r0, r1 = (5, 9) if bd == 12 else (3, 11)
                    off = 1 << (bd + 7 - r0 - 1)
                    lim = (1 << (bd + 1 + 7 - r0)) - 1
                    hs = sum(ht[t] * win[:, t:t + w] for t in range(7))
                    inter = np.clip((hs + (1 << (r0 - 1))) >> r0, -off, lim - off)
                    vs = sum(vt[t] * inter[t:t + h, :] for t in range(7))
                    res = np.clip((vs + (1 << (r1 - 1))) >> r1, 0, maxv)
                else:
                    st = int(u[1])
                    xq = u[2:4].view(np.int8).astype(np.int64)
                    w0, w1 = int(xq[0]), int(xq[1])
                    w2 = 128 - w0 - w1
                    ue = win[3:3 + h, 3:3 + w] << 4
                    f0 = sgr_pass(win, SGR[st][0], SGR[st][1], bd, h, w, y0, 0) if SGR[st][0] else ue
                    f1 = sgr_pass(win, SGR[st][2], SGR[st][3], bd, h, w, y0, 1) if SGR[st][2] else ue
                    v = w1 * ue + w0 * f0 + w2 * f1
                    res = np.clip((v + 1024) >> 11, 0, maxv)
                out[p][y0:y1 + 1, x0:x1 + 1] = res
    return out


def test_frame_lr_matches_independent_formulation():
    cases = [dict(width=130, height=140, bit_depth=8, qindex=100, seed=41, lr_unit_size=64),
             dict(width=100, height=75, bit_depth=10, qindex=160, seed=42, lr_unit_size=64, ssx=0, ssy=0),
             dict(width=96, height=200, bit_depth=12, qindex=60, seed=43, lr_unit_size=128),
             dict(width=70, height=90, bit_depth=10, ssx=1, ssy=0, qindex=130, seed=44, lr_unit_size=64),
             dict(width=150, height=66, bit_depth=8, mono=True, qindex=200, seed=45, lr_unit_size=256),
             dict(width=257, height=129, bit_depth=10, qindex=90, seed=46, lr_unit_size=64)]
    restored = 0
    for kw in cases:
        f = common.make_frame(lr=2, grain=False, **kw)
        deb = pyoracle.loop_filter(f, [None if a is None else a.copy() for a in f.rec])
        cd = pyoracle.cdef(f, deb)
        got = pyoracle.lr(f, deb, cd)
        want = lr_frame(f, deb, cd)
        for p in range(f.num_planes):
            rows, cols = f.crop_shape(p)
            assert np.array_equal(got[p][:rows, :cols], want[p][:rows, :cols]), (kw, p)
            restored += int((got[p][:rows, :cols] != cd[p][:rows, :cols]).sum())
    assert restored > 10000


def test_identity_parameters():
    for bd in (8, 10, 12):
        f = common.make_frame(width=96, height=80, bit_depth=bd, lr=2, grain=False, seed=50 + bd)
        deb = pyoracle.loop_filter(f, [None if a is None else a.copy() for a in f.rec])
        cd = pyoracle.cdef(f, deb)
        for p in range(f.num_planes):
            U = f.lr_units[p].reshape(-1, 12)
            for i in range(U.shape[0]):
                if i % 2 == 0:
                    U[i, 0] = abi.RESTORE_WIENER
                    U[i, 4:10] = 0
                else:
                    U[i, 0] = abi.RESTORE_SGRPROJ
                    U[i, 1] = 14 + (i // 2) % 2          # r1 = 0
                    U[i, 2] = 0                           # w0 = 0
                    U[i, 3] = np.int8(40).view(np.uint8)  # any w1
        got = pyoracle.lr(f, deb, cd)
        for p in range(f.num_planes):
            rows, cols = f.crop_shape(p)
            assert np.array_equal(got[p][:rows, :cols], cd[p][:rows, :cols]), (bd, p)
