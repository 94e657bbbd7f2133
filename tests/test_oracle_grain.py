"""Pins the oracle's film grain synthesis.

1. Random number process (7.18.3.2): the new bit of every step obeys the
   linear recurrence b[n + 16] = b[n] ^ b[n + 1] ^ b[n + 3] ^ b[n + 12]
   of the feedback taps, and the generator has full period 65535.
2. Templates (7.18.3.3): the Gaussian values at the generator's 11-bit
   outputs and the auto-regressive recursion re-derived here.
3. Scaling (7.18.3.4): the LUT passes through every point, is constant
   outside them and interpolates between them.
4. Noise and blend (7.18.3.5): a block-by-block formulation of the noise
   image (each 32x32 luma block copied from the template at its random
   offset, the 2-column / 2-row overlaps blended) and of the blend.
5. Zero scaling points leave a plane unchanged.
"""
import os
import re

import numpy as np

import common
import pyoracle

HERE = os.path.dirname(os.path.abspath(__file__))


def gaussian_table():
    src = open(os.path.join(os.path.dirname(HERE), "rivet-av1_amd", "csrc", "av1_tables.h")).read()
    m = re.search(r"AV1_GAUSSIAN_SEQUENCE\[2048\] = \{(.*?)\};", src, re.S)
    return np.array([int(v) for v in re.findall(r"-?\d+", m.group(1))])


G = gaussian_table()


def r2(x, n):
    return x if n == 0 else (x + (1 << (n - 1))) >> n


class Lfsr:
    def __init__(self, seed):
        self.r = seed

    def get(self, bits):
        r = self.r
        bit = (r ^ (r >> 1) ^ (r >> 3) ^ (r >> 12)) & 1
        self.r = (r >> 1) | (bit << 15)
        return (self.r >> (16 - bits)) & ((1 << bits) - 1)


def test_lfsr_recurrence_and_period():
    s = 0x1234
    bits = []
    for _ in range(400):
        s = pyoracle.rng_step(s)
        bits.append((s >> 15) & 1)
    for n in range(len(bits) - 16):
        assert bits[n + 16] == bits[n] ^ bits[n + 1] ^ bits[n + 3] ^ bits[n + 12]
    s0, s, n = 1, pyoracle.rng_step(1), 1
    while s != s0:
        s = pyoracle.rng_step(s)
        n += 1
        assert n <= 65535
    assert n == 65535


def templates(P):
    g = P.grain
    bd = P.bit_depth
    gmin, gmax = -(128 << (bd - 8)), (256 << (bd - 8)) - 1 - (128 << (bd - 8))
    shift = 12 - bd + g.grain_scale_shift
    lag = g.ar_coeff_lag
    ars = g.ar_coeff_shift_minus_6 + 6
    L = np.zeros((73, 82), dtype=np.int64)
    rng = Lfsr(g.grain_seed)
    for y in range(73):
        for x in range(82):
            L[y, x] = r2(int(G[rng.get(11)]), shift) if g.num_y_points > 0 else 0
    taps = [(dr, dc) for dr in range(-lag, 1) for dc in range(-lag, lag + 1) if (dr, dc) < (0, 0)]
    for y in range(3, 73):
        for x in range(3, 79):
            s = sum(int(L[y + dr, x + dc]) * (g.ar_coeffs_y_plus_128[i] - 128) for i, (dr, dc) in enumerate(taps))
            L[y, x] = min(max(L[y, x] + r2(s, ars), gmin), gmax)
    C = [np.zeros((73, 82), dtype=np.int64), np.zeros((73, 82), dtype=np.int64)]
    if not P.mono_chrome:
        cw, ch = (44 if P.ssx else 82), (38 if P.ssy else 73)
        for k, (seedx, npts, coef) in enumerate([(0xb524, g.num_cb_points, g.ar_coeffs_cb_plus_128),
                                                  (0x49d8, g.num_cr_points, g.ar_coeffs_cr_plus_128)]):
            rng = Lfsr(g.grain_seed ^ seedx)
            for y in range(ch):
                for x in range(cw):
                    C[k][y, x] = r2(int(G[rng.get(11)]), shift) if (npts > 0 or g.chroma_scaling_from_luma) else 0
        for y in range(3, ch):
            for x in range(3, cw - 3):
                for k, coef in enumerate([g.ar_coeffs_cb_plus_128, g.ar_coeffs_cr_plus_128]):
                    s = sum(int(C[k][y + dr, x + dc]) * (coef[i] - 128) for i, (dr, dc) in enumerate(taps))
                    if g.num_y_points > 0:
                        ly, lx = ((y - 3) << P.ssy) + 3, ((x - 3) << P.ssx) + 3
                        l = int(L[ly:ly + 1 + P.ssy, lx:lx + 1 + P.ssx].sum())
                        s += r2(l, P.ssx + P.ssy) * (coef[len(taps)] - 128)
                    C[k][y, x] = min(max(C[k][y, x] + r2(s, ars), gmin), gmax)
    return L, C[0], C[1]


CASES = [dict(width=100, height=70, bit_depth=8, seed=61),
         dict(width=97, height=65, bit_depth=10, seed=62, grain_lag=2),
         dict(width=64, height=96, bit_depth=12, seed=63, ssx=0, ssy=0, grain_lag=1),
         dict(width=90, height=50, bit_depth=10, seed=64, ssx=1, ssy=0),
         dict(width=80, height=66, bit_depth=12, seed=65, mono=True),
         dict(width=75, height=40, bit_depth=8, seed=66, grain_lag=0, mc_identity=True)]


def test_templates_match():
    for kw in CASES:
        f = common.make_frame(lr=0, grain=True, **kw)
        want = templates(f.params)
        got = pyoracle.grain_templates(f.params)
        cw, ch = (44 if f.params.ssx else 82), (38 if f.params.ssy else 73)
        assert np.array_equal(got[0], want[0]), kw
        if not f.params.mono_chrome:
            assert np.array_equal(got[1][:ch, :cw], want[1][:ch, :cw]), kw
            assert np.array_equal(got[2][:ch, :cw], want[2][:ch, :cw]), kw


def test_scaling_lut_shape():
    for kw in CASES:
        f = common.make_frame(lr=0, grain=True, **kw)
        g = f.params.grain
        lut = pyoracle.grain_scaling_luts(f.params)
        n = g.num_y_points
        xs, ys = list(g.point_y_value[:n]), list(g.point_y_scaling[:n])
        if n == 0:
            continue
        assert all(lut[0, x] == ys[0] for x in range(0, xs[0] + 1))
        assert all(lut[0, x] == ys[-1] for x in range(xs[-1], 256))
        for i in range(n - 1):
            assert lut[0, xs[i]] == ys[i]
            seg = lut[0, xs[i]:xs[i + 1] + 1]
            lo, hi = min(ys[i], ys[i + 1]), max(ys[i], ys[i + 1])
            assert seg.min() >= lo and seg.max() <= hi
            # linear within rounding
            t = (np.arange(xs[i], xs[i + 1] + 1) - xs[i]) / (xs[i + 1] - xs[i])
            assert np.abs(seg - (ys[i] + t * (ys[i + 1] - ys[i]))).max() <= 1.0


def noise_image(P, T, plane, w, h):
    """Block-by-block noise of one plane (7.18.3.5)."""
    g = P.grain
    bd = P.bit_depth
    gmin, gmax = -(128 << (bd - 8)), (256 << (bd - 8)) - 1 - (128 << (bd - 8))
    sx, sy = (P.ssx, P.ssy) if plane else (0, 0)
    pw, ph = (w + sx) >> sx, (h + sy) >> sy
    bw, bh = 32 >> sx, 32 >> sy          # block pitch in this plane
    ow, oh = (2 >> sx), (2 >> sy)        # overlap
    nstripes = -(-((h + 1) // 2) // 16)
    nblocks = -(-((w + 1) // 2) // 16)
    stripes = []
    for ln in range(nstripes):
        rng = Lfsr(g.grain_seed ^ ((((ln * 37 + 178) & 255) << 8)) ^ ((ln * 173 + 105) & 255))
        S = np.zeros((bh + oh, nblocks * bw + ow + 40), dtype=np.int64)
        for b in range(nblocks):
            rnd = rng.get(8)
            ox, oy = rnd >> 4, rnd & 15
            px = 6 + ox if sx else 9 + 2 * ox
            py = 6 + oy if sy else 9 + 2 * oy
            blk = T[py:py + bh + oh, px:px + bw + ow].astype(np.int64)
            x0 = b * bw
            if g.overlap_flag and b > 0:
                old = S[:, x0:x0 + ow].copy()
                if sx:
                    mix = old * 23 + blk[:, :1] * 22
                else:
                    mix = np.stack([old[:, 0] * 27 + blk[:, 0] * 17, old[:, 1] * 17 + blk[:, 1] * 27], axis=1)
                blk = blk.copy()
                blk[:, :ow] = np.clip((mix + 16) >> 5, gmin, gmax)
            S[:, x0:x0 + bw + ow] = blk
        stripes.append(S)
    img = np.zeros((ph, pw), dtype=np.int64)
    for y in range(ph):
        ln, i = y // bh, y % bh
        row = stripes[ln][i, :pw].copy()
        if g.overlap_flag and ln > 0 and i < oh:
            old = stripes[ln - 1][i + bh, :pw]
            if sy:
                mix = old * 23 + row * 22
            else:
                mix = old * 27 + row * 17 if i == 0 else old * 17 + row * 27
            row = np.clip((mix + 16) >> 5, gmin, gmax)
        img[y] = row
    return img


def apply_grain(P, w, h, planes):
    g = P.grain
    bd = P.bit_depth
    T = templates(P)
    lut = pyoracle.grain_scaling_luts(P).astype(np.int64)

    def scale(p, idx):
        sh = bd - 8
        x = idx >> sh
        rem = idx - (x << sh)
        if bd == 8:
            return lut[p][x]
        nxt = lut[p][np.minimum(x + 1, 255)]
        v = lut[p][x] + ((nxt - lut[p][x]) * rem + ((1 << sh) >> 1) >> sh)
        return np.where(x == 255, lut[p][x], v)

    if g.clip_to_restricted_range:
        lo, yhi = 16 << (bd - 8), 235 << (bd - 8)
        chi = yhi if P.matrix_coefficients == 0 else 240 << (bd - 8)
    else:
        lo, yhi, chi = 0, (256 << (bd - 8)) - 1, (256 << (bd - 8)) - 1
    sh = g.grain_scaling_minus_8 + 8
    out = [None if a is None else a.astype(np.int64) for a in planes]
    Y = planes[0].astype(np.int64)
    if not P.mono_chrome:
        cw, ch = (w + P.ssx) >> P.ssx, (h + P.ssy) >> P.ssy
        lx = np.arange(cw) << P.ssx
        ly = np.arange(ch) << P.ssy
        if P.ssx:
            avg = (Y[np.ix_(ly, lx)] + Y[np.ix_(ly, np.minimum(lx + 1, w - 1))] + 1) >> 1
        else:
            avg = Y[np.ix_(ly, lx)]
        for p, (npts, mult, lmult, off) in ((1, (g.num_cb_points, g.cb_mult, g.cb_luma_mult, g.cb_offset)),
                                           (2, (g.num_cr_points, g.cr_mult, g.cr_luma_mult, g.cr_offset))):
            if not (npts > 0 or g.chroma_scaling_from_luma):
                continue
            orig = planes[p].astype(np.int64)
            if g.chroma_scaling_from_luma:
                merged = avg
            else:
                merged = np.clip((avg * (lmult - 128) + orig * (mult - 128) >> 6) + ((off - 256) << (bd - 8)), 0,
                                 (1 << bd) - 1)
            noise = noise_image(P, T[p], p, w, h)
            sn = (scale(p, merged) * noise + (1 << (sh - 1))) >> sh
            out[p] = np.clip(orig + sn, lo, chi)
    if g.num_y_points > 0:
        noise = noise_image(P, T[0], 0, w, h)
        sn = (scale(0, Y) * noise + (1 << (sh - 1))) >> sh
        out[0] = np.clip(Y + sn, lo, yhi)
    return [None if a is None else a.astype(np.uint16) for a in out]


def test_apply_matches_block_formulation():
    for kw in CASES + [dict(width=130, height=200, bit_depth=10, seed=67, grain_lag=3)]:
        f = common.make_frame(lr=0, grain=True, **kw)
        P = f.params
        w, h = P.upscaled_width, P.frame_height
        base = [None if p >= f.num_planes else
                np.ascontiguousarray(f.rec[p][:f.crop_shape(p)[0], :f.crop_shape(p)[1]]) for p in range(3)]
        got = pyoracle.grain_apply(P, w, h, [None if a is None else a.copy() for a in base])
        want = apply_grain(P, w, h, base)
        for p in range(f.num_planes):
            assert np.array_equal(got[p], want[p]), (kw, p)
        assert any(not np.array_equal(got[p], base[p]) for p in range(f.num_planes))


def test_zero_points_identity():
    f = common.make_frame(width=64, height=48, bit_depth=10, lr=0, grain=True, seed=70)
    P = f.params
    P.grain.num_y_points = 0
    P.grain.num_cb_points = 0
    P.grain.num_cr_points = 0
    P.grain.chroma_scaling_from_luma = 0
    base = [np.ascontiguousarray(f.rec[p][:f.crop_shape(p)[0], :f.crop_shape(p)[1]]) for p in range(3)]
    got = pyoracle.grain_apply(P, 64, 48, [a.copy() for a in base])
    for p in range(3):
        assert np.array_equal(got[p], base[p])
