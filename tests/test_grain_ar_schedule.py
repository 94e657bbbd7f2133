"""The film grain prepare kernel's auto-regressive schedule, replayed on the host.

grain.hip filters each 73x82 template on one wave along the diagonals
T = K*y + x (K = lag + 1): lane l owns rows (l + 3) mod 64 and that + 64,
keeps the last 16 emitted values of its own row and of the three rows
above (passed between lanes by one rotation each per diagonal), and works
in chunks of 16 diagonals whose LDS accesses are one base index per lane
plus the step (ar_place). Steps whose column lies outside the lane's row
read and write neighbouring template positions, padding columns or a
spare run after the template; chroma units run in barrier rounds after
the luma chunks they read (chroma_first_round).

The replay below runs that schedule lane for lane on the host, on the
kernel's LDS layout (TFRONT samples before each template, 64 after, the
padding and spare samples filled with random values), and checks:
  - every LDS index stays inside the template's allocation;
  - a chroma sample reads co-located luma samples only after the round in
    which the luma wave filtered them;
  - the rows the fourth wave copies out in a round (rows_final) already hold
    their final values, whatever the other waves do in that round;
  - the resulting templates equal the oracle's (orc_grain_templates) at
    every lag, chroma layout and bit depth, with and without luma points
    (the chroma luma tap) and with chroma scaling from luma.
Any garbage that reached a filtered sample (a spare or padding value, a
stale write, a tap from the wrong lane or diagonal) would change it.
"""
import numpy as np
import pytest

import common
import pyoracle

TW, TH = 82, 73
TFRONT = 16
TPAD = TFRONT + TH * TW + 64
AR_ROUND = 16
LANES = np.arange(64)


def gaussian_table():
    import os
    import re
    here = os.path.dirname(os.path.abspath(__file__))
    src = open(os.path.join(os.path.dirname(here), "rivet-av1_amd", "csrc", "av1_tables.h")).read()
    m = re.search(r"AV1_GAUSSIAN_SEQUENCE\[2048\] = \{(.*?)\};", src, re.S)
    return np.array([int(v) for v in re.findall(r"-?\d+", m.group(1))])


G = gaussian_table()


def r2(x, n):
    return x if n == 0 else (x + (1 << (n - 1))) >> n


def raw_template(seed, rows, cols, shift, on):
    """Gaussian template before the AR filter (7.18.3.3, raster LFSR order)."""
    t = np.zeros((rows, cols), dtype=np.int64)
    r = seed
    for y in range(rows):
        for x in range(cols):
            bit = (r ^ (r >> 1) ^ (r >> 3) ^ (r >> 12)) & 1
            r = (r >> 1) | (bit << 15)
            t[y, x] = r2(int(G[(r >> 5) & 2047]), shift) if on else 0
    return t


def ar_taps(lag):
    """ArTaps<LAG>: per dot product, the row above (d), the diagonal distance
    of its pair (m) and the coefficient indices of its two halves."""
    K = lag + 1
    out = []
    for dd in range(lag, -1, -1):
        base = (lag - dd) * (2 * lag + 1)
        first, last = -lag, (lag if dd else -1)
        q = []
        dc = last
        while dc >= first:
            q.append((base + (dc - 1 + lag) if dc - 1 >= first else -1, base + (dc + lag)))
            dc -= 2
        for k in range(len(q) - 1, -1, -1):
            dc_hi = last - 2 * k
            out.append((dd, dd * K - (dc_hi - 1), q[k][0], q[k][1]))
    return out


class Lds:
    """One template's LDS array, indexed from the template's origin."""

    def __init__(self, raw, rows, cols, rng):
        self.a = rng.integers(-32768, 32768, TPAD).astype(np.int64)
        for y in range(rows):
            self.a[TFRONT + y * TW:TFRONT + y * TW + cols] = raw[y]

    def idx(self, i):
        i = np.asarray(i)
        assert (i >= -TFRONT).all() and (i < TH * TW + 64).all(), "LDS index out of the template's allocation"
        return TFRONT + i

    def read(self, i):
        return self.a[self.idx(i)]

    def write(self, i, v):
        # lanes writing one address in one instruction: any of them may win
        self.a[self.idx(i)] = ((np.asarray(v) + 32768) & 0xFFFF) - 32768

    def template(self, rows, cols):
        return np.stack([self.a[TFRONT + y * TW:TFRONT + y * TW + cols] for y in range(rows)])


def ar_first(K):
    return 3 * K + 3


def ar_chunks(K, rows, cols):
    return (K * (rows - 1) + cols - 4 - ar_first(K)) // AR_ROUND + 1


def ar_place(K, T0, rows, cols):
    ya = (LANES + 3) & 63
    y = ya.copy()
    x0 = T0 - K * ya
    wrap = x0 >= cols
    y[wrap] += 64
    x0[wrap] -= 64 * K
    inside = y < rows
    lo = np.clip(3 - x0, 0, AR_ROUND)
    hi = np.clip(cols - 3 - x0, 0, AR_ROUND)
    mask = np.where(inside & (y >= 3) & (hi > lo), (1 << hi) - (1 << lo), 0)
    base = np.where(inside, y * TW + x0, TH * TW)
    return base, mask, y, x0


def rows_final(K, D, rows, cols):
    n = (D - (cols - 4)) // K + 1 if D >= cols - 4 else 0
    return min(max(n, 3), rows)


def chroma_first_round(K, ssx, ssy, cols, units, halves):
    TF, L = ar_first(K), AR_ROUND // halves
    last_luma = K * (TH - 1) + TW - 4
    first = 0
    for u in range(units):
        b = TF + L * u + L - 1
        d = 2 * b - 2 * K - 2 if ssy else (b + cols - 6 if ssx else b)
        d = min(d, last_luma)
        dep = (d - TF) // AR_ROUND if d >= TF else -1
        first = max(first, dep + 1 - u)
    return first


class Wave:
    """One AR wave: its lanes' history registers and schedule."""

    def __init__(self, lag, lds, rows, cols, coef, ars, gmin, gmax, chroma=None):
        self.lag, self.K = lag, lag + 1
        self.lds, self.rows, self.cols = lds, rows, cols
        self.taps = ar_taps(lag)
        self.coef = coef
        self.ars, self.gmin, self.gmax = ars, gmin, gmax
        self.chroma = chroma  # (luma Lds, ssx, ssy, luma_tap, luma_done) or None
        z = lambda: np.zeros((AR_ROUND, 64), dtype=np.int64)  # noqa: E731
        # history pairs (value of diagonal j, value of j + 1) of the lane's row
        # (d = 0) and of the rows 1, 2, 3 above
        self.lo = [z() for _ in range(4)]
        self.hi = [z() for _ in range(4)]
        self.prev = np.zeros(64, dtype=np.int64)

    def load(self, T0):
        """The chunk's 16 samples (load_chunk), read when the previous chunk starts."""
        base = ar_place(self.K, T0, self.rows, self.cols)[0]
        return [self.lds.read(base + s) for s in range(AR_ROUND)]

    def steps(self, T0, s0, s1, filt, rnd, tvc, luma_round=None):
        base, mask, y, x0 = ar_place(self.K, T0, self.rows, self.cols)
        tv = tvc[s0:s1]
        lv = None
        if self.chroma is not None and filt and self.chroma[3]:
            luma, ssx, ssy, _, done = self.chroma
            lb = np.where(mask != 0, (((y - 3) << ssy) + 3) * TW + ((x0 - 3) << ssx) + 3, 0)
            lv = []
            for s in range(s0, s1):
                o = lb + (s << ssx)
                pos = [o] + ([o + 1] if ssx else []) + ([o + TW] if ssy else []) + ([o + TW + 1] if ssx and ssy else [])
                tot = sum(luma.read(p) for p in pos)
                # filtered steps read luma samples the luma wave finished in an earlier round
                f = ((mask >> s) & 1) == 1
                for p in pos:
                    pf = p[f]
                    ly, lx = pf // TW, pf % TW
                    filt_luma = (ly >= 3) & (lx >= 3) & (lx < TW - 3)
                    assert (done[ly[filt_luma], lx[filt_luma]] < rnd).all(), "chroma read luma before it was filtered"
                lv.append(r2(tot, ssx + ssy))
        for k, s in enumerate(range(s0, s1)):
            out = tv[k]
            if filt:
                acc = np.full(64, 1 << (self.ars - 1), dtype=np.int64)
                if lv is not None:
                    acc = acc + lv[k] * self.coef[2 * self.lag * self.K]
                for (d, m, clo, chi) in self.taps:
                    idx = (s - m) & (AR_ROUND - 1)
                    acc = acc + self.lo[d][idx] * (self.coef[clo] if clo >= 0 else 0) + self.hi[d][idx] * self.coef[chi]
                val = np.clip(tv[k] + (acc >> self.ars), self.gmin, self.gmax)
                am = ((mask >> s) & 1) == 1
                out = np.where(am, val, tv[k])
                self.lds.write(base + s, out)
                if luma_round is not None:
                    yy, xx = (base + s)[am] // TW, (base + s)[am] % TW
                    luma_round[yy, xx] = rnd
            j = (s - 1) & (AR_ROUND - 1)
            self.lo[0][j], self.hi[0][j] = self.prev, out
            self.prev = out
            for d in range(1, self.lag + 1):
                # lane l receives lane l - 1's pair (DPP wave_ror:1)
                self.lo[d][j] = np.roll(self.lo[d - 1][j], 1)
                self.hi[d][j] = np.roll(self.hi[d - 1][j], 1)


def replay(P, seed=0):
    g = P.grain
    bd, lag = P.bit_depth, g.ar_coeff_lag
    K = lag + 1
    gmin, gmax = -(128 << (bd - 8)), (256 << (bd - 8)) - 1 - (128 << (bd - 8))
    shift = 12 - bd + g.grain_scale_shift
    ars = g.ar_coeff_shift_minus_6 + 6
    chroma = not P.mono_chrome
    cw, ch = (44 if P.ssx else 82), (38 if P.ssy else 73)
    rng = np.random.default_rng(seed)
    luma = Lds(raw_template(g.grain_seed, TH, TW, shift, g.num_y_points > 0), TH, TW, rng)
    cl = []
    for seedx, npts in ((0xb524, g.num_cb_points), (0x49d8, g.num_cr_points)):
        on = npts > 0 or g.chroma_scaling_from_luma
        cl.append(Lds(raw_template(g.grain_seed ^ seedx, ch, cw, shift, on), ch, cw, rng))
    cy = [c - 128 for c in g.ar_coeffs_y_plus_128]
    ccb = [c - 128 for c in g.ar_coeffs_cb_plus_128]
    ccr = [c - 128 for c in g.ar_coeffs_cr_plus_128]
    luma_tap = g.num_y_points > 0
    halves = 2 if P.ssy else 1
    nl = ar_chunks(K, TH, TW)
    cu = ar_chunks(K, ch, cw) * halves if chroma else 0
    first = chroma_first_round(K, P.ssx, P.ssy, cw, cu, halves) if chroma and luma_tap else 0
    rounds = max(nl, first + cu)
    done = np.full((TH, TW), 10 ** 6)
    wl = Wave(lag, luma, TH, TW, cy, ars, gmin, gmax)
    wc = [Wave(lag, cl[i], ch, cw, (ccb, ccr)[i], ars, gmin, gmax, chroma=(luma, P.ssx, P.ssy, luma_tap, done))
          for i in range(2)] if chroma else []
    TF = ar_first(K)
    chunks = {id(wl): nl}
    for w in wc:
        chunks[id(w)] = cu // halves
    for w in [wl] + wc:
        w.steps(TF - AR_ROUND, 0, AR_ROUND, False, -1, w.load(TF - AR_ROUND))
        w.tvn = w.load(TF)

    def run_unit(w, u, hv, r, luma_round=None):
        c = u if hv == 1 else u >> 1
        if hv == 1 or u % 2 == 0:
            w.tvc = w.tvn
            if c + 1 < chunks[id(w)]:
                w.tvn = w.load(TF + AR_ROUND * (c + 1))
        s0 = 0 if hv == 1 or u % 2 == 0 else AR_ROUND // 2
        w.steps(TF + AR_ROUND * c, s0, s0 + AR_ROUND // hv, True, r, w.tvc, luma_round)
    L = AR_ROUND // halves
    # with full-height chroma (4:4:4, 4:2:2) the fourth wave copies rows
    # out during the rounds; otherwise every row is copied afterwards
    behind = chroma and not P.ssy
    copied = []  # (Lds, rows, cols, y, row as copied)
    dl = dc = 0
    for r in range(rounds):
        # the fourth wave copies rows final by the previous round
        fl = rows_final(K, TF + AR_ROUND * min(r, nl) - 1, TH, TW)
        fc = rows_final(K, TF + L * min(max(r - first, 0), cu) - 1, ch, cw) if chroma else 0
        if behind:
            copied += [(luma, TW, y, luma.a[TFRONT + y * TW:TFRONT + y * TW + TW].copy()) for y in range(dl, fl)]
            copied += [(c, cw, y, c.a[TFRONT + y * TW:TFRONT + y * TW + cw].copy()) for c in cl for y in range(dc, fc)]
            dl, dc = fl, fc
        # the chroma waves run beside the luma wave: they may only rely on
        # luma chunks of earlier rounds (checked through `done`)
        for w in wc:
            u = r - first
            if 0 <= u < cu:
                run_unit(w, u, halves, r)
        if r < nl:
            run_unit(wl, r, 1, r, luma_round=done)
    for lds, cols, y, row in copied:
        assert np.array_equal(row, lds.a[TFRONT + y * TW:TFRONT + y * TW + cols]), "row copied before it was final"
    if behind:
        assert dl == rows_final(K, TF + AR_ROUND * min(rounds - 1, nl) - 1, TH, TW)
    out = [luma.template(TH, TW)]
    out += [c.template(ch, cw) for c in cl] if chroma else []
    return out, (cw, ch), rounds


CASES = [
    dict(width=64, height=48, bit_depth=10, grain_lag=3),
    dict(width=64, height=48, bit_depth=8, grain_lag=2),
    dict(width=64, height=48, bit_depth=12, grain_lag=1),
    dict(width=64, height=48, bit_depth=10, ssx=1, ssy=0, grain_lag=3),
    dict(width=64, height=48, bit_depth=8, ssx=1, ssy=0, grain_lag=1),
    dict(width=64, height=48, bit_depth=12, ssx=0, ssy=0, grain_lag=3),
    dict(width=64, height=48, bit_depth=10, ssx=0, ssy=0, grain_lag=2),
    dict(width=64, height=48, bit_depth=12, mono=True, grain_lag=3),
]


@pytest.mark.parametrize("kw", CASES, ids=lambda kw: "lag%d_%s_%db" % (
    kw["grain_lag"], "mono" if kw.get("mono") else "%d%d" % (kw.get("ssx", 1), kw.get("ssy", 1)), kw["bit_depth"]))
@pytest.mark.parametrize("variant", ["points", "no_luma_points", "cfl"])
def test_ar_schedule_matches_oracle(kw, variant):
    f = common.make_frame(lr=0, grain=True, seed=300 + kw["grain_lag"] * 7 + kw["bit_depth"], **kw)
    P = f.params
    if variant == "no_luma_points":
        if P.mono_chrome:
            pytest.skip("monochrome has no chroma luma tap")
        P.grain.num_y_points = 0
    elif variant == "cfl":
        if P.mono_chrome:
            pytest.skip("monochrome has no chroma")
        P.grain.chroma_scaling_from_luma = 1
        P.grain.num_cb_points = 0
        P.grain.num_cr_points = 0
    got, (cw, ch), _ = replay(P, seed=kw["grain_lag"])
    want = pyoracle.grain_templates(P)
    assert np.array_equal(got[0], want[0][:TH, :TW])
    if not P.mono_chrome:
        assert np.array_equal(got[1], want[1][:ch, :cw])
        assert np.array_equal(got[2], want[2][:ch, :cw])


def test_rounds_per_layout():
    """Barrier rounds per template set: the luma chunks, or the chroma
    units after their first round if chroma ends later."""
    f = common.make_frame(width=64, height=48, bit_depth=10, lr=0, grain=True, grain_lag=3, seed=5)
    _, _, rounds = replay(f.params)
    assert rounds == 24  # 4:2:0, lag 3: 22 luma chunks, 22 chroma half chunks from round 2
    assert ar_chunks(4, TH, TW) == 22
    assert chroma_first_round(4, 0, 0, 82, ar_chunks(4, TH, TW), 1) == 1
