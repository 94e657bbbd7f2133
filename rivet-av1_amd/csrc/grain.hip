/*
 * grain.hip — film grain synthesis (AV1 7.18.3) for gfx950.
 *
 * Follows rivet-av1's grain::apply (src/decoder/grain.rs:8-354).
 *
 * grain_prepare_kernel: per frame, workgroup 0 (four waves) builds the
 * templates and workgroups 1 .. TABLE_ITEMS the scaling functions
 * (7.18.3.4, at every sample value) and the per-stripe random block
 * offsets beside it.
 *   - Gaussian templates (7.18.3.3): luma 73x82, cb/cr 38x44 (4:2:0) or
 *     larger. The 16-bit LFSR is linear over GF(2), so the thread of a
 *     template row (or of its second half) starts from A^n seed (a
 *     precomputed matrix per row start) and steps along its row.
 *   - The auto-regressive filter is a recursion: sample (y, x) reads rows
 *     y-lag..y and columns x-lag..x+lag. Processing it along the diagonals
 *     T = (lag+1)*y + x visits every dependency at an earlier T, so all
 *     samples of one T run in parallel (one lane per template row). The
 *     rows above a lane's row belong to the lanes below it, so the taps
 *     are passed between lanes by DPP rotations and stay in registers.
 *     Luma runs on wave 0, cb and cr on waves 1 and 2, in barrier rounds a
 *     fixed number behind the luma samples they read.
 *   - The filtered templates go to the scratch in 8 copies, each shifted
 *     by one more sample (wave 3 copies finished rows during the filter
 *     where it does not slow it, the rest follow it).
 * grain_apply_kernel: one workgroup per (noise stripe, 64 luma columns).
 *   The noise of a sample is rebuilt directly from the template: its 32x32
 *   block's offset, the horizontal blend with the block to the left, then
 *   the vertical blend with the stripe above (whose value is itself
 *   horizontally blended), exactly the order in which the reference fills
 *   its noise stripes and noise image. A thread takes 8-sample row
 *   segments that lie in one noise block, so their noise is 8 consecutive
 *   template samples: one aligned 16-byte load from one of 8 copies of the
 *   templates, each shifted by one more sample, that the prepare kernel
 *   writes. Chroma reads the un-noised luma of the input, as the reference
 *   blends chroma first. Planes whose rows are not aligned for segment
 *   accesses take a sample-by-sample variant.
 */
#include "av1_tables.h"
#include "rvpf_internal.h"

namespace rvpf {

namespace {

/* --- LFSR jump tables ------------------------------------------------ */
constexpr uint16_t lfsr_step(uint16_t r) {
  return (uint16_t)((r >> 1) | ((((r) ^ (r >> 1) ^ (r >> 3) ^ (r >> 12)) & 1) << 15));
}

constexpr uint16_t apply_cols(const uint16_t *cols, uint16_t s) {
  uint16_t r = 0;
  for (int j = 0; j < 16; j++)
    if ((s >> j) & 1) r ^= cols[j];
  return r;
}

/* LFSR states at the start of every template row: col[w][y][j] = A^(y W) e_j
 * for rows of W = 82 (w = 0) or 44 (w = 1) samples, so the state before
 * sample (y, 0) is one matrix-vector product with the seed; col[2][y][j] =
 * A^(82 y + FILL_SEG) e_j, the state before sample (y, FILL_SEG) of an
 * 82-column template. */
constexpr int FILL_SEG = 48; /* columns per fill item */

struct alignas(16) RowJump {
  uint16_t col[3][73][16];
};

constexpr RowJump make_row_jump() {
  RowJump t{};
  const int widths[2] = {82, 44};
  for (int w = 0; w < 2; w++) {
    uint16_t m1[16] = {}; /* A^W */
    for (int j = 0; j < 16; j++) {
      uint16_t v = (uint16_t)(1u << j);
      for (int k = 0; k < widths[w]; k++) v = lfsr_step(v);
      m1[j] = v;
    }
    for (int j = 0; j < 16; j++) t.col[w][0][j] = (uint16_t)(1u << j);
    for (int y = 1; y < 73; y++)
      for (int j = 0; j < 16; j++) t.col[w][y][j] = apply_cols(m1, t.col[w][y - 1][j]);
  }
  uint16_t ms[16] = {}; /* A^FILL_SEG */
  for (int j = 0; j < 16; j++) {
    uint16_t v = (uint16_t)(1u << j);
    for (int k = 0; k < FILL_SEG; k++) v = lfsr_step(v);
    ms[j] = v;
  }
  for (int y = 0; y < 73; y++)
    for (int j = 0; j < 16; j++) t.col[2][y][j] = apply_cols(ms, t.col[0][y][j]);
  return t;
}

constexpr RowJump ROW_JUMP = make_row_jump();

__device__ __forceinline__ uint32_t step(uint32_t r) {
  const uint32_t bit = (r ^ (r >> 1) ^ (r >> 3) ^ (r >> 12)) & 1;
  return (r >> 1) | (bit << 15);
}

__device__ __forceinline__ int round2(int x, int n) { return (x + ((1 << n) >> 1)) >> n; }
__device__ __forceinline__ int clip3(int lo, int hi, int x) { return min(max(x, lo), hi); }

constexpr int TW = 82, TH = 73;
/* four waves: the AR filter of luma, cb and cr on waves 0-2, wave 3
 * writes finished template rows out meanwhile */
constexpr int PREP_THREADS = 256;

/* Row y of a Gaussian template (7.18.3.3, grain.rs:40-56): the LFSR
 * sequence from `seed` runs over the template in raster order, so the row
 * starts at state A^(y cols) seed (row_state). The LFSR's bit stream b obeys
 * b[k + 16] = b[k] ^ b[k + 1] ^ b[k + 3] ^ b[k + 12], so 4 new bits at a
 * time come from bits already known: 16 steps are 4 such groups, after
 * which a 32-bit window holds b[0 .. 31] and the state after step m is
 * b[m .. m + 15]; its table index (state >> 5) & 2047 is the 11-bit field at
 * m + 5. The 16 table reads are issued before the 16 template writes.
 * Template rows that are off hold 0. */
__device__ void fill_row(int16_t *t, const int16_t *gauss, uint32_t s, int y, int xb, int cols, int shift, bool on) {
  const int keep = on ? -1 : 0;
  int16_t *row = t + y * TW;
  for (int x0 = xb; x0 < cols; x0 += 16) {
    uint32_t w = s;
#pragma unroll
    for (int g = 0; g < 4; g++) {
      const int n = 4 * g;
      const uint32_t nb = ((w >> n) ^ (w >> (n + 1)) ^ (w >> (n + 3)) ^ (w >> (n + 12))) & 15u;
      w |= nb << (16 + n);
    }
    int v[16];
#pragma unroll
    for (int m = 1; m <= 16; m++) v[m - 1] = gauss[__builtin_amdgcn_ubfe(w, m + 5, 11)];
#pragma unroll
    for (int k = 0; k < 16; k++)
      if (x0 + k < cols) row[x0 + k] = (int16_t)(round2(v[k], shift) & keep);
    s = w >> 16;
  }
}

/* The LFSR state before sample (y, 0) of a template of 82 (wsel 0) or 44
 * (wsel 1) columns, or before sample (y, FILL_SEG) of an 82-column one
 * (wsel 2): one GF(2) matrix-vector product with the seed, from the row's
 * 16 columns (two 16-byte loads). */
__device__ __forceinline__ uint32_t row_state(uint32_t seed, int wsel, int y) {
  const uint4 *c = reinterpret_cast<const uint4 *>(ROW_JUMP.col[wsel][y]);
  const uint4 a = c[0], b = c[1];
  const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
  uint32_t s = 0;
#pragma unroll
  for (int j = 0; j < 16; j++) {
    const uint32_t col = (w[j >> 1] >> ((j & 1) * 16)) & 0xffffu;
    s ^= col & (0u - ((seed >> j) & 1u));
  }
  return s;
}

/* One chroma AR sample (grain.rs:101-134) with no neighbour taps (lag 0):
 * cb and cr from the co-located (averaged) luma alone. */
__device__ __forceinline__ void chroma_lag0(int16_t *tcb, int16_t *tcr, const int16_t *tl, const GrainHdr &h, int y,
                                            int x, int gmin, int gmax) {
  int a0 = 0, a1 = 0;
  if (h.num_y > 0) {
    const int lx = ((x - 3) << h.ssx) + 3, ly = ((y - 3) << h.ssy) + 3;
    int l = 0;
    for (int i = 0; i <= h.ssy; i++)
      for (int j = 0; j <= h.ssx; j++) l += tl[(ly + i) * TW + lx + j];
    l = round2(l, h.ssx + h.ssy);
    a0 = l * h.ar_cb[0];
    a1 = l * h.ar_cr[0];
  }
  tcb[y * TW + x] = (int16_t)clip3(gmin, gmax, tcb[y * TW + x] + round2(a0, h.ar_shift));
  tcr[y * TW + x] = (int16_t)clip3(gmin, gmax, tcr[y * TW + x] + round2(a1, h.ar_shift));
}

/* Lane l receives the value of lane (l - 1) mod 64 (DPP wave_ror:1). */
__device__ __forceinline__ int from_lane_below(int v) {
  return __builtin_amdgcn_mov_dpp(v, 0x13C, 0xF, 0xF, false);
}

/* Diagonals per chunk of the AR recursion. */
constexpr int AR_ROUND = 16;

/* The first filtered diagonal of every template, that of sample (3, 3). */
constexpr int ar_first(int K) { return 3 * K + 3; }

/* Chunks of AR_ROUND diagonals from the first filtered diagonal through
 * the last one, that of sample (rows - 1, cols - 4). */
__device__ __forceinline__ int ar_chunks(int K, int rows, int cols) {
  return (K * (rows - 1) + cols - 4 - ar_first(K)) / AR_ROUND + 1;
}

typedef short short2_g __attribute__((ext_vector_type(2)));

/* acc + a.lo * b.lo + a.hi * b.hi (signed 16-bit halves; v_dot2_i32_i16) */
__device__ __forceinline__ int dot2(uint32_t a, uint32_t b, int acc) {
  return __builtin_amdgcn_sdot2(__builtin_bit_cast(short2_g, a), __builtin_bit_cast(short2_g, b), acc, false);
}

/* clip3(lo, hi, x) as one v_med3_i32 (lo <= hi) */
__device__ __forceinline__ int med3(int lo, int hi, int x) {
  int r;
  asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(lo), "v"(hi));
  return r;
}

/* a.lo * b.lo + a.hi * b.hi: the VOP3P form with a zero addend, so no
 * register is set to 0 first */
__device__ __forceinline__ int dot2_0(uint32_t a, uint32_t b) {
  int r;
  asm("v_dot2_i32_i16 %0, %1, %2, 0" : "=v"(r) : "v"(a), "v"(b));
  return r;
}

/* (a & m) | (b & ~m) as one v_bfi_b32 */
__device__ __forceinline__ int bfi(int m, int a, int b) {
  int r;
  asm("v_bfi_b32 %0, %1, %2, %3" : "=v"(r) : "v"(m), "v"(a), "v"(b));
  return r;
}

/* (x << sh) + c as one v_lshl_add_u32 (c in a VGPR, so the shift may be an
 * SGPR) */
__device__ __forceinline__ int lshl_add(int x, int sh, int c) {
  int r;
  asm("v_lshl_add_u32 %0, %1, %2, %3" : "=v"(r) : "v"(x), "s"(sh), "v"(c));
  return r;
}

/* (lo & 0xffff) | (hi << 16) */
__device__ __forceinline__ uint32_t pack16(int lo, int hi) {
  return __builtin_amdgcn_perm((uint32_t)hi, (uint32_t)lo, 0x05040100u);
}

/* The taps of the AR filter as dot products of packed pairs. Row d above
 * (d = LAG .. 1) has the 2 LAG + 1 taps dc = -LAG .. LAG, the row itself
 * (d = 0) the LAG taps dc = -LAG .. -1. Tap (d, dc) of the sample on
 * diagonal T lies on diagonal T - (d K - dc), so the taps of a row are
 * consecutive diagonals; they are paired from the latest one down, (dc - 1,
 * dc), and an odd one out takes the upper half of a pair whose lower
 * coefficient is 0. A pair (dc_lo, dc_lo + 1) is the packed history entry
 * of diagonal T - m, m = d K - dc_lo. */
template <int LAG>
struct ArTaps {
  static constexpr int K = LAG + 1;
  static constexpr int NP = LAG * (LAG + 1) + (LAG + 1) / 2; /* dot products per sample */
  int d[NP], m[NP], lo[NP], hi[NP];                        /* coefficient index of each half, -1 = 0 */
  constexpr ArTaps() : d{}, m{}, lo{}, hi{} {
    int n = 0;
    /* oldest pairs first, so the latest taps come last in the sum */
    for (int dd = LAG; dd >= 0; dd--) {
      const int base = (LAG - dd) * (2 * LAG + 1);
      const int first = -LAG, last = dd ? LAG : -1;
      int q[2 * LAG + 2][2] = {}, nq = 0;
      for (int dc = last; dc >= first; dc -= 2) {
        q[nq][0] = dc - 1 >= first ? base + (dc - 1 + LAG) : -1;
        q[nq][1] = base + (dc + LAG);
        nq++;
      }
      for (int k = nq - 1; k >= 0; k--) {
        const int dc_hi = last - 2 * k;
        d[n] = dd;
        m[n] = dd * K - (dc_hi - 1);
        lo[n] = q[k][0];
        hi[n] = q[k][1];
        n++;
      }
    }
  }
};

/* Where a lane works in the chunk of diagonals T0 .. T0 + 15. Lane l owns
 * the template rows ya = (l + 3) mod 64 and ya + 64; on diagonal T its
 * sample is x = T - K y. It takes row ya while the chunk starts inside it
 * (x0 < cols), then ya + 64 (the two rows are never both inside the
 * template in one chunk); a row past the template sends the lane to a
 * 16-sample spare run after it. The chunk's accesses are then base + s for
 * step s (immediate offsets). A step whose x lies outside the row reads
 * and writes a neighbouring position (the end of the row above, the start
 * of the row below, or a padding column): it writes back the value it read
 * at the chunk's start, and such a position is written by its own lane at
 * least 82 - K diagonals away, in another chunk, so nothing changes. Bit s
 * of mask is set where step s is a filtered sample (y >= 3, 3 <= x < cols -
 * 3); every tap of a filtered sample is a template position, emitted by its
 * own lane on its own diagonal. */
struct ArPlace {
  int base;      /* template index of the sample of step 0 */
  uint32_t mask; /* filtered steps */
  int y, x0;     /* row, column of step 0 */
};

template <int K>
__device__ __forceinline__ ArPlace ar_place(int T0, int rows, int cols) {
  const int ya = (threadIdx.x + 3) & 63;
  int y = ya, x0 = T0 - K * ya;
  if (x0 >= cols) {
    y += 64;
    x0 -= 64 * K;
  }
  const bool in = y < rows;
  const int lo = min(max(3 - x0, 0), AR_ROUND), hi = min(max(cols - 3 - x0, 0), AR_ROUND);
  ArPlace p;
  p.mask = (in && y >= 3 && hi > lo) ? (1u << hi) - (1u << lo) : 0u;
  p.base = in ? y * TW + x0 : TH * TW;
  p.y = y;
  p.x0 = x0;
  return p;
}

/* volatile: one 16-bit LDS access per sample; merged into wider
 * (unaligned) accesses they would cost VALU packing and unpacking */
typedef __attribute__((address_space(3))) volatile int16_t lds_i16;

/* Steps S0 .. S1 - 1 of one chunk of a template's AR filter (see ar_wave).
 * Every sample of these steps is read before any is written. FILTER =
 * false is the prologue: the history of the 16 diagonals before the first
 * filtered one, which hold template values. CHROMA adds the tap on the
 * co-located luma average of subsampling (SSX, SSY), read from the luma
 * template's finished samples. */

template <int LAG, bool CHROMA, int SSX, int SSY, bool FILTER, int S0, int S1>
__device__ __forceinline__ void ar_steps(int16_t *t, const ArPlace &pl, const uint32_t *cp, int luma_coef,
                                         const int16_t *tl, bool luma_tap, int ars, int gmin, int gmax,
                                         uint32_t (&PO)[AR_ROUND], uint32_t (&PN1)[AR_ROUND],
                                         uint32_t (&PN2)[AR_ROUND], uint32_t (&PN3)[AR_ROUND], int &prev) {
  constexpr ArTaps<LAG> TP{};
  constexpr int NP = ArTaps<LAG>::NP;
  constexpr int N = S1 - S0;
  lds_i16 *ts = (lds_i16 *)(t + pl.base);
  const int rc = 1 << (ars - 1);
  int tv[N], lv[N];
#pragma unroll
  for (int k = 0; k < N; k++) {
    tv[k] = ts[S0 + k];
    lv[k] = 0;
  }
  if (CHROMA && FILTER && luma_tap) {
    /* only filtered steps use the tap; with none the reads go to index 0 */
    const int lb = pl.mask ? (((pl.y - 3) << SSY) + 3) * TW + ((pl.x0 - 3) << SSX) + 3 : 0;
    const int16_t *q = tl + lb;
#pragma unroll
    for (int k = 0; k < N; k++) {
      const int o = (S0 + k) << SSX;
      int l = q[o];
      if (SSX) l += q[o + 1];
      if (SSY) l += q[o + TW];
      if (SSX && SSY) l += q[o + TW + 1];
      lv[k] = round2(l, SSX + SSY);
    }
  }
#pragma unroll
  for (int k = 0; k < N; k++) {
    const int s = S0 + k;
    int out = tv[k];
    if constexpr (FILTER) {
      /* three independent sums of the older pairs, then the two latest
       * pairs (rows 0 and 1, diagonals T - 2 and T - 1) on two of them,
       * so only one dot product and one add follow the latest values */
      /* tv + (sum >> ars) == (sum + (tv << ars)) >> ars: the sample
       * starts the first sum with the rounding constant; the second sum
       * (and the third, for luma) starts with a zero-addend dot product */
      int acc[3] = {lshl_add(tv[k], ars, rc), 0, CHROMA ? __mul24(lv[k], luma_coef) : 0};
      bool started[3] = {true, false, CHROMA};
#pragma unroll
      for (int i = 0; i < NP; i++) {
        const int idx = (s - TP.m[i]) & (AR_ROUND - 1);
        const uint32_t pv = TP.d[i] == 0 ? PO[idx] : TP.d[i] == 1 ? PN1[idx] : TP.d[i] == 2 ? PN2[idx] : PN3[idx];
        const int a = i >= NP - 2 ? i - (NP - 2) : i % 3;
        acc[a] = started[a] ? dot2(pv, cp[i], acc[a]) : dot2_0(pv, cp[i]);
        started[a] = true;
      }
      const int val = med3(gmin, gmax, (acc[0] + acc[1] + acc[2]) >> ars);
      const int am = __builtin_amdgcn_sbfe((int)pl.mask, s, 1); /* -1 where filtered */
      out = bfi(am, val, tv[k]);
      ts[s] = (int16_t)out;
    }
    const int j = (s - 1) & (AR_ROUND - 1);
    PO[j] = pack16(prev, out);
    prev = out;
    if (LAG >= 1) PN1[j] = (uint32_t)from_lane_below((int)PO[j]);
    if (LAG >= 2) PN2[j] = (uint32_t)from_lane_below((int)PN1[j]);
    if (LAG >= 3) PN3[j] = (uint32_t)from_lane_below((int)PN2[j]);
  }
}

/* The auto-regressive filter of 7.18.3.3 (grain.rs:57-73 luma, 101-134
 * chroma) over one template, on one wave. Samples are visited along the
 * diagonals T = K*y + x (K = LAG + 1): every sample (y - d, x + dc) that
 * (y, x) reads lies on an earlier diagonal, T - (d*K - dc).
 *
 * Each lane emits one value per diagonal (ar_place): the filtered sample
 * inside the filtered area, the unfiltered value elsewhere. It keeps its
 * last 16 emitted values as packed pairs of consecutive diagonals (PO[j] =
 * (value of j, value of j + 1), j mod 16) and the same pairs of the lanes
 * 1, 2 and 3 below it (PN1, PN2, PN3, each filled by one lane rotation of
 * the previous array per diagonal), which are exactly the rows above its
 * row; the sum of a sample is then NP dot products of pairs with packed
 * coefficients, in registers whose indices are constants of the unrolled
 * steps. The filtered value replaces the sample only inside the filtered
 * area (a bitfield select, no branch).
 *
 * The diagonals run from the first filtered one in chunks of AR_ROUND,
 * after a prologue that loads the 16 diagonals before it; unit u of the
 * wave (a chunk, or half a chunk for subsampled chroma) runs in barrier
 * round first + u, then writer(r) copies rows out. coef holds the 2 LAG K
 * neighbour coefficients and, for chroma, the luma one. */
template <int LAG, bool CHROMA, int SSX, int SSY, typename Writer>
__device__ void ar_wave(int16_t *t, int rows, int cols, const int *coef, const int16_t *tl, bool luma_tap, int ars,
                        int gmin, int gmax, int first, int halves, int rounds, Writer &writer) {
  constexpr int K = LAG + 1;
  constexpr ArTaps<LAG> TP{};
  constexpr int NP = ArTaps<LAG>::NP;
  constexpr int NL = 2 * LAG * K;
  constexpr int TF = ar_first(K);
  uint32_t cp[NP];
#pragma unroll
  for (int i = 0; i < NP; i++) cp[i] = pack16(TP.lo[i] >= 0 ? coef[TP.lo[i]] : 0, coef[TP.hi[i]]);
  const int luma_coef = CHROMA ? coef[NL] : 0;
  uint32_t PO[AR_ROUND], PN1[AR_ROUND], PN2[AR_ROUND], PN3[AR_ROUND];
#pragma unroll
  for (int i = 0; i < AR_ROUND; i++) PO[i] = PN1[i] = PN2[i] = PN3[i] = 0;
  int prev = 0; /* value this lane emitted on the previous diagonal */
  {
    const ArPlace pl = ar_place<K>(TF - AR_ROUND, rows, cols);
    ar_steps<LAG, CHROMA, SSX, SSY, false, 0, AR_ROUND>(t, pl, cp, luma_coef, tl, luma_tap, ars, gmin, gmax, PO,
                                                        PN1, PN2, PN3, prev);
  }
  const int units = ar_chunks(K, rows, cols) * halves;
  for (int r = 0; r < rounds; r++) {
    const int u = r - first;
    if (u >= 0 && u < units) {
      if (halves == 1) {
        const ArPlace pl = ar_place<K>(TF + AR_ROUND * u, rows, cols);
        ar_steps<LAG, CHROMA, SSX, SSY, true, 0, AR_ROUND>(t, pl, cp, luma_coef, tl, luma_tap, ars, gmin, gmax, PO,
                                                           PN1, PN2, PN3, prev);
      } else {
        const ArPlace pl = ar_place<K>(TF + AR_ROUND * (u >> 1), rows, cols);
        if ((u & 1) == 0)
          ar_steps<LAG, CHROMA, SSX, SSY, true, 0, AR_ROUND / 2>(t, pl, cp, luma_coef, tl, luma_tap, ars, gmin, gmax,
                                                                 PO, PN1, PN2, PN3, prev);
        else
          ar_steps<LAG, CHROMA, SSX, SSY, true, AR_ROUND / 2, AR_ROUND>(t, pl, cp, luma_coef, tl, luma_tap, ars, gmin,
                                                                        gmax, PO, PN1, PN2, PN3, prev);
      }
    }
    writer(r);
    __syncthreads();
  }
}

/* The round in which chroma unit 0 can run: unit u (diagonals a .. b of
 * TF + L u .. TF + L u + L - 1, L = 16 / halves) reads the luma samples
 * co-located with its filtered samples, the latest of which lies on luma
 * diagonal 2 b - 2 K - 2 (4:2:0), b + cols - 6 (4:2:2) or b (4:4:4), or on
 * the last filtered luma diagonal at most; luma chunk i, run in round i,
 * finishes diagonals through TF + 16 i + 15, so the unit can run in any
 * round after that chunk's. */
__device__ __forceinline__ int chroma_first_round(int K, int ssx, int ssy, int cols, int units, int halves) {
  const int TF = ar_first(K), L = AR_ROUND / halves;
  const int last_luma = K * (TH - 1) + TW - 4;
  int first = 0;
  for (int u = 0; u < units; u++) {
    const int b = TF + L * u + L - 1;
    int d = ssy ? 2 * b - 2 * K - 2 : ssx ? b + cols - 6 : b;
    d = min(d, last_luma);
    const int dep = d >= TF ? (d - TF) / AR_ROUND : -1;
    first = max(first, dep + 1 - u);
  }
  return first;
}

/* Rows y0 .. y1 - 1 of one template (in LDS, row pitch TW) into all 8
 * shifted copies of its scratch plane: copy s at column x holds sample
 * x + s, 0 past the row (GRAIN_TWS columns per row). Item (row, 8-column
 * chunk) per lane, lanes lane, lane + nl, ... */
__device__ __forceinline__ void write_rows(const int16_t *t, int16_t *dst, int y0, int y1, int lane, int nl) {
  constexpr int Q = GRAIN_TWS / 8;
  for (int i = lane; i < (y1 - y0) * Q; i += nl) {
    const int r = i / Q, q = i - r * Q, y = y0 + r;
    const uint32_t *row = reinterpret_cast<const uint32_t *>(t + y * TW) + q * 4; /* TW and TFRONT even */
    uint32_t w[8];
#pragma unroll
    for (int k = 0; k < 8; k++) w[k] = q * 8 + 2 * k < TW ? row[k] : 0u;
    int16_t *d = dst + y * GRAIN_TWS + q * 8;
#pragma unroll
    for (int sft = 0; sft < 8; sft++) {
      uint32_t c[4];
#pragma unroll
      for (int k = 0; k < 4; k++) {
        const int a2 = (sft >> 1) + k;
        c[k] = (sft & 1) ? __builtin_amdgcn_alignbyte(w[a2 + 1], w[a2], 2) : w[a2];
      }
      *reinterpret_cast<uint4 *>(d + sft * GRAIN_TCOPY) = make_uint4(c[0], c[1], c[2], c[3]);
    }
  }
}

/* Rows of a template that are final once its diagonals through D are
 * filtered: rows 0-2 are never filtered, row y >= 3 is final after its
 * last filtered sample, on diagonal K y + cols - 4. */
__device__ __forceinline__ int rows_final(int K, int D, int rows, int cols) {
  const int n = D >= cols - 4 ? (D - (cols - 4)) / K + 1 : 0;
  return min(max(n, 3), rows);
}

/* The AR stage of the three templates: wave 0 luma, waves 1 and 2 cb and
 * cr, each chroma unit in a round after the luma chunks it reads; wave 3
 * writes the template rows that became final by the previous round into
 * the scratch (wl luma and wc chroma rows are written when it returns).
 * Lag 0 has no recursion: luma samples are only clipped and each chroma
 * sample depends on the finished luma alone; its rows are all written
 * afterwards. */
template <int LAG>
__device__ void ar_all(int16_t *tl, int16_t *tcb, int16_t *tcr, const GrainHdr &h, const GrainScratch &sc,
                       bool chroma, int cw, int ch, int gmin, int gmax, int &wl, int &wc) {
  constexpr int K = LAG + 1;
  const int wave = threadIdx.x >> 6;
  wl = wc = 0;
  if constexpr (LAG == 0) {
    for (int i = threadIdx.x; i < 70 * 76; i += blockDim.x) {
      const int y = 3 + i / 76, x = 3 + i % 76;
      tl[y * TW + x] = (int16_t)clip3(gmin, gmax, tl[y * TW + x]);
    }
    __syncthreads();
    if (chroma) {
      const int rw = cw - 6, nr = ch - 3;
      for (int i = threadIdx.x; i < rw * nr; i += blockDim.x) chroma_lag0(tcb, tcr, tl, h, 3 + i / rw, 3 + i % rw, gmin, gmax);
    }
    __syncthreads();
  } else {
    /* vertically subsampled chroma reads two luma diagonals per diagonal:
     * it runs half a chunk per round */
    const int halves = h.ssy ? 2 : 1;
    const bool luma_tap = h.num_y > 0;
    const int lr = ar_chunks(K, TH, TW);
    const int cu = chroma ? ar_chunks(K, ch, cw) * halves : 0;
    const int first = chroma && luma_tap ? chroma_first_round(K, h.ssx, h.ssy, cw, cu, halves) : 0;
    const int rounds = max(lr, first + cu);
    /* rows final after the luma chunks and chroma units of rounds 0 .. r - 1 */
    constexpr int TF = ar_first(K);
    const int L = AR_ROUND / halves;
    auto final_l = [&](int r) { return rows_final(K, TF + AR_ROUND * min(r, lr) - 1, TH, TW); };
    auto final_c = [&](int r) {
      return chroma ? rows_final(K, TF + L * min(max(r - first, 0), cu) - 1, ch, cw) : 0;
    };
    wl = final_l(rounds - 1);
    wc = final_c(rounds - 1);
    /* Wave 3 copies the rows final by the previous round out while the
     * chroma waves set the pace (full-height chroma: 4:4:4, 4:2:2); where
     * the luma wave does (4:2:0, monochrome) the copying slowed the rounds
     * by more than it saves, and all rows are copied afterwards. */
    const bool behind = chroma && !h.ssy;
    if (!behind) wl = wc = 0;
    const int lane = threadIdx.x & 63;
    int dl = 0, dc = 0;
    auto writer = [&](int r) {
      if (behind && wave == 3) {
        const int fl = final_l(r);
        write_rows(tl, sc.tmpl, dl, fl, lane, 64);
        dl = fl;
        const int fc = final_c(r);
        write_rows(tcb, sc.tmpl + GRAIN_TPLANE, dc, fc, lane, 64);
        write_rows(tcr, sc.tmpl + 2 * GRAIN_TPLANE, dc, fc, lane, 64);
        dc = fc;
      }
    };
    if (wave == 0) {
      ar_wave<LAG, false, 0, 0>(tl, TH, TW, h.ar_y, tl, false, h.ar_shift, gmin, gmax, 0, 1, rounds, writer);
    } else if ((wave == 1 || wave == 2) && chroma) {
      const int *cf = wave == 1 ? h.ar_cb : h.ar_cr;
      int16_t *t = wave == 1 ? tcb : tcr;
      if (h.ssx && h.ssy)
        ar_wave<LAG, true, 1, 1>(t, ch, cw, cf, tl, luma_tap, h.ar_shift, gmin, gmax, first, halves, rounds, writer);
      else if (h.ssx)
        ar_wave<LAG, true, 1, 0>(t, ch, cw, cf, tl, luma_tap, h.ar_shift, gmin, gmax, first, halves, rounds, writer);
      else
        ar_wave<LAG, true, 0, 0>(t, ch, cw, cf, tl, luma_tap, h.ar_shift, gmin, gmax, first, halves, rounds, writer);
    } else {
      for (int r = 0; r < rounds; r++) {
        writer(r);
        __syncthreads();
      }
    }
  }
}

/* Each template in LDS has TFRONT samples before it (a lane of the AR
 * prologue may start up to 3 K - 3 samples before row 0's) and 64 spare
 * samples after it (the 16-sample run of lanes whose row is past the
 * template, and the luma-tap reads past the last row of lanes with no
 * filtered step). */
constexpr int TFRONT = 16;
constexpr int TPAD = TFRONT + TH * TW + 64;

struct alignas(16) GaussTable {
  int16_t v[2048];
};

constexpr GaussTable make_gauss() {
  GaussTable g{};
  for (int i = 0; i < 2048; i++) g.v[i] = AV1_GAUSSIAN_SEQUENCE[i];
  return g;
}

constexpr GaussTable GAUSS = make_gauss();

struct PrepShared {
  int16_t tl[TPAD], tcb[TPAD], tcr[TPAD];
  alignas(16) int16_t gauss[2048];
};

/* Phase 1 of the prepare stage: the Gaussian templates into LDS. An item
 * is FILL_SEG columns of a template row (two per 82-column row, one per
 * 44-column row), one per thread; the first item's LFSR state is loaded
 * with the Gaussian table's copy into LDS in flight. */
__device__ __forceinline__ void prep_fill(const GrainHdr &h, PrepShared &sm) {
  const int shift = 12 - h.bd + h.grain_scale_shift;
  const bool chroma = h.nplanes > 1;
  const int cw = h.ssx ? 44 : 82, ch = h.ssy ? 38 : 73;
  const int csegs = cw == 82 ? 2 : 1;
  const int litems = 2 * TH, nitems = litems + (chroma ? 2 * ch * csegs : 0);
  struct Item {
    int16_t *t;
    uint32_t seed;
    int y, g, cols;
    bool on;
  };
  auto item = [&](int i) -> Item {
    if (i < litems) return {sm.tl + TFRONT, h.seed, i >> 1, i & 1, TW, h.num_y > 0};
    const int r = i - litems, per = ch * csegs, pl = r >= per, q = r - pl * per;
    return {(pl ? sm.tcr : sm.tcb) + TFRONT, h.seed ^ (pl ? 0x49d8u : 0xb524u), q / csegs, q % csegs, cw,
            (pl ? h.num_cr : h.num_cb) > 0 || h.cfl != 0};
  };
  auto state = [&](const Item &it) -> uint32_t {
    return row_state(it.seed, it.g ? 2 : (it.cols == 82 ? 0 : 1), it.y);
  };
  const int i0 = threadIdx.x;
  const uint32_t s0 = i0 < nitems ? state(item(i0)) : 0u;
  {
    const uint4 *g = reinterpret_cast<const uint4 *>(GAUSS.v);
    reinterpret_cast<uint4 *>(sm.gauss)[threadIdx.x] = g[threadIdx.x]; /* 256 x 16 bytes */
  }
  __syncthreads();
  for (int i = i0; i < nitems; i += PREP_THREADS) {
    const Item it = item(i);
    const uint32_t st = i == i0 ? s0 : state(it);
    const int xb = it.g * FILL_SEG, xe = it.g ? it.cols : min(FILL_SEG, it.cols);
    fill_row(it.t, sm.gauss, st, it.y, xb, xe, shift, it.on);
  }
  __syncthreads();
}

/* Phase 2: the auto-regressive filter of the three templates, in LDS. */
__device__ __forceinline__ void prep_ar(const GrainHdr &h, const GrainScratch &sc, PrepShared &sm, int &wl, int &wc) {
  const int bd = h.bd;
  const int gmin = -(128 << (bd - 8)), gmax = (256 << (bd - 8)) - 1 - (128 << (bd - 8));
  const bool chroma = h.nplanes > 1;
  const int cw = h.ssx ? 44 : 82, ch = h.ssy ? 38 : 73;
  int16_t *tl = sm.tl + TFRONT, *tcb = sm.tcb + TFRONT, *tcr = sm.tcr + TFRONT;
  switch (h.ar_lag) {
    case 0: ar_all<0>(tl, tcb, tcr, h, sc, chroma, cw, ch, gmin, gmax, wl, wc); break;
    case 1: ar_all<1>(tl, tcb, tcr, h, sc, chroma, cw, ch, gmin, gmax, wl, wc); break;
    case 2: ar_all<2>(tl, tcb, tcr, h, sc, chroma, cw, ch, gmin, gmax, wl, wc); break;
    default: ar_all<3>(tl, tcb, tcr, h, sc, chroma, cw, ch, gmin, gmax, wl, wc); break;
  }
}

/* Phase 3: the template rows the AR stage left (rows wl .. 72 of luma, wc
 * .. ch - 1 of chroma) into the 8 shifted copies, on every thread. */
__device__ __forceinline__ void prep_out(const GrainHdr &h, const GrainScratch &sc, PrepShared &sm, int wl, int wc) {
  write_rows(sm.tl + TFRONT, sc.tmpl, wl, TH, threadIdx.x, PREP_THREADS);
  if (h.nplanes > 1) {
    const int ch = h.ssy ? 38 : 73;
    write_rows(sm.tcb + TFRONT, sc.tmpl + GRAIN_TPLANE, wc, ch, threadIdx.x, PREP_THREADS);
    write_rows(sm.tcr + TFRONT, sc.tmpl + 2 * GRAIN_TPLANE, wc, ch, threadIdx.x, PREP_THREADS);
  }
}

/* Workgroups per frame of the tables stage. */
constexpr int TABLE_ITEMS = 8;

/* Item j of the tables stage of one frame (256 threads): its share of
 * the scaling functions and the block offsets of the noise stripes j,
 * j + TABLE_ITEMS, ... */
__device__ __forceinline__ void grain_tables_item(const GrainHdr &h, const GrainScratch &sc, int j, int16_t *l256) {
  /* scaling LUTs (grain.rs:137-165) at the 256 8-bit points, then the
   * scaling function of scale_lut (grain.rs:167-177) at every sample value
   * 0 .. 2^bd - 1, interpolated between them: item j writes values
   * j * 256 + t, j * 256 + t + TABLE_ITEMS * 256, ... of every plane */
  for (int p = 0; p < h.nplanes; p++) {
    const int np = h.npoints[p];
    const int *xs = h.points[p][0], *ys = h.points[p][1];
    const int x = threadIdx.x;
    int v = 0;
    if (np > 0) {
      if (x < xs[0]) {
        v = ys[0];
      } else if (x >= xs[np - 1]) {
        v = ys[np - 1];
      } else {
        int i = 0;
        while (i + 1 < np && xs[i + 1] <= x) i++;
        const int dy = ys[i + 1] - ys[i], dx = xs[i + 1] - xs[i];
        const int delta = dy * ((65536 + (dx >> 1)) / dx);
        v = ys[i] + (((x - xs[i]) * delta + 32768) >> 16);
      }
    }
    l256[p * 256 + x] = (int16_t)v;
  }
  __syncthreads();
  {
    const int sh = h.bd - 8, n = 1 << h.bd;
    for (int p = 0; p < h.nplanes; p++)
      for (int idx = j * 256 + threadIdx.x; idx < n; idx += TABLE_ITEMS * 256) {
        const int16_t *lut = l256 + p * 256;
        const int x = idx >> sh, rem = idx - (x << sh);
        int v = lut[x];
        if (sh && x < 255) v += round2((lut[x + 1] - lut[x]) * rem, sh);
        sc.lut[p * GRAIN_LUT_N + idx] = (int16_t)v;
      }
  }
  /* block offsets per noise stripe (grain.rs:186-199) */
  const int nstripes = ((h.h + 1) / 2 + 15) / 16;
  const int nblocks = ((h.w + 1) / 2 + 15) / 16;
  for (int ln = j * 256 + threadIdx.x; ln < nstripes; ln += TABLE_ITEMS * 256) {
    uint32_t r = h.seed;
    r ^= (uint32_t)(((ln * 37 + 178) & 255) << 8);
    r ^= (uint32_t)((ln * 173 + 105) & 255);
    for (int b = 0; b < nblocks; b++) {
      r = step(r);
      sc.offsets[ln * sc.max_blocks + b] = (uint8_t)((r >> 8) & 255);
    }
  }
}

/* Templates, LUTs and block offsets of one frame (PREP_THREADS threads). */
__device__ __forceinline__ void grain_prepare_body(const GrainHdr &h, const GrainScratch &sc, PrepShared &sm) {
  /* the AR filter is a serial chain on three waves: they win issue
   * arbitration against the filters' waves that share their SIMDs */
  __builtin_amdgcn_s_setprio(3);
  prep_fill(h, sm);
  int wl, wc;
  prep_ar(h, sc, sm, wl, wc);
  __syncthreads();
  prep_out(h, sc, sm, wl, wc);
}


/* The scaling functions of the frame's planes into LDS: plane p's value
 * at sample v at lut[(p << bd) + v], in 16-byte chunks. */
__device__ __forceinline__ void load_luts(const GrainHdr &h, const GrainScratch &sc, int16_t *lut) {
  const int lg = h.bd - 3; /* 8-entry chunks per plane, log2 */
  const int n = h.nplanes << lg;
  for (int c = threadIdx.x; c < n; c += 256) {
    const int p = c >> lg, k = c - (p << lg);
    reinterpret_cast<uint4 *>(lut)[c] = reinterpret_cast<const uint4 *>(sc.lut + p * GRAIN_LUT_N)[k];
  }
}

/* Template blend of two noise values (the horizontal and vertical
 * overlaps of 7.18.3.5): weights 23/22 for a one-sample overlap, 27/17
 * and 17/27 for the two samples of a two-sample overlap. */
__device__ __forceinline__ int overlap_blend(int old, int gv, bool one, int k, int gmin, int gmax) {
  /* |noise| <= 2048: 24-bit products */
  const int wo = one ? 23 : (k == 0 ? 27 : 17), wn = one ? 22 : (k == 0 ? 17 : 27);
  return clip3(gmin, gmax, round2(__mul24(old, wo) + __mul24(gv, wn), 5));
}

/* Template sample (y, x) of a block whose offset byte is o: the block's
 * origin is (9 + 2 oy, 9 + 2 ox), or 6 + o in a subsampled direction. */
__device__ __forceinline__ int origin_x(int o, int psx) { return psx ? 6 + (o >> 4) : 9 + (o >> 4) * 2; }
__device__ __forceinline__ int origin_y(int o, int psy) { return psy ? 6 + (o & 15) : 9 + (o & 15) * 2; }

/* 8 template samples of plane template pt at row y, columns x .. x + 7:
 * one aligned 16-byte load from the copy shifted by x & 7. */
__device__ __forceinline__ void tmpl8(const int16_t *pt, int y, int x, int *o) {
  const uint4 v = *reinterpret_cast<const uint4 *>(pt + (x & 7) * GRAIN_TCOPY + y * GRAIN_TWS + (x & ~7));
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    o[2 * k] = (int)(int16_t)(w[k] & 0xffff);
    o[2 * k + 1] = (int)(int16_t)(w[k] >> 16);
  }
}

/* Noise of 8 consecutive samples (block bx, columns j0 .. j0 + 7 of the
 * block, row i of noise stripe ln): the template samples, blended with the
 * block to the left in the overlap columns, then with the stripe above
 * (itself blended horizontally) in the overlap rows. Every offset and
 * template origin is the same for the 8 samples. */
__device__ __forceinline__ void seg_noise(const int16_t *pt, const uint8_t *offs, int mb, int ln, int bx, int j0,
                                          int i, int psx, int psy, bool ov, int gmin, int gmax, int *g) {
  const int bs = 32 >> psx, bsy = 32 >> psy;
  const int o = offs[ln * mb + bx];
  tmpl8(pt, origin_y(o, psy) + i, origin_x(o, psx) + j0, g);
  const bool hov = ov && bx > 0 && j0 == 0;
  const int nh = psx ? 1 : 2;
  if (hov) {
    const int o2 = offs[ln * mb + bx - 1];
    const int16_t *lp = pt + (origin_y(o2, psy) + i) * GRAIN_TWS + origin_x(o2, psx) + bs;
    for (int k = 0; k < nh; k++) g[k] = overlap_blend(lp[k], g[k], psx, k, gmin, gmax);
  }
  if (ov && ln > 0 && i < (psy ? 1 : 2)) {
    const int ou = offs[(ln - 1) * mb + bx];
    int up[8];
    tmpl8(pt, origin_y(ou, psy) + i + bsy, origin_x(ou, psx) + j0, up);
    if (hov) {
      const int ou2 = offs[(ln - 1) * mb + bx - 1];
      const int16_t *lp = pt + (origin_y(ou2, psy) + i + bsy) * GRAIN_TWS + origin_x(ou2, psx) + bs;
      for (int k = 0; k < nh; k++) up[k] = overlap_blend(lp[k], up[k], psx, k, gmin, gmax);
    }
#pragma unroll
    for (int k = 0; k < 8; k++) g[k] = overlap_blend(up[k], g[k], psy, i, gmin, gmax);
  }
}

/* 8 consecutive samples of a plane row: one 16-byte (u16) or 8-byte (u8)
 * access when the whole segment lies inside the row, sample by sample at
 * the row's end. */
template <typename T>
struct Seg8;
template <>
struct Seg8<uint16_t> {
  static __device__ __forceinline__ void load(const uint16_t *p, int n, int *o) {
    if (n >= 8) {
      const uint4 v = *reinterpret_cast<const uint4 *>(p);
      const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
      for (int k = 0; k < 4; k++) {
        o[2 * k] = w[k] & 0xffff;
        o[2 * k + 1] = w[k] >> 16;
      }
    } else {
      for (int k = 0; k < 8; k++) o[k] = k < n ? p[k] : 0;
    }
  }
  static __device__ __forceinline__ void store(uint16_t *p, int n, const int *o) {
    if (n >= 8) {
      *reinterpret_cast<uint4 *>(p) = make_uint4((uint32_t)o[0] | ((uint32_t)o[1] << 16), (uint32_t)o[2] | ((uint32_t)o[3] << 16),
                                                 (uint32_t)o[4] | ((uint32_t)o[5] << 16), (uint32_t)o[6] | ((uint32_t)o[7] << 16));
    } else {
      for (int k = 0; k < n; k++) p[k] = (uint16_t)o[k];
    }
  }
};
template <>
struct Seg8<uint8_t> {
  static __device__ __forceinline__ void load(const uint8_t *p, int n, int *o) {
    if (n >= 8) {
      const uint2 v = *reinterpret_cast<const uint2 *>(p);
#pragma unroll
      for (int k = 0; k < 4; k++) {
        o[k] = (v.x >> (8 * k)) & 0xff;
        o[4 + k] = (v.y >> (8 * k)) & 0xff;
      }
    } else {
      for (int k = 0; k < 8; k++) o[k] = k < n ? p[k] : 0;
    }
  }
  static __device__ __forceinline__ void store(uint8_t *p, int n, const int *o) {
    if (n >= 8) {
      uint32_t a = 0, b = 0;
#pragma unroll
      for (int k = 0; k < 4; k++) {
        a |= (uint32_t)o[k] << (8 * k);
        b |= (uint32_t)o[4 + k] << (8 * k);
      }
      *reinterpret_cast<uint2 *>(p) = make_uint2(a, b);
    } else {
      for (int k = 0; k < n; k++) p[k] = (uint8_t)o[k];
    }
  }
};

/* 8 consecutive samples of a plane row as loaded, unpacked later so that
 * every load of a tile is in flight before the first is used: one 16-byte
 * (u16) or 8-byte (u8) access when the whole segment lies inside the row,
 * else the n samples present and zeros. */
template <typename T>
struct Raw8;
template <>
struct Raw8<uint16_t> {
  uint4 v;
  __device__ __forceinline__ void load(const uint16_t *p, int n) {
    if (n >= 8) {
      v = *reinterpret_cast<const uint4 *>(p);
    } else {
      uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (k < n) w[k >> 1] |= (uint32_t)p[k] << (16 * (k & 1));
      v = make_uint4(w[0], w[1], w[2], w[3]);
    }
  }
  __device__ __forceinline__ void unpack(int *o) const {
    const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int k = 0; k < 4; k++) {
      o[2 * k] = w[k] & 0xffff;
      o[2 * k + 1] = w[k] >> 16;
    }
  }
};
template <>
struct Raw8<uint8_t> {
  uint2 v;
  __device__ __forceinline__ void load(const uint8_t *p, int n) {
    if (n >= 8) {
      v = *reinterpret_cast<const uint2 *>(p);
    } else {
      uint32_t w[2] = {0u, 0u};
#pragma unroll
      for (int k = 0; k < 8; k++)
        if (k < n) w[k >> 2] |= (uint32_t)p[k] << (8 * (k & 3));
      v = make_uint2(w[0], w[1]);
    }
  }
  __device__ __forceinline__ void unpack(int *o) const {
#pragma unroll
    for (int k = 0; k < 4; k++) {
      o[k] = (v.x >> (8 * k)) & 0xff;
      o[4 + k] = (v.y >> (8 * k)) & 0xff;
    }
  }
};

/* The block offsets a segment's noise reads (grain.rs:186-199 draws them),
 * one byte each in one word: its own block's (o), and in the overlaps
 * those of the block to the left (l), above (u) and above-left (ul). */
struct SegOffs {
  uint32_t w;
  __device__ __forceinline__ int o() const { return w & 0xff; }
  __device__ __forceinline__ int l() const { return (w >> 8) & 0xff; }
  __device__ __forceinline__ int u() const { return (w >> 16) & 0xff; }
  __device__ __forceinline__ int ul() const { return w >> 24; }
};

__device__ __forceinline__ SegOffs seg_offs(const uint8_t *offs, int mb, int ln, int bx, bool hov, bool vov) {
  const uint8_t *row = offs + ln * mb + bx;
  const uint32_t o = row[0], l = hov ? row[-1] : 0u, u = vov ? row[-mb] : 0u, ul = hov && vov ? row[-mb - 1] : 0u;
  return SegOffs{o | (l << 8) | (u << 16) | (ul << 24)};
}

/* Template sample (y, x) .. (y, x + 7) of plane template pt: the copy
 * shifted by x & 7 makes them one aligned 16-byte word. */
__device__ __forceinline__ const int16_t *tmpl_at(const int16_t *pt, int y, int x) {
  return pt + (x & 7) * GRAIN_TCOPY + y * GRAIN_TWS + (x & ~7);
}

/* The template words of a segment (block row i, block columns j0 ..
 * j0 + 7): its own 8 samples and, in the horizontal overlap, the left
 * block's samples just past its width (the two columns it blends with). */
struct SegTaps {
  uint4 m;
  uint32_t l;
};

__device__ __forceinline__ SegTaps seg_taps(const int16_t *pt, const SegOffs &s, int i, int j0, int psx, int psy,
                                            bool hov) {
  SegTaps t;
  t.m = *reinterpret_cast<const uint4 *>(tmpl_at(pt, origin_y(s.o(), psy) + i, origin_x(s.o(), psx) + j0));
  t.l = hov ? *reinterpret_cast<const uint32_t *>(
                  tmpl_at(pt, origin_y(s.l(), psy) + i, origin_x(s.l(), psx) + (32 >> psx)))
            : 0u;
  return t;
}

__device__ __forceinline__ void unpack_s16(const uint4 &v, int *g) {
  const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    g[2 * k] = (int)(int16_t)(w[k] & 0xffff);
    g[2 * k + 1] = (int)w[k] >> 16;
  }
}

__device__ __forceinline__ int lo16(uint32_t w) { return (int)(int16_t)(w & 0xffff); }
__device__ __forceinline__ int hi16(uint32_t w) { return (int)w >> 16; }

/* The horizontal overlap (7.18.3.5) of the first word of a row segment
 * (its first two samples, or one when subsampled) with the two samples
 * of the left block in word l. */
__device__ __forceinline__ uint32_t hblend_word(uint32_t w, uint32_t l, int psx, int gmin, int gmax) {
  const int b0 = overlap_blend(lo16(l), lo16(w), psx, 0, gmin, gmax);
  const int b1 = psx ? hi16(w) : overlap_blend(hi16(l), hi16(w), psx, 1, gmin, gmax);
  return pack16(b0, b1);
}

/* Noise of a segment from its template words (seg_noise's blends,
 * 7.18.3.5), as 4 words of two samples: the horizontal overlap with the
 * left block, then, in the first rows of a stripe, the vertical one with
 * the stripe above (its words are read here: only the lanes of those rows
 * get there). The words stay packed, so the blends hold few registers. */
__device__ __forceinline__ uint4 seg_blend(const int16_t *pt, const SegOffs &s, const SegTaps &t, int i, int j0,
                                           int psx, int psy, bool hov, bool vov, int gmin, int gmax) {
  const int bs = 32 >> psx, bsy = 32 >> psy;
  uint32_t g[4] = {t.m.x, t.m.y, t.m.z, t.m.w};
  if (hov) g[0] = hblend_word(g[0], t.l, psx, gmin, gmax);
  if (vov) {
    const uint4 uv =
        *reinterpret_cast<const uint4 *>(tmpl_at(pt, origin_y(s.u(), psy) + i + bsy, origin_x(s.u(), psx) + j0));
    uint32_t u[4] = {uv.x, uv.y, uv.z, uv.w};
    if (hov)
      u[0] = hblend_word(
          u[0],
          *reinterpret_cast<const uint32_t *>(tmpl_at(pt, origin_y(s.ul(), psy) + i + bsy, origin_x(s.ul(), psx) + bs)),
          psx, gmin, gmax);
#pragma unroll
    for (int w = 0; w < 4; w++)
      g[w] = pack16(overlap_blend(lo16(u[w]), lo16(g[w]), psy, i, gmin, gmax),
                    overlap_blend(hi16(u[w]), hi16(g[w]), psy, i, gmin, gmax));
  }
  return make_uint4(g[0], g[1], g[2], g[3]);
}

/* A chroma segment of the noise-applying stage through its three rounds:
 * chroma_load (its samples, the luma it averages, its block offsets),
 * chroma_index once they have landed (the packed index into the scaling
 * function, 7.18.3.5, so the luma is dead before the next round),
 * chroma_taps (its template words) and chroma_out. */
template <typename T>
struct ChromaSeg {
  int p, x, y, i; /* plane, column, row, row in the noise stripe */
  bool on, hov, vov, noise;
  Raw8<T> raw, l0, l1;
  SegOffs so;
  uint32_t m[4];
  SegTaps t;
};

/* Item it of the tile (columns x0 .. of noise stripe ln) of an LS tile:
 * per items per plane, segs = 1 << lsegs of them per row. per is a power
 * of two >= 64, so a wave's items all lie in one plane and the plane and
 * its parameters are uniform. */
template <typename T>
__device__ __forceinline__ void chroma_load(ChromaSeg<T> &c, const GrainHdr &h, const GrainScratch &sc,
                                            const PlanePtr &i0, const PlanePtr &i1, const PlanePtr &i2, int it,
                                            int x0, int y0, int ln, int lsegs, int per) {
  const int ssx = h.ssx, ssy = h.ssy;
  const int cw = (h.w + ssx) >> ssx, ch = (h.h + ssy) >> ssy, cbs = 32 >> ssx;
  c.p = __builtin_amdgcn_readfirstlane(1 + (it >= per));
  const int r = it - (c.p - 1) * per;
  c.i = r >> lsegs;
  c.x = (x0 >> ssx) + ((r & ((1 << lsegs) - 1)) << 3);
  c.y = (y0 >> ssy) + c.i;
  c.on = it < 2 * per && c.x < cw && c.y < ch;
  c.hov = h.overlap && c.x >= cbs && (c.x & (cbs - 1)) == 0;
  c.vov = h.overlap && ln > 0 && c.i < (ssy ? 1 : 2);
  c.noise = (c.p == 1 ? h.num_cb : h.num_cr) > 0 || h.cfl;
  if (!c.on) return;
  const PlanePtr &ip = c.p == 1 ? i1 : i2;
  c.raw.load(gptr(reinterpret_cast<const T *>(ip.data)) + (int64_t)c.y * ip.stride + c.x, cw - c.x);
  if (!c.noise) return;
  const T *lrow = gptr(reinterpret_cast<const T *>(i0.data)) + (int64_t)(c.y << ssy) * i0.stride + (c.x << ssx);
  const int nl = h.w - (c.x << ssx);
  c.l0.load(lrow, nl);
  if (ssx) c.l1.load(lrow + 8, nl - 8);
  c.so = seg_offs(sc.offsets, sc.max_blocks, ln, c.x >> (5 - ssx), c.hov, c.vov);
}

template <typename T>
__device__ __forceinline__ void chroma_index(ChromaSeg<T> &c, const GrainHdr &h) {
  if (!c.on || !c.noise) return;
  const int ssx = h.ssx, bd = h.bd, maxv = (1 << bd) - 1;
  int orig[8], avg[8], m[8];
  c.raw.unpack(orig);
  if (ssx) {
    int l[16];
    c.l0.unpack(l);
    c.l1.unpack(l + 8);
    const int nl = h.w - (c.x << 1);
    if (nl >= 16) {
#pragma unroll
      for (int k = 0; k < 8; k++) avg[k] = (l[2 * k] + l[2 * k + 1] + 1) >> 1;
    } else {
#pragma unroll
      for (int k = 0; k < 8; k++) {
        /* the right neighbour of the last luma column is itself */
        const int b = 2 * k + 1 < nl ? l[2 * k + 1] : l[2 * k];
        avg[k] = (l[2 * k] + b + 1) >> 1;
      }
    }
  } else {
    c.l0.unpack(avg);
  }
  if (h.cfl) {
#pragma unroll
    for (int k = 0; k < 8; k++) m[k] = avg[k];
  } else {
    /* |weights| <= 128, samples <= 4095: 24-bit products */
    const int wl = (c.p == 1 ? h.cb_luma_mult : h.cr_luma_mult) - 128;
    const int wc = (c.p == 1 ? h.cb_mult : h.cr_mult) - 128;
    const int off = ((c.p == 1 ? h.cb_offset : h.cr_offset) - 256) * (1 << (bd - 8));
#pragma unroll
    for (int k = 0; k < 8; k++) m[k] = med3(0, maxv, ((__mul24(avg[k], wl) + __mul24(orig[k], wc)) >> 6) + off);
  }
#pragma unroll
  for (int k = 0; k < 4; k++) c.m[k] = (uint32_t)m[2 * k] | ((uint32_t)m[2 * k + 1] << 16);
}

template <typename T>
__device__ __forceinline__ void chroma_taps(ChromaSeg<T> &c, const GrainHdr &h, const GrainScratch &sc) {
  if (c.on && c.noise)
    c.t = seg_taps(sc.tmpl + (size_t)c.p * GRAIN_TPLANE, c.so, c.i, c.x & ((32 >> h.ssx) - 1), h.ssx, h.ssy, c.hov);
}

/* noised sample: orig + round2(scale * noise, shift), clipped; the scale
 * is at most 255 and |noise| <= 2048, so the product is one 24-bit
 * multiply-add; the 8 scaling reads go out together */
__device__ __forceinline__ void add_noise8(const int16_t *pl, const int *idx, const int *orig, const int *z, int rnd,
                                           int sh, int lo, int hi, int *v) {
  int sv[8];
#pragma unroll
  for (int k = 0; k < 8; k++) sv[k] = pl[idx[k]];
#pragma unroll
  for (int k = 0; k < 8; k++) v[k] = med3(lo, hi, orig[k] + ((__mul24(sv[k], z[k]) + rnd) >> sh));
}

template <typename T, typename TO>
__device__ __forceinline__ void chroma_out(const ChromaSeg<T> &c, const GrainHdr &h, const GrainScratch &sc,
                                           const PlanePtr &o1, const PlanePtr &o2, int16_t *lut) {
  if (!c.on) return;
  const int ssx = h.ssx, ssy = h.ssy, bd = h.bd;
  const int cw = (h.w + ssx) >> ssx;
  int orig[8], v[8];
  c.raw.unpack(orig);
  if (c.noise) {
    const int gmin = -(128 << (bd - 8)), gmax = (256 << (bd - 8)) - 1 - (128 << (bd - 8));
    int z[8], idx[8];
    unpack_s16(seg_blend(sc.tmpl + (size_t)c.p * GRAIN_TPLANE, c.so, c.t, c.i, c.x & ((32 >> ssx) - 1), ssx, ssy,
                         c.hov, c.vov, gmin, gmax),
               z);
#pragma unroll
    for (int k = 0; k < 8; k++) idx[k] = (c.m[k >> 1] >> (16 * (k & 1))) & 0xffff;
    const int sh = h.scaling_shift;
    add_noise8(lut + (c.p << bd), idx, orig, z, (1 << sh) >> 1, sh, h.min_value, h.max_chroma, v);
  } else {
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = orig[k];
  }
  const PlanePtr &op = c.p == 1 ? o1 : o2;
  Seg8<TO>::store(gptr(reinterpret_cast<TO *>(op.data)) + (int64_t)c.y * op.stride + c.x, cw - c.x, v);
}

/* One tile of the noise-applying stage: noise stripe ln (32 luma rows) and
 * luma columns [x0, x0 + 64 LS), with the chroma samples they cover. A
 * thread takes LS 8-sample row segments of luma, 64 columns apart (8 per
 * row and 64 columns, 32 rows), and CI chroma segments per pass (items
 * threadIdx.x + 256 (u + CI pass) of the 2 per items of both planes:
 * one pass, or two for 4:4:4 with CI = 1). LS = 2 for 4:2:0 gives every
 * thread one chroma segment. Every segment lies in one 32-column noise
 * block, so its noise is 8 consecutive template samples. The first
 * pass's accesses go out in three rounds, each before anything waits on
 * it: the plane samples and block offsets (with the scaling tables' copy
 * into LDS), then the template words the offsets address, then, in the
 * first rows of a stripe only, the stripe above's. The planes' rows must
 * be 16-byte (u16) or 8-byte (u8) aligned. Chroma reads the un-noised luma
 * of the input, as the reference blends chroma first. */
template <typename T, typename TO, int CI, int LS, bool TWO_PASS = false>
__device__ __forceinline__ void grain_apply_tile(const GrainHdr &h, const GrainScratch &sc, const PlanePtr &i0,
                                                 const PlanePtr &i1, const PlanePtr &i2, const PlanePtr &o0,
                                                 const PlanePtr &o1, const PlanePtr &o2, int tile_id, int16_t *lut) {
  constexpr int TW = 64 * LS; /* luma columns per tile */
  const int ntx = (h.w + TW - 1) / TW;
  const int ln = tile_id / ntx, x0 = (tile_id - ln * ntx) * TW, y0 = ln * 32;
  const int bd = h.bd;
  const int gmin = -(128 << (bd - 8)), gmax = (256 << (bd - 8)) - 1 - (128 << (bd - 8));
  const bool ov = h.overlap;
  const T *lin = gptr(reinterpret_cast<const T *>(i0.data));
  const int64_t ls = i0.stride;

  /* round 1 */
  const int lx0 = x0 + (threadIdx.x & 7) * 8, li = threadIdx.x >> 3, ly = y0 + li;
  const bool lvov = ov && ln > 0 && li < 2;
  const bool noise_y = h.num_y > 0;
  bool lon[LS], lhov[LS];
  Raw8<T> lraw[LS];
  SegOffs lso[LS];
#pragma unroll
  for (int q = 0; q < LS; q++) {
    const int lx = lx0 + 64 * q;
    lon[q] = lx < h.w && ly < h.h;
    lhov[q] = ov && lx >= 32 && (lx & 31) == 0;
    if (lon[q]) {
      lraw[q].load(lin + (int64_t)ly * ls + lx, h.w - lx);
      if (noise_y) lso[q] = seg_offs(sc.offsets, sc.max_blocks, ln, lx >> 5, lhov[q], lvov);
    }
  }
  const bool chroma = h.nplanes > 1;
  /* chroma segments per tile row (4, 8 or 16) and per plane */
  const int lsegs = chroma ? 3 + (LS == 2) - h.ssx : 0, per = (1 << lsegs) * (32 >> (chroma ? h.ssy : 0));
  ChromaSeg<T> cs[CI];
#pragma unroll
  for (int u = 0; u < CI; u++) {
    cs[u].on = false;
    if (chroma) chroma_load(cs[u], h, sc, i0, i1, i2, threadIdx.x + 256 * u, x0, y0, ln, lsegs, per);
  }
  load_luts(h, sc, lut);
  __syncthreads(); /* round 1 has landed */
#pragma unroll
  for (int u = 0; u < CI; u++) chroma_index(cs[u], h);

  /* round 2 */
  SegTaps lt[LS];
#pragma unroll
  for (int q = 0; q < LS; q++)
    if (lon[q] && noise_y) lt[q] = seg_taps(sc.tmpl, lso[q], li, (lx0 + 64 * q) & 31, 0, 0, lhov[q]);
#pragma unroll
  for (int u = 0; u < CI; u++) chroma_taps(cs[u], h, sc);

  /* luma out first: its samples and noise are dead before the chroma */
  const int sh = h.scaling_shift, rnd = (1 << sh) >> 1;
#pragma unroll
  for (int q = 0; q < LS; q++) {
    if (!lon[q]) continue;
    const int lx = lx0 + 64 * q;
    int orig[8], v[8];
    lraw[q].unpack(orig);
    if (noise_y) {
      int z[8];
      unpack_s16(seg_blend(sc.tmpl, lso[q], lt[q], li, lx & 31, 0, 0, lhov[q], lvov, gmin, gmax), z);
      add_noise8(lut, orig, orig, z, rnd, sh, h.min_value, h.max_luma, v);
    } else {
#pragma unroll
      for (int k = 0; k < 8; k++) v[k] = orig[k];
    }
    Seg8<TO>::store(gptr(reinterpret_cast<TO *>(o0.data)) + (int64_t)ly * o0.stride + lx, h.w - lx, v);
  }
#pragma unroll
  for (int u = 0; u < CI; u++) chroma_out<T, TO>(cs[u], h, sc, o1, o2, lut);

  /* a second pass (TWO_PASS: 4:4:4 with CI = 1, which only the batched
   * stage runs; every other layout has at most 256 CI items): the same
   * rounds, one after another */
  if constexpr (TWO_PASS)
  for (int it0 = 256 * CI; chroma && it0 < 2 * per; it0 += 256 * CI) {
#pragma unroll
    for (int u = 0; u < CI; u++) {
      chroma_load(cs[u], h, sc, i0, i1, i2, it0 + threadIdx.x + 256 * u, x0, y0, ln, lsegs, per);
      chroma_index(cs[u], h);
      chroma_taps(cs[u], h, sc);
      chroma_out<T, TO>(cs[u], h, sc, o1, o2, lut);
    }
  }
}

/* Sample-by-sample variant for planes whose rows are not aligned for the
 * segment accesses: thread t of the tile takes samples t, t + 256, ... of
 * the tile's luma (64 x 32) and then of each chroma plane. */
template <typename T, typename TO>
__device__ __forceinline__ void grain_apply_tile_unaligned(const GrainHdr &h, const GrainScratch &sc,
                                                           const PlanePtr &i0, const PlanePtr &i1, const PlanePtr &i2,
                                                           const PlanePtr &o0, const PlanePtr &o1, const PlanePtr &o2,
                                                           int tile_id, int16_t *lut) {
  load_luts(h, sc, lut);
  __syncthreads();
  const int ntx = (h.w + 63) >> 6;
  const int ln = tile_id / ntx, x0 = (tile_id - ln * ntx) * 64, y0 = ln * 32;
  const int bd = h.bd;
  const int gmin = -(128 << (bd - 8)), gmax = (256 << (bd - 8)) - 1 - (128 << (bd - 8));
  const bool ov = h.overlap;
  const int mb = sc.max_blocks;
  const T *lin = gptr(reinterpret_cast<const T *>(i0.data));
  const int64_t ls = i0.stride;
  const int maxv = (1 << bd) - 1;
  const int np = h.nplanes > 1 ? 3 : 1;
  for (int p = 1; p < np; p++) {
    const int ssx = h.ssx, ssy = h.ssy;
    const int cw = (h.w + ssx) >> ssx, ch = (h.h + ssy) >> ssy;
    const int tw = 64 >> ssx, th = 32 >> ssy;
    const PlanePtr &ip = p == 1 ? i1 : i2;
    const PlanePtr &op = p == 1 ? o1 : o2;
    const int16_t *pt = sc.tmpl + (size_t)p * GRAIN_TPLANE;
    for (int t = threadIdx.x; t < tw * th; t += 256) {
      const int ci = t / tw, c = t - ci * tw;
      const int cx = (x0 >> ssx) + c, cy = (y0 >> ssy) + ci;
      if (cx >= cw || cy >= ch) continue;
      const int orig = gptr(reinterpret_cast<const T *>(ip.data))[(int64_t)cy * ip.stride + cx];
      int v = orig;
      const int npts = p == 1 ? h.num_cb : h.num_cr;
      if (npts > 0 || h.cfl) {
        const T *lrow = lin + (int64_t)(cy << ssy) * ls;
        const int lx = cx << ssx, lnx = min(lx + 1, h.w - 1);
        const int avg = ssx ? round2((int)lrow[lx] + (int)lrow[lnx], 1) : (int)lrow[lx];
        int merged;
        if (h.cfl) {
          merged = avg;
        } else {
          const int lm = p == 1 ? h.cb_luma_mult : h.cr_luma_mult;
          const int mu = p == 1 ? h.cb_mult : h.cr_mult;
          const int of = p == 1 ? h.cb_offset : h.cr_offset;
          const int combined = avg * (lm - 128) + orig * (mu - 128);
          merged = clip3(0, maxv, (combined >> 6) + (of - 256) * (1 << (bd - 8)));
        }
        const int bs = 32 >> ssx;
        int z[8];
        seg_noise(pt, sc.offsets, mb, ln, cx / bs, (cx % bs) & ~7, ci, ssx, ssy, ov, gmin, gmax, z);
        const int sn = round2(lut[(p << bd) + merged] * z[cx & 7], h.scaling_shift);
        v = clip3(h.min_value, h.max_chroma, orig + sn);
      }
      gptr(reinterpret_cast<TO *>(op.data))[(int64_t)cy * op.stride + cx] = (TO)v;
    }
  }
  for (int t = threadIdx.x; t < 64 * 32; t += 256) {
    const int i = t >> 6, x = x0 + (t & 63), y = y0 + i;
    if (x >= h.w || y >= h.h) continue;
    const int orig = lin[(int64_t)y * ls + x];
    int v = orig;
    if (h.num_y > 0) {
      int z[8];
      seg_noise(sc.tmpl, sc.offsets, mb, ln, x >> 5, x & 24, i, 0, 0, ov, gmin, gmax, z);
      const int sn = round2(lut[orig] * z[x & 7], h.scaling_shift);
      v = clip3(h.min_value, h.max_luma, orig + sn);
    }
    gptr(reinterpret_cast<TO *>(o0.data))[(int64_t)y * o0.stride + x] = (TO)v;
  }
}

/* Workgroup 0: the templates (grain_prepare_body); workgroups 1 ..
 * TABLE_ITEMS: the scaling functions and block offsets beside it. */
__global__ __launch_bounds__(PREP_THREADS) void grain_prepare_kernel(GrainHdr h, GrainScratch sc) {
  __shared__ PrepShared sm;
  static_assert(PREP_THREADS == 256, "the tables items and the Gaussian copy take 256 threads");
  if (blockIdx.x == 0)
    grain_prepare_body(h, sc, sm);
  else
    grain_tables_item(h, sc, blockIdx.x - 1, sm.gauss);
}

/* KIND (apply_kind): 1 (segments, 64 luma columns per tile: 4:2:2, mono),
 * 2 (segments, two chroma segments per thread: 4:4:4), 3 (segments, 128
 * luma columns per tile: 4:2:0), 0 (sample by sample). Input samples T,
 * output samples TO (8-bit output from 16-bit planes for the packed 8-bit
 * frame). The luma samples are written before the chroma pass, so kind 3
 * needs 61 VGPRs (8 waves per SIMD). The scaling functions take
 * 2 * (nplanes << bd) bytes of dynamic LDS. */
template <typename T, typename TO, int ALIGNED>
__global__ __launch_bounds__(256) void grain_apply_kernel(GrainHdr h, GrainScratch sc, PlanePtr i0, PlanePtr i1,
                                                          PlanePtr i2, PlanePtr o0, PlanePtr o1, PlanePtr o2) {
  extern __shared__ int16_t lut_dyn[]; /* 2 * (nplanes << bd) bytes */
  int16_t *lut = lut_dyn;
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  if constexpr (ALIGNED != 0)
    grain_apply_tile<T, TO, ALIGNED == 2 ? 2 : 1, ALIGNED == 3 ? 2 : 1>(h, sc, i0, i1, i2, o0, o1, o2, tile, lut);
  else
    grain_apply_tile_unaligned<T, TO>(h, sc, i0, i1, i2, o0, o1, o2, tile, lut);
}

struct PrepRecord {
  GrainHdr h;
  GrainScratch s;
};

struct ApplyRecord {
  GrainHdr h;
  GrainScratch s;
  PlanePtr in[3], out[3];
  int kind; /* apply_kind */
};

template <bool DIRECT>
__global__ __launch_bounds__(PREP_THREADS) void grain_prepare_batch(BatchWork w) {
  __shared__ PrepShared sm;
  batch_loop<DIRECT>(w, [&](int f, int) {
    const PrepRecord &r = batch_table<PrepRecord>(w.rec, f);
    grain_prepare_body(r.h, r.s, sm);
  });
}

/* The tables stage of every frame: TABLE_ITEMS workgroups per frame. */
__global__ __launch_bounds__(256) void grain_tables_batch(BatchWork w) {
  const int f = blockIdx.x / TABLE_ITEMS, j = blockIdx.x - f * TABLE_ITEMS;
  const PrepRecord &r = batch_table<PrepRecord>(w.rec, f);
  __shared__ int16_t l256[3 * 256];
  grain_tables_item(r.h, r.s, j, l256);
}

/* At most 5 waves per SIMD in the grab loop (it takes 73 VGPRs, 6 waves;
 * the bound only caps the persistent form's registers). */
template <typename T, bool DIRECT>
__global__ __launch_bounds__(256, 5) void grain_apply_batch(BatchWork w) {
  __shared__ __align__(16) int16_t lut[3 * GRAIN_LUT_N];
  batch_loop<DIRECT>(w, [&](int f, int j) {
    const ApplyRecord &r = batch_table<ApplyRecord>(w.rec, f);
    if (r.kind == 3)
      grain_apply_tile<T, T, 1, 2>(r.h, r.s, r.in[0], r.in[1], r.in[2], r.out[0], r.out[1], r.out[2], j, lut);
    else if (r.kind != 0)
      grain_apply_tile<T, T, 1, 1, true>(r.h, r.s, r.in[0], r.in[1], r.in[2], r.out[0], r.out[1], r.out[2], j, lut);
    else
      grain_apply_tile_unaligned<T, T>(r.h, r.s, r.in[0], r.in[1], r.in[2], r.out[0], r.out[1], r.out[2], j, lut);
  });
}

/* Tiles of the noise-applying stage: 64 luma columns (128 for kind 3) of
 * one noise stripe. */
int apply_tiles(const GrainHdr &h, int kind) {
  const int tw = kind == 3 ? 128 : 64;
  return ((h.w + tw - 1) / tw) * ((h.h + 31) / 32);
}

/* Every plane row in and out starts on a segment boundary (8 samples). */
bool rows_aligned(const GrainHdr &h, const PlanePtr in[3], const PlanePtr out[3], int in_bps, int out_bps) {
  for (int p = 0; p < h.nplanes; p++) {
    if (((uintptr_t)in[p].data % (8 * (uintptr_t)in_bps)) != 0 || (in[p].stride % 8) != 0) return false;
    if (((uintptr_t)out[p].data % (8 * (uintptr_t)out_bps)) != 0 || (out[p].stride % 8) != 0) return false;
  }
  return true;
}

/* Tile kind of a frame: 0 (rows not aligned: sample by sample), 2 (4:4:4),
 * 3 (4:2:0), 1 (4:2:2 and mono). */
int apply_kind(const GrainHdr &h, const PlanePtr in[3], const PlanePtr out[3], int in_bps, int out_bps) {
  if (!rows_aligned(h, in, out, in_bps, out_bps)) return 0;
  if (h.nplanes == 1) return 1;
  if (!h.ssx) return 2;
  return h.ssy ? 3 : 1;
}

}  // namespace

int launch_grain_prepare(const GrainHdr &h, const GrainScratch &s, hipStream_t st) {
  grain_prepare_kernel<<<1 + TABLE_ITEMS, PREP_THREADS, 0, st>>>(h, s);
  return (int)hipGetLastError();
}

size_t grain_prep_record_size() { return sizeof(PrepRecord); }

int grain_prep_plan(const GrainHdr &h, const GrainScratch &s, void *rec) {
  PrepRecord &r = *static_cast<PrepRecord *>(rec);
  r.h = h;
  r.s = s;
  return 1;
}

int launch_grain_prepare_batch(const BatchWork &w, const StagePlan &sp, hipStream_t st) {
  if (w.total == 0) return 0;
  if (batch_direct(w, sp.grid))
    grain_prepare_batch<true><<<sp.grid, PREP_THREADS, 0, st>>>(w);
  else
    grain_prepare_batch<false><<<sp.grid, PREP_THREADS, 0, st>>>(w);
  /* one prepare item per frame: records are indexed by frame */
  grain_tables_batch<<<w.n * TABLE_ITEMS, 256, 0, st>>>(w);
  return (int)hipGetLastError();
}

size_t grain_apply_record_size() { return sizeof(ApplyRecord); }

int grain_apply_plan(const GrainHdr &h, const GrainScratch &s, const PlanePtr in[3], const PlanePtr out[3],
                     int bps, void *rec) {
  ApplyRecord &r = *static_cast<ApplyRecord *>(rec);
  r.h = h;
  r.s = s;
  for (int p = 0; p < 3; p++) {
    r.in[p] = in[p];
    r.out[p] = out[p];
  }
  r.kind = apply_kind(h, in, out, bps, bps);
  /* the batched kernel runs kind 2 (4:4:4) with one chroma segment per
   * thread and pass */
  if (r.kind == 2) r.kind = 1;
  return apply_tiles(h, r.kind);
}

int launch_grain_apply_batch(const BatchWork &w, const StagePlan &sp, int bps, hipStream_t st) {
  if (w.total == 0) return 0;
  if (bps == 1)
    if (batch_direct(w, sp.grid))
      grain_apply_batch<uint8_t, true><<<sp.grid, 256, 0, st>>>(w);
    else
      grain_apply_batch<uint8_t, false><<<sp.grid, 256, 0, st>>>(w);
  else
    if (batch_direct(w, sp.grid))
      grain_apply_batch<uint16_t, true><<<sp.grid, 256, 0, st>>>(w);
    else
      grain_apply_batch<uint16_t, false><<<sp.grid, 256, 0, st>>>(w);
  return (int)hipGetLastError();
}

template <typename T, typename TO>
void apply_launch(int kind, int grid, const GrainHdr &h, const GrainScratch &s, const PlanePtr in[3],
                  PlanePtr out[3], hipStream_t st) {
  const size_t lds = (size_t)(h.nplanes << h.bd) * sizeof(int16_t); /* the scaling functions */
  if (kind == 1)
    grain_apply_kernel<T, TO, 1><<<grid, 256, lds, st>>>(h, s, in[0], in[1], in[2], out[0], out[1], out[2]);
  else if (kind == 3)
    grain_apply_kernel<T, TO, 3><<<grid, 256, lds, st>>>(h, s, in[0], in[1], in[2], out[0], out[1], out[2]);
  else if (kind == 2)
    grain_apply_kernel<T, TO, 2><<<grid, 256, lds, st>>>(h, s, in[0], in[1], in[2], out[0], out[1], out[2]);
  else
    grain_apply_kernel<T, TO, 0><<<grid, 256, lds, st>>>(h, s, in[0], in[1], in[2], out[0], out[1], out[2]);
}

int launch_grain_apply(const GrainHdr &h, const GrainScratch &s, const PlanePtr in[3], PlanePtr out[3], int bps,
                       hipStream_t st, int out_bps) {
  if (out_bps == 0) out_bps = bps;
  if (out_bps > bps) return (int)hipErrorInvalidValue;
  const int kind = apply_kind(h, in, out, bps, out_bps);
  const int grid = apply_tiles(h, kind);
  if (bps == 1)
    apply_launch<uint8_t, uint8_t>(kind, grid, h, s, in, out, st);
  else if (out_bps == 1)
    apply_launch<uint16_t, uint8_t>(kind, grid, h, s, in, out, st);
  else
    apply_launch<uint16_t, uint16_t>(kind, grid, h, s, in, out, st);
  return (int)hipGetLastError();
}

}  // namespace rvpf
