/*
 * lr.hip — loop restoration (AV1 7.17) for gfx950.
 *
 * Follows rivet-av1's loop_restoration / lr_stripe / lr_unit_cols /
 * restore_stripe / LrCtx::sample / restore_rect_from_window /
 * box_filter_rect / wiener_coefficient (src/decoder/postfilter.rs:572-1004).
 *
 * One workgroup (256 threads, 4 waves) per (plane, stripe, slice): a
 * stripe is 64 luma rows offset by -8, a slice is min(64, unit size)
 * columns, so a slice lies in one restoration unit (the last unit of a row
 * absorbs the remainder: its slices map to it by clamping the unit column).
 * The workgroup stages its window (slice + 3-sample border) in LDS with the
 * stripe rule of 7.17.6 (rows above / below the stripe come from the
 * deblocked frame, at most 2 rows out). Threads form SW-wide lane groups,
 * one per band of rows. Every LDS read of the window is dword-aligned (the
 * LDS's unaligned path stalls its command queue), and pairs of 16-bit
 * samples go through the packed dot products:
 *   Wiener: a horizontal 7-tap pass writes pairs of columns into an int16
 *           LDS buffer (the clip3 range fits 16 bits at 8/10/12 bits), each
 *           output a sum of packed sample pairs times packed tap pairs;
 *           then a vertical 7-tap pass down each column's band over a
 *           7-row register window.
 *   SGR:    per part of 32 output rows, a walk down each column computes,
 *           two rows at a time, the 5- and 3-sample horizontal sums (sum
 *           and sum of squares) of the new window rows, their 5- and 3-row
 *           sums, and A and B of both passes (pass 0, r = 2, on odd rows
 *           only), packed as (256 - A) << 21 | B so that the sum of two
 *           neighbours is one add; the 24-bit multiplies use per-set
 *           constants from a table. A second walk applies the 3x3 weights
 *           (6/5 on odd rows for r = 2, 4/3 for r = 1) over row pairs and
 *           the projection. The set's passes select one of three code
 *           paths (both, r = 1 alone, r = 2 alone).
 * Samples outside restored units, beyond the frame but inside the
 * mode-info area, are copied from the CDEF output.
 */
#include "av1_tables.h"
#include "rvpf_internal.h"

namespace rvpf {

namespace {

constexpr int NT = 256;        /* threads per workgroup */
constexpr int MAXW = 64;       /* max slice width */
constexpr int WS = MAXW + 16;  /* window row stride (samples) */
/* Window column c (sample x0 - 3 + c) sits at LDS column c + WOFF, so the
 * 8-sample chunk starting at x0 - 8 + 8k lands on an aligned column 8k. */
constexpr int WOFF = 5;
constexpr int WR = 64 + 6;     /* window rows: stripe + 3 above and below */
constexpr int ABW = MAXW + 2;  /* A/B row stride: positions -1..w */
/* The self-guided A/B are computed and used in parts of SGR_PART rows of
 * the stripe (rows [r0, r1)), so their buffers hold the positions of one
 * part: r0 - 1 .. r1, at most SGR_PART + 2 rows. */
constexpr int SGR_PART = 32;
constexpr int ABR = SGR_PART + 2; /* A/B rows of a part */

struct LrLaunch {
  Geom g;
  LrHdr h;
  PlanePtr deb[3], cdef[3], out[3];
  int plane_id[3];
  int blk_start[4]; /* padded to multiples of 8 per plane */
  int nseg;
  int nslices[3];
  int nstripes;
  int wide[3]; /* rows of the plane's deblocked, CDEF and output planes are 8-sample aligned */
};

template <typename T>
__device__ __forceinline__ const T *cpp(const PlanePtr &p) {
  return gptr(reinterpret_cast<const T *>(p.data));
}

/* LDS of one workgroup. */
struct LrShared {
  uint16_t win[WR * WS];
  union {
    int16_t hbuf[WR * MAXW]; /* Wiener: horizontal pass output */
    struct {
      uint32_t ab0[(ABR / 2 + 1) * ABW]; /* pass 0 (r = 2): odd rows i at (i - r0 + 1) >> 1 */
      uint32_t ab1[ABR * ABW];           /* pass 1 (r = 1): row i at i - r0 + 1 */
      uint8_t t[256];                    /* 256 - SGR_A2, for sgr_ab */
    } sgr;
  };
};

/* Rows [lo, hi) of n split evenly over nb bands: band b. */
__device__ __forceinline__ void band_rows(int n, int nb, int b, int &lo, int &hi) {
  lo = (b * n) / nb;
  hi = ((b + 1) * n) / nb;
}

/* out[y][x] = in[y][x] for rows y = y, y + NB, ... <= y_last of one
 * column, COPIES rows per batch with every load issued (rows past y_last
 * repeat it) before the stores. */
template <typename T, int NB>
__device__ __forceinline__ void copy_column(T *out, int64_t os, const T *in, int64_t is, int x, int y, int y_last) {
  constexpr int COPIES = 8;
  for (; y <= y_last; y += COPIES * NB) {
    T v[COPIES];
#pragma unroll
    for (int k = 0; k < COPIES; k++) v[k] = in[(int64_t)min(y + k * NB, y_last) * is + x];
#pragma unroll
    for (int k = 0; k < COPIES; k++)
      if (y + k * NB <= y_last) out[(int64_t)(y + k * NB) * os + x] = v[k];
  }
}

/* 8 consecutive samples as one 16-byte (u16) or 8-byte (u8) access. */
template <typename T>
struct Chunk;
template <>
struct Chunk<uint16_t> {
  typedef uint4 V;
  static __device__ __forceinline__ uint4 to_u16(V v) { return v; }
};
template <>
struct Chunk<uint8_t> {
  typedef uint2 V;
  static __device__ __forceinline__ uint4 to_u16(V v) {
    const uint32_t a = v.x, b = v.y;
    return make_uint4((a & 0xff) | ((a << 8) & 0xff0000), ((a >> 16) & 0xff) | ((a >> 8) & 0xff0000),
                      (b & 0xff) | ((b << 8) & 0xff0000), ((b >> 16) & 0xff) | ((b >> 8) & 0xff0000));
  }
};

/* out = in over rows [ya, yb] x columns [x0, x0 + wc) in 8-sample chunks
 * (rows 8-sample aligned); every load of a thread before its stores. The
 * columns past the last whole chunk are copied sample by sample. */
template <typename T>
__device__ __forceinline__ void copy_rect_wide(T *out, int64_t os, const T *in, int64_t is, int x0, int wc, int ya,
                                               int yb) {
  typedef typename Chunk<T>::V V;
  const int nch = wc >> 3, rows = yb - ya + 1, n = nch * rows;
  /* three chunks per thread and step, loaded before they are stored */
  auto chunk = [&](int it, int64_t st_) { const int r = it / nch; return (int64_t)(ya + r) * st_ + x0 + (it - r * nch) * 8; };
  for (int i0 = threadIdx.x; i0 < n; i0 += 3 * NT) {
    const int i1 = i0 + NT, i2 = i0 + 2 * NT;
    const V v0 = *reinterpret_cast<const V *>(in + chunk(i0, is));
    const V v1 = *reinterpret_cast<const V *>(in + chunk(min(i1, n - 1), is));
    const V v2 = *reinterpret_cast<const V *>(in + chunk(min(i2, n - 1), is));
    *reinterpret_cast<V *>(out + chunk(i0, os)) = v0;
    if (i1 < n) *reinterpret_cast<V *>(out + chunk(i1, os)) = v1;
    if (i2 < n) *reinterpret_cast<V *>(out + chunk(i2, os)) = v2;
  }
  const int rem = wc - nch * 8;
  for (int it = threadIdx.x; it < rem * rows; it += NT) {
    const int r = it / rem, c = x0 + nch * 8 + (it - r * rem);
    out[(int64_t)(ya + r) * os + c] = in[(int64_t)(ya + r) * is + c];
  }
}

/* Self-guided constants of one pass (postfilter.rs:915-935): n = (2r+1)^2,
 * the scale s, the clamp pmax on p, 1/n in 12-bit fixed point and the
 * rounding of the sums to 8 bits. z = min((p s + 2^19) >> 20, 255) equals
 * (min(p, pmax) s + 2^19) >> 20 with pmax the smallest p at which it
 * reaches 255: below pmax it is at most 254, and pmax s + 2^19 stays below
 * 256 * 2^20 because s < 2^20. The table's s are 22..3236 and pmax < 2^24,
 * so z is one 24-bit multiply-add. */
struct SgrK {
  uint32_t n, s, pmax, oon;
};
struct SgrTable {
  SgrK k[16][2];
};
constexpr SgrK make_sgr_k(int r, int eps) {
  SgrK q{};
  if (r == 0) eps = 1; /* pass off: never used */
  const uint32_t n = (uint32_t)((2 * r + 1) * (2 * r + 1));
  const uint32_t n2e = n * n * (uint32_t)eps;
  q.n = n;
  q.s = ((1u << 20) + n2e / 2) / n2e;
  q.pmax = ((255u << 20) - (1u << 19) + q.s - 1) / q.s;
  q.oon = ((1u << 12) + n / 2) / n;
  return q;
}
constexpr SgrTable make_sgr_table() {
  SgrTable t{};
  for (int set = 0; set < 16; set++) {
    t.k[set][0] = make_sgr_k(2 * (AV1_SGR_PARAMS[set][0] != 0), AV1_SGR_PARAMS[set][1]);
    t.k[set][1] = make_sgr_k(AV1_SGR_PARAMS[set][2] != 0, AV1_SGR_PARAMS[set][3]);
  }
  return t;
}
constexpr SgrTable SGR_K = make_sgr_table();

/* 256 - SGR_A2 (postfilter.rs:884-898): 0 at 255, 255 at 0, else
 * 256 - ((z << 8) + z / 2) / (z + 1). A byte per entry. */
struct SgrT {
  uint8_t v[256];
};
constexpr SgrT make_sgr_t() {
  SgrT t{};
  for (int z = 0; z < 256; z++)
    t.v[z] = (uint8_t)(256 - (z >= 255 ? 256 : (z == 0 ? 1 : ((z << 8) + z / 2) / (z + 1))));
  return t;
}
constexpr SgrT SGR_T = make_sgr_t();

/* Rounding of the box sums to 8 bits: sum >> sh and sq >> 2 sh, rounded. */
struct SgrRound {
  uint32_t r1, sh1, r2, sh2;
};

/* A and B at one position (postfilter.rs:940-965) from the box sum and sum
 * of squares, packed as (t << 21) | B with t = 256 - A: t <= 255 and
 * B < 2^20 at 8/10/12 bits, so the sum of two packed values keeps both
 * fields apart. B = (t (sum / n) + 2048) >> 12 in 12-bit fixed point;
 * t * oon < 2^17 and the full product is below 2^32. */
__device__ __forceinline__ uint32_t sgr_ab(uint32_t sum, uint32_t sq, const SgrRound &rd, const SgrK &k,
                                           const uint8_t *tt) {
  const uint32_t av = (sq + rd.r2) >> rd.sh2;
  const uint32_t d = (sum + rd.r1) >> rd.sh1;
  const uint32_t an = __umul24(av, k.n), dd = __umul24(d, d);
  const uint32_t p = an > dd ? an - dd : 0;
  const uint32_t z = (__umul24(min(p, k.pmax), k.s) + (1u << 19)) >> 20;
  const uint32_t t = tt[z];
  /* oon < 2^10: the mask lets the compiler keep a 24-bit multiply */
  const uint32_t b = (__umul24(__umul24(t, k.oon & 0x3FFu), sum) + 2048u) >> 12;
  return (t << 21) | b;
}

typedef unsigned short ushort2_l __attribute__((ext_vector_type(2)));

/* acc + a.lo * b.lo + a.hi * b.hi (unsigned 16-bit halves; v_dot2_u32_u16) */
__device__ __forceinline__ uint32_t udot2(uint32_t a, uint32_t b, uint32_t acc) {
  return __builtin_amdgcn_udot2(__builtin_bit_cast(ushort2_l, a), __builtin_bit_cast(ushort2_l, b), acc, false);
}

/* Window samples c .. c + 4 of one row as packed pairs (c, c+1), (c+2, c+3)
 * and the fifth sample, from three dword-aligned LDS reads: an 8-byte
 * read at a 2-byte boundary takes the LDS's unaligned path, which stalls
 * its command queue. sh = 16 (c & 1) selects the half-dword phase. */
struct Win5 {
  uint32_t p12, p34, w5;
};
__device__ __forceinline__ Win5 win5(const uint16_t *row, int c, uint32_t sh) {
  const uint32_t *d = reinterpret_cast<const uint32_t *>(row) + (c >> 1);
  const uint32_t d0 = d[0], d1 = d[1], d2 = d[2];
  return {__builtin_amdgcn_alignbit(d1, d0, sh), __builtin_amdgcn_alignbit(d2, d1, sh), (d2 >> sh) & 0xFFFFu};
}

/* Horizontal sums of one window row for one A/B column: over the five
 * samples (h5, q5: sum and sum of squares) and the middle three (h3, q3). */
struct HRow {
  uint32_t h3, h5, q3, q5;
};
__device__ __forceinline__ HRow hrow(const Win5 &v) {
  HRow r;
  const uint32_t y34 = udot2(v.p34, 0x00010001u, 0);
  r.h3 = udot2(v.p12, 0x00010000u, y34);
  r.h5 = udot2(v.p12, 0x00010001u, y34 + v.w5);
  const uint32_t s34 = udot2(v.p34, v.p34, 0);
  r.q3 = udot2(v.p12 & 0xFFFF0000u, v.p12, s34);
  r.q5 = __umul24(v.w5, v.w5) + udot2(v.p12, v.p12, s34);
  return r;
}
/* the sums of window row `row` at LDS column c */
__device__ __forceinline__ HRow hrow_at(const uint16_t *row, int c, uint32_t sh) { return hrow(win5(row, c, sh)); }

/* Weighting terms of one row of packed A/B, at columns j - 1, j, j + 1
 * (p[0..2]): E = left + right, C = centre, each split into its t and B
 * fields. */
__device__ __forceinline__ void split_row(const uint32_t *p, int &et, int &eb, int &ct, int &cb) {
  const uint32_t l = p[0], c = p[1], r = p[2];
  const uint32_t e = l + r;
  et = (int)(e >> 21);
  eb = (int)(e & 0x1FFFFFu);
  ct = (int)(c >> 21);
  cb = (int)(c & 0x1FFFFFu);
}

/* 24-bit signed multiply (-add): values carried around a loop lose the
 * range the compiler would need to pick the 24-bit instruction itself. */
__device__ __forceinline__ int mul24i(int a, int b) {
  int r;
  asm("v_mul_i32_i24 %0, %1, %2" : "=v"(r) : "v"(a), "v"(b));
  return r;
}
__device__ __forceinline__ int mad24i(int a, int b, int c) {
  int r;
  asm("v_mad_i32_i24 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "v"(c));
  return r;
}
/* clip3(lo, hi, x) as one v_med3_i32 (lo <= hi) */
__device__ __forceinline__ int med3i(int x, int lo, int hi) {
  int r;
  asm("v_med3_i32 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(lo), "v"(hi));
  return r;
}

/* Pass 0 (r = 2) on one odd row: a = the sum of 6 A(j) + 5 (A(j-1) +
 * A(j+1)), as 16 * 256 less the same weights of t, and b = the same sum of
 * B plus 128, the rounding of an odd output row. */
struct P0Row {
  int a, b;
};
__device__ __forceinline__ P0Row p0_row(const uint32_t *p) {
  int et, eb, ct, cb;
  split_row(p, et, eb, ct, cb);
  return {mad24i(ct, -6, mad24i(et, -5, 4096)), mad24i(cb, 6, mad24i(eb, 5, 128))};
}

/* Pass 1 (r = 1) terms of one row: as an outer row (weights 3 4 3) x =
 * -(3 Et + 4 Ct) of t and z = 3 Eb + 4 Cb + 128 of B, as the centre row
 * (weights 4 4 4) y = -4 (Et + Ct) and wb = 4 (Eb + Cb). */
struct P1Row {
  int x, y, z, wb;
};
__device__ __forceinline__ P1Row p1_row(const uint32_t *p) {
  int et, eb, ct, cb;
  split_row(p, et, eb, ct, cb);
  const int x = mad24i(et, -3, mul24i(ct, -4));
  const int z = mad24i(eb, 3, mad24i(cb, 4, 128));
  return {x, x - et, z, z + eb - 128};
}

/* Pass 1 output from rows m, c, n: A = 8192 + xm + xn + yc (the nine
 * weights of A = 256 - t sum to 32), B = zm + zn + wbc (with the rounding
 * 256); round2(A u + B, 9). */
__device__ __forceinline__ int p1_out(const P1Row &m, const P1Row &c, const P1Row &n, int u) {
  return mad24i(m.x + n.x + c.y, u, (u << 13) + (m.z + n.z + c.wb)) >> 9;
}

/* A/B of both passes (P0: r = 2 on odd rows, P1: r = 1) at positions
 * (i, j), i in [r0 - 1, r0 + npos - 2], j in [-1, w]: pass 0 into ab0 row
 * (i - r0 + 1) >> 1, pass 1 into ab1 row i - r0 + 1, column j + 1. The
 * stripe starts on an even frame row, so odd i are odd frame rows.
 * Positions come in pairs (odd i, i + 1): lane group `band` walks pairs
 * [plo, phi) down column j = col - 1, reading two new window rows per
 * pair (rows i + 4, i + 5; the pair needs rows i + 1 .. i + 5). Columns
 * SW - 1 .. w, past the lanes, are computed position by position. */
template <int SW, bool P0, bool P1>
__device__ __forceinline__ void sgr_ab_part(LrShared &sm, int band, int col, int w, int r0, int npair,
                                            const SgrRound &rd, const SgrK &k0, const SgrK &k1) {
  constexpr int NB = NT / SW;
  const uint8_t *tt = sm.sgr.t;
  int plo, phi;
  band_rows(npair, NB, band, plo, phi);
  const int j = col - 1;
  if (j <= w && plo < phi) {
    const int cw = WOFF + j + 1; /* LDS column of the box's first sample */
    const uint32_t sh = (cw & 1) << 4;
    const uint16_t *wp = &sm.win[(r0 + 2 * plo) * WS];
    const HRow a = hrow_at(wp, cw, sh);
    HRow b = hrow_at(wp + WS, cw, sh), c = hrow_at(wp + 2 * WS, cw, sh);
    wp += 3 * WS;
    uint32_t *p0 = &sm.sgr.ab0[plo * ABW + j + 1];
    uint32_t *p1 = &sm.sgr.ab1[2 * plo * ABW + j + 1];
    uint32_t ps = a.h5 + b.h5, pq = a.q5 + b.q5; /* rows i + 1, i + 2 */
#pragma unroll 2
    for (int m = plo; m < phi; m++) {
      const HRow x = hrow_at(wp, cw, sh), y = hrow_at(wp + WS, cw, sh);
      wp += 2 * WS;
      if (P0) {
        *p0 = sgr_ab(ps + c.h5 + x.h5 + y.h5, pq + c.q5 + x.q5 + y.q5, rd, k0, tt);
        ps = c.h5 + x.h5;
        pq = c.q5 + x.q5;
      }
      if (P1) {
        p1[0] = sgr_ab(b.h3 + c.h3 + x.h3, b.q3 + c.q3 + x.q3, rd, k1, tt);
        p1[ABW] = sgr_ab(c.h3 + x.h3 + y.h3, c.q3 + x.q3 + y.q3, rd, k1, tt);
      }
      b = x;
      c = y;
      p0 += ABW;
      p1 += 2 * ABW;
    }
  }
  const int npos = 2 * npair, nl = w - SW + 2;
  for (int t = threadIdx.x; t < nl * npos; t += NT) {
    const int q = t / npos, i = t - q * npos; /* position r0 - 1 + i */
    const int jj = SW - 1 + q;
    const int cw = WOFF + jj + 1;
    const uint32_t sh = (cw & 1) << 4;
    const uint16_t *p = &sm.win[(r0 + i) * WS];
    const HRow h1 = hrow_at(p + WS, cw, sh), h2 = hrow_at(p + 2 * WS, cw, sh), h3 = hrow_at(p + 3 * WS, cw, sh);
    if (P0 && (i & 1) == 0) { /* odd position: rows 1 .. 5 */
      const HRow h0 = hrow_at(p, cw, sh), h4 = hrow_at(p + 4 * WS, cw, sh);
      sm.sgr.ab0[(i >> 1) * ABW + jj + 1] =
          sgr_ab(h0.h5 + h1.h5 + h2.h5 + h3.h5 + h4.h5, h0.q5 + h1.q5 + h2.q5 + h3.q5 + h4.q5, rd, k0, tt);
    }
    if (P1) sm.sgr.ab1[i * ABW + jj + 1] = sgr_ab(h1.h3 + h2.h3 + h3.h3, h1.q3 + h2.q3 + h3.q3, rd, k1, tt);
  }
}

/* The 3x3 weighting of both passes and the projection (postfilter.rs:823-851)
 * for output rows [r0, r1) at column j = col: lane group `band` takes row
 * pairs (even i, i + 1) [qlo, qhi). acc = cu u + w0 f0 + w2 f1, where cu
 * folds w1 and the weight of a pass that is off (its f is u << 4). */
template <typename T, int SW, bool P0, bool P1>
__device__ __forceinline__ void sgr_weight_part(const LrShared &sm, int band, int col, int w, int r0, int r1,
                                                T *out, int64_t os, int cu, int w0, int w2, int maxv) {
  constexpr int NB = NT / SW;
  const int j = col;
  int qlo, qhi;
  band_rows((r1 - r0 + 1) >> 1, NB, band, qlo, qhi);
  if (j >= w || qlo >= qhi) return;
  int i = r0 + 2 * qlo;
  const uint32_t *a0 = &sm.sgr.ab0[qlo * ABW + j];       /* odd row i - 1 */
  const uint32_t *a1 = &sm.sgr.ab1[2 * qlo * ABW + j];   /* rows i - 1, i */
  P0Row h0 = {0, 0};
  P1Row m1 = {0, 0, 0, 0}, c1 = {0, 0, 0, 0};
  if (P0) h0 = p0_row(a0);
  if (P1) {
    m1 = p1_row(a1);
    c1 = p1_row(a1 + ABW);
  }
  a0 += ABW;
  a1 += 2 * ABW;
  const uint16_t *up = &sm.win[(i + 3) * WS + WOFF + j + 3];
  T *o = out + (int64_t)i * os;
  /* output rows i and i + 1; unrolled by two so that the row terms
   * carried from one pair to the next need no register moves */
  auto step = [&]() {
    const int u0 = up[0], u1 = up[WS];
    up += 2 * WS;
    int acc0 = mad24i(cu, u0, 1024), acc1 = mad24i(cu, u1, 1024);
    if (P0) {
      const P0Row n0 = p0_row(a0);
      a0 += ABW;
      acc0 = mad24i(w0, mad24i(h0.a + n0.a, u0, h0.b + n0.b) >> 9, acc0);
      acc1 = mad24i(w0, mad24i(n0.a, u1, n0.b) >> 8, acc1);
      h0 = n0;
    }
    if (P1) {
      const P1Row n1 = p1_row(a1), n2 = p1_row(a1 + ABW);
      a1 += 2 * ABW;
      acc0 = mad24i(w2, p1_out(m1, c1, n1, u0), acc0);
      acc1 = mad24i(w2, p1_out(c1, n1, n2, u1), acc1);
      m1 = n1;
      c1 = n2;
    }
    o[0] = (T)med3i(acc0 >> 11, 0, maxv);
    if (i + 1 < r1) o[os] = (T)med3i(acc1 >> 11, 0, maxv);
    o += 2 * os;
    i += 2;
  };
  int q = qlo;
  for (; q + 1 < qhi; q += 2) {
    step();
    step();
  }
  if (q < qhi) step();
}

/* Self-guided restoration of one tile (box_filter_rect,
 * postfilter.rs:901-991) in parts of SGR_PART output rows: the A/B of the
 * part's positions, then the weighting. */
template <typename T, int SW, bool P0, bool P1>
__device__ __forceinline__ void sgr_tile(LrShared &sm, const rvpf_lr_unit *u, int set, int band, int col, int w,
                                         int h, int bd, T *out, int64_t os) {
  const SgrK k0 = SGR_K.k[set][0], k1 = SGR_K.k[set][1];
  const int sh = bd - 8;
  const SgrRound rd = {(1u << sh) >> 1, (uint32_t)sh, (1u << (2 * sh)) >> 1, (uint32_t)(2 * sh)};
  const int w0 = u->sgr_xqd[0], w1 = u->sgr_xqd[1];
  const int w2 = 128 - w0 - w1;
  const int cu = (w1 + (P0 ? 0 : w0) + (P1 ? 0 : w2)) << 4;
  const int maxv = (1 << bd) - 1;
  for (int r0 = 0; r0 < h; r0 += SGR_PART) {
    const int r1 = min(h, r0 + SGR_PART);
    if (r0) __syncthreads(); /* the previous part's weighting has read A/B */
    sgr_ab_part<SW, P0, P1>(sm, band, col, w, r0, (r1 - r0 + 3) >> 1, rd, k0, k1);
    __syncthreads();
    sgr_weight_part<T, SW, P0, P1>(sm, band, col, w, r0, r1, out, os, cu, w0, w2, maxv);
  }
}

typedef short short2_l __attribute__((ext_vector_type(2)));

/* acc + a.lo * b.lo + a.hi * b.hi (signed 16-bit halves; v_dot2_i32_i16) */
__device__ __forceinline__ int sdot2(uint32_t a, uint32_t b, int acc) {
  return __builtin_amdgcn_sdot2(__builtin_bit_cast(short2_l, a), __builtin_bit_cast(short2_l, b), acc, false);
}

__device__ __forceinline__ uint32_t pack16(int lo, int hi) { return ((uint32_t)lo & 0xFFFFu) | ((uint32_t)hi << 16); }

/* Wiener restoration of one tile (wiener_coefficient, postfilter.rs:994-1004,
 * with rounding_variables, predict.rs:1153-1166; symmetric 7-tap filters).
 * Horizontal pass: a thread filters an aligned pair of output columns (j,
 * j + 1) of one window row from five aligned dwords of the window: the
 * taps of column j + 1 are whole sample pairs, those of column j the same
 * pairs shifted by one sample (v_alignbit), each pair times a packed pair
 * of taps (v_dot2_i32_i16); the clipped result fits 16 bits and the pair is
 * stored as one dword. Vertical pass: a thread walks its column down a
 * band of rows with a 7-row register window. `o` points at the thread's
 * column in the tile's first row. */
template <typename T, int SW>
__device__ __forceinline__ void wiener_tile(LrShared &sm, const rvpf_lr_unit *u, int band, int col, int w, int h,
                                            int wh, int bd, T *o, int64_t os) {
  constexpr int NB = NT / SW;
  constexpr int NPAIR = SW / 2, LPAIR = SW == 64 ? 5 : 4; /* column pairs per row, log2 */
  const int round0 = bd == 12 ? 5 : 3, round1 = bd == 12 ? 9 : 11;
  const int v0 = u->wiener[0][0], v1 = u->wiener[0][1], v2 = u->wiener[0][2];
  const int h0 = u->wiener[1][0], h1 = u->wiener[1][1], h2 = u->wiener[1][2];
  const int v3 = 128 - 2 * (v0 + v1 + v2), h3 = 128 - 2 * (h0 + h1 + h2);
  const int offset = 1 << (bd + 7 - round0 - 1);
  const int limit = (1 << (bd + 1 + 7 - round0)) - 1;
  const int cmin = -offset, cmax = limit - offset;
  {
    const uint32_t t01 = pack16(h0, h1), t23 = pack16(h2, h3), t21 = pack16(h2, h1);
    const uint32_t t0l = pack16(h0, 0), t0h = pack16(0, h0);
    const int rnd = (1 << round0) >> 1;
    const int npairs = (w + 1) >> 1;
    uint32_t *hb = reinterpret_cast<uint32_t *>(sm.hbuf);
    for (int it = threadIdx.x; it < wh * NPAIR; it += NT) {
      const int r = it >> LPAIR, q = it & (NPAIR - 1);
      if (q >= npairs) continue;
      /* output column j = 2q: window columns j .. j + 7 at LDS columns
       * j + 5 .. j + 12, inside dwords (j + 4) / 2 .. (j + 12) / 2 */
      const uint32_t *d = reinterpret_cast<const uint32_t *>(&sm.win[r * WS]) + q + 2;
      const uint32_t d2 = d[0], d3 = d[1], d4 = d[2], d5 = d[3], d6 = d[4];
      int s0 = sdot2(__builtin_amdgcn_alignbit(d3, d2, 16), t01, rnd);
      s0 = sdot2(__builtin_amdgcn_alignbit(d4, d3, 16), t23, s0);
      s0 = sdot2(__builtin_amdgcn_alignbit(d5, d4, 16), t21, s0);
      s0 = sdot2(d5, t0h, s0);
      int s1 = sdot2(d3, t01, rnd);
      s1 = sdot2(d4, t23, s1);
      s1 = sdot2(d5, t21, s1);
      s1 = sdot2(d6, t0l, s1);
      hb[r * (MAXW / 2) + q] = pack16(min(max(s0 >> round0, cmin), cmax), min(max(s1 >> round0, cmin), cmax));
    }
  }
  __syncthreads();
  if (col >= w) return;
  constexpr int RB = (64 + NB - 1) / NB; /* rows per band at most */
  int lo, hi;
  band_rows(h, NB, band, lo, hi);
  const int16_t *hb = &sm.hbuf[lo * MAXW + col];
  const int rnd1 = (1 << round1) >> 1, maxv = (1 << bd) - 1;
  int r0 = hb[0], r1 = hb[MAXW], r2 = hb[2 * MAXW], r3 = hb[3 * MAXW], r4 = hb[4 * MAXW], r5 = hb[5 * MAXW];
  o += (int64_t)lo * os;
#pragma unroll
  for (int k = 0; k < RB; k++) {
    if (lo + k >= hi) break;
    const int r6 = hb[(k + 6) * MAXW];
    const int sum = __mul24(v0, r0 + r6) + __mul24(v1, r1 + r5) + __mul24(v2, r2 + r4) + __mul24(v3, r3) + rnd1;
    o[0] = (T)min(max(sum >> round1, 0), maxv);
    o += os;
    r0 = r1;
    r1 = r2;
    r2 = r3;
    r3 = r4;
    r4 = r5;
    r5 = r6;
  }
}

/* Tile (plane, stripe st, slice sl) with slice width SW (64, or 32 for
 * 32-sample units). */
template <typename T, int SW>
__device__ __forceinline__ void lr_tile(const LrLaunch &a, int plane, int st, int sl, LrShared &sm) {
  constexpr int LG = SW == 64 ? 6 : 5;
  constexpr int NB = NT / SW; /* lane groups = row bands */
  const int col = threadIdx.x & (SW - 1), band = threadIdx.x >> LG;
  const Geom &g = a.g;
  const int ssx = plane ? g.ssx : 0, ssy = plane ? g.ssy : 0;
  const int pex = ((g.width + ssx) >> ssx) - 1;
  const int pey = ((g.height + ssy) >> ssy) - 1;
  const int area_w = (g.mi_cols * 4) >> ssx, area_h = (g.mi_rows * 4) >> ssy;
  const int unit_size = a.h.unit_size[plane];
  /* lr_stripe, postfilter.rs:640-661 */
  const int stripe_start = (-8 + st * 64) >> ssy;
  const int stripe_end = stripe_start + (64 >> ssy) - 1;
  const int y0 = max(stripe_start, 0);
  const int y1 = min(stripe_end, pey);
  const int y1c = st == a.nstripes - 1 ? area_h - 1 : min(stripe_end, area_h - 1); /* rows written */
  const int x0 = sl * SW;
  const int x1c = min(x0 + SW, area_w) - 1; /* columns written */
  const int xr = min(x1c, pex);             /* columns restored */
  if (y0 > y1c || x0 > x1c) return;
  const T *cdef = cpp<T>(a.cdef[plane]);
  const int64_t cs_ = a.cdef[plane].stride;
  T *out = gptr(reinterpret_cast<T *>(a.out[plane].data));
  const int64_t os = a.out[plane].stride;

  int type = 0;
  const rvpf_lr_unit *u = nullptr;
  if (y0 <= y1 && x0 <= xr) {
    const int luma_row = max(st * 64 - 8, 0);
    const int unit_row = min(a.h.unit_rows[plane] - 1, ((luma_row + 8) >> ssy) / unit_size);
    const int unit_col = min(a.h.unit_cols[plane] - 1, x0 >> (31 - __clz(unit_size)));
    u = &gptr(a.h.units[plane])[unit_row * a.h.unit_cols[plane] + unit_col];
    type = u->type;
    if (type != RVPF_RESTORE_WIENER && type != RVPF_RESTORE_SGRPROJ) type = 0;
  }

  /* copy what is not restored: everything (no filter), else the rows
   * below the frame and the columns right of it inside the mode-info area */
  if (type == 0 && a.wide[plane]) {
    copy_rect_wide<T>(out, os, cdef, cs_, x0, x1c - x0 + 1, y0, y1c);
    return;
  }
  {
    const int x = x0 + col;
    const int ya = type ? y1 + 1 : y0;
    if (x <= x1c) copy_column<T, NB>(out, os, cdef, cs_, x, ya + band, y1c);
    if (type && x > xr && x <= x1c) copy_column<T, NB>(out, os, cdef, cs_, x, y0 + band, y1);
  }
  if (type == 0) return;

  const int w = xr - x0 + 1, h = y1 - y0 + 1;
  const int ww = w + 6, wh = h + 6;
  /* LrCtx::sample (postfilter.rs:584-596) over the window */
  {
    const T *deb = cpp<T>(a.deb[plane]);
    const int64_t ds = a.deb[plane].stride;
    /* source row of window row r (rows outside the stripe: the deblocked
     * frame, at most 2 out) */
    auto src_row = [&](int r) -> const T * {
      const int yc = min(max(y0 + r - 3, 0), pey);
      const bool above = yc < stripe_start, below = yc > stripe_end;
      const int y = above ? max(stripe_start - 2, yc) : (below ? min(stripe_end + 2, yc) : yc);
      return (above || below) ? deb + (int64_t)y * ds : cdef + (int64_t)y * cs_;
    };
    constexpr int NCH = (SW + 16) / 8; /* chunks from x0 - 8 to x0 + SW + 8 */
    if (a.wide[plane] && x0 >= 8 && x0 + SW + 8 <= pex + 1) {
      /* interior tile: aligned 8-sample chunks, every load of a thread
       * before its LDS stores */
      typedef typename Chunk<T>::V V;
      const int n = NCH * wh;
      constexpr int PER = (NCH * WR + NT - 1) / NT;
      V v[PER];
#pragma unroll
      for (int k = 0; k < PER; k++) {
        const int it = min((int)threadIdx.x + k * NT, n - 1);
        const int r = it / NCH, c = it - r * NCH;
        v[k] = *reinterpret_cast<const V *>(src_row(r) + x0 - 8 + 8 * c);
      }
#pragma unroll
      for (int k = 0; k < PER; k++) {
        const int it = threadIdx.x + k * NT;
        if (it < n) {
          const int r = it / NCH, c = it - r * NCH;
          *reinterpret_cast<uint4 *>(&sm.win[r * WS + 8 * c]) = Chunk<T>::to_u16(v[k]);
        }
      }
    } else {
      const int xa = min(max(x0 + col - 3, 0), pex);
      const int xb = min(max(x0 + col + SW - 3, 0), pex);
      const bool second = col + SW < ww;
      /* LOADS rows per batch: every load is issued (rows past the window
       * repeat its last row) before the LDS stores */
      constexpr int LOADS = 6;
      for (int r0 = band; r0 < wh; r0 += LOADS * NB) {
        uint16_t va[LOADS], vb[LOADS];
#pragma unroll
        for (int k = 0; k < LOADS; k++) {
          const T *src = src_row(min(r0 + k * NB, wh - 1));
          va[k] = src[xa];
          vb[k] = src[xb];
        }
#pragma unroll
        for (int k = 0; k < LOADS; k++) {
          const int r = r0 + k * NB;
          if (r < wh) {
            uint16_t *wr = &sm.win[r * WS + WOFF + col];
            wr[0] = va[k];
            if (second) wr[SW] = vb[k];
          }
        }
      }
    }
  }
  const int bd = g.bd;
  if (type == RVPF_RESTORE_WIENER) {
    __syncthreads();
    wiener_tile<T, SW>(sm, u, band, col, w, h, wh, bd, out + (int64_t)y0 * os + x0 + col, os);
    return;
  }

  /* self-guided */
  for (int i = threadIdx.x; i < 256; i += NT) sm.sgr.t[i] = SGR_T.v[i];
  __syncthreads();
  const int set = u->sgr_set & 15;
  T *ob = out + (int64_t)y0 * os + x0 + col;
  if (AV1_SGR_PARAMS[set][0] == 0)
    sgr_tile<T, SW, false, true>(sm, u, set, band, col, w, h, bd, ob, os);
  else if (AV1_SGR_PARAMS[set][2] == 0)
    sgr_tile<T, SW, true, false>(sm, u, set, band, col, w, h, bd, ob, os);
  else
    sgr_tile<T, SW, true, true>(sm, u, set, band, col, w, h, bd, ob, os);
}

/* Slice lb of grid plane gp (stripe-major). */
template <typename T>
__device__ __forceinline__ void lr_slice(const LrLaunch &a, int gp, int lb, LrShared &sm) {
  const int plane = a.plane_id[gp];
  const int st = lb / a.nslices[gp], sl = lb - st * a.nslices[gp];
  if (a.h.unit_size[plane] >= MAXW)
    lr_tile<T, 64>(a, plane, st, sl, sm);
  else
    lr_tile<T, 32>(a, plane, st, sl, sm);
}

template <typename T>
__global__ __launch_bounds__(NT) void lr_kernel(LrLaunch a) {
  __shared__ LrShared sm;
  int lb;
  const int gp = xcd_segment(blockIdx.x, a.blk_start, a.nseg, &lb);
  if (gp >= a.nseg || lb >= a.nslices[gp] * a.nstripes) return; /* padding */
  lr_slice<T>(a, gp, lb, sm);
}

/* Batched: item j of a frame is slice j of its (unpadded) grid. The
 * persistent form is held at 6 waves per SIMD, the occupancy the LDS
 * allows (80 VGPRs and a 48-byte spill; unbounded it takes 110, 4 waves). */
template <typename T, bool DIRECT>
__global__ __launch_bounds__(NT, DIRECT ? 1 : 6) void lr_batch(BatchWork w) {
  __shared__ LrShared sm;
  batch_loop<DIRECT>(w, [&](int f, int j) {
    const LrLaunch &a = batch_table<LrLaunch>(w.rec, f);
    int gp = 0;
    while (gp + 1 < a.nseg && j >= a.blk_start[gp + 1]) gp++;
    lr_slice<T>(a, gp, j - a.blk_start[gp], sm);
  });
}

bool rows_aligned8(const PlanePtr &q, int bps) {
  return q.data && ((uintptr_t)q.data % (8 * (uintptr_t)bps)) == 0 && q.stride % 8 == 0;
}

int plan(const Geom &g, const LrHdr &h, const PlanePtr deblocked[3], const PlanePtr cdef[3],
         const PlanePtr out[3], int bps, bool pad8, LrLaunch &a) {
  a.g = g;
  a.h = h;
  for (int p = 0; p < 3; p++) {
    a.deb[p] = deblocked[p];
    a.cdef[p] = cdef[p];
    a.out[p] = out[p];
    a.wide[p] = p < g.nplanes && rows_aligned8(deblocked[p], bps) && rows_aligned8(cdef[p], bps) &&
                rows_aligned8(out[p], bps);
  }
  a.nstripes = (g.height + 8 + 63) / 64;
  int n = 0, total = 0;
  for (int p = 0; p < g.nplanes; p++) {
    if (h.type[p] == 0) continue;
    const int ssx = p ? g.ssx : 0;
    const int area_w = (g.mi_cols * 4) >> ssx;
    const int sw = h.unit_size[p] < MAXW ? h.unit_size[p] : MAXW;
    a.plane_id[n] = p;
    a.nslices[n] = (area_w + sw - 1) / sw;
    a.blk_start[n] = total;
    total += pad8 ? (a.nslices[n] * a.nstripes + 7) & ~7 : a.nslices[n] * a.nstripes;
    n++;
  }
  a.nseg = n;
  for (int i = n; i < 3; i++) {
    a.plane_id[i] = 0;
    a.nslices[i] = 1;
    a.blk_start[i] = total;
  }
  a.blk_start[3] = total;
  return total;
}

}  // namespace

size_t lr_record_size() { return sizeof(LrLaunch); }

int lr_plan(const Geom &g, const LrHdr &h, const PlanePtr deblocked[3], const PlanePtr cdef[3],
            const PlanePtr out[3], int bps, void *rec) {
  return plan(g, h, deblocked, cdef, out, bps, false, *static_cast<LrLaunch *>(rec));
}

int launch_lr_batch(const BatchWork &w, const StagePlan &sp, int bps, hipStream_t st) {
  if (w.total == 0) return 0;
  if (bps == 1)
    if (batch_direct(w, sp.grid))
      lr_batch<uint8_t, true><<<sp.grid, NT, 0, st>>>(w);
    else
      lr_batch<uint8_t, false><<<sp.grid, NT, 0, st>>>(w);
  else
    if (batch_direct(w, sp.grid))
      lr_batch<uint16_t, true><<<sp.grid, NT, 0, st>>>(w);
    else
      lr_batch<uint16_t, false><<<sp.grid, NT, 0, st>>>(w);
  return (int)hipGetLastError();
}

int launch_lr(const Geom &g, const LrHdr &h, const PlanePtr deblocked[3], const PlanePtr cdef[3],
              PlanePtr out[3], int bps, hipStream_t st) {
  LrLaunch a;
  const int total = plan(g, h, deblocked, cdef, out, bps, true, a);
  if (total == 0) return 0;
  if (bps == 1)
    lr_kernel<uint8_t><<<total, NT, 0, st>>>(a);
  else
    lr_kernel<uint16_t><<<total, NT, 0, st>>>(a);
  return (int)hipGetLastError();
}

}  // namespace rvpf
