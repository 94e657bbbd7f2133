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
 * one per band of rows; a thread owns one column and walks down its band,
 * so every vertical stencil slides through registers and reads one new row
 * per output row:
 *   Wiener: horizontal 7-tap pass into an int16 LDS buffer (the clip3
 *           range fits 16 bits at 8/10/12 bits), then the vertical 7-tap
 *           pass over a 7-row register window.
 *   SGR:    one walk per column computes the 5- and 3-row sums of the
 *           5- and 3-sample horizontal sums (sum and sum of squares) and
 *           from them A and B of both passes (pass 0, r = 2, on odd rows
 *           only), each packed as (A << 20) | B; a second walk applies
 *           the 3x3 weights (6/5 on odd rows for r = 2, 4/3 for r = 1)
 *           from a 3-row register window and the projection.
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

__device__ __forceinline__ int round2(int x, int n) { return (x + ((1 << n) >> 1)) >> n; }

/* SGR_A2 (postfilter.rs:884-898): 256 at 255 and above, 1 at 0, else
 * ((z << 8) + z / 2) / (z + 1). */
struct A2Table {
  uint16_t v[256];
};
constexpr A2Table make_a2() {
  A2Table t{};
  for (int z = 0; z < 256; z++) t.v[z] = (uint16_t)(z >= 255 ? 256 : (z == 0 ? 1 : ((z << 8) + z / 2) / (z + 1)));
  return t;
}
constexpr A2Table SGR_A2 = make_a2();

/* LDS of one workgroup. */
struct LrShared {
  uint16_t win[WR * WS];
  union {
    int16_t hbuf[WR * MAXW]; /* Wiener: horizontal pass output */
    struct {
      uint32_t ab0[(ABR / 2 + 1) * ABW]; /* pass 0 (r = 2): odd rows i at (i - r0 + 1) >> 1 */
      uint32_t ab1[ABR * ABW];           /* pass 1 (r = 1): row i at i - r0 + 1 */
      uint16_t a2[256];                  /* the a2 table of sgr_ab */
    } sgr;
  };
};

/* Per-pass constants of the self-guided filter (postfilter.rs:915-935). */
struct SgrPass {
  uint32_t n, s, one_over_n, p_lim;
};

__device__ __forceinline__ SgrPass sgr_pass(int r, int eps) {
  SgrPass q;
  const int n = (2 * r + 1) * (2 * r + 1);
  const int n2e = n * n * eps;
  q.n = (uint32_t)n;
  q.s = ((1u << 20) + (uint32_t)(n2e / 2)) / (uint32_t)n2e;
  q.one_over_n = ((1u << 12) + (uint32_t)(n / 2)) / (uint32_t)n;
  /* p * s fits in 32 bits up to this bound; beyond it z saturates */
  q.p_lim = (0xFFFFFFFFu - (1u << 19)) / q.s;
  return q;
}

/* A and B at one position from the box sum and sum of squares
 * (postfilter.rs:940-965), packed as (A << 20) | B: A <= 256 and
 * B <= (255 * 25 * 4095 * 164 + 2048) >> 12 < 2^20. Every product fits
 * 32 bits unsigned at 8/10/12 bits. */
__device__ __forceinline__ uint32_t sgr_ab(uint32_t sum, uint32_t sq, int bd, const SgrPass &q, const uint16_t *a2t) {
  const int sh = bd - 8;
  const uint32_t av = (sq + ((1u << (2 * sh)) >> 1)) >> (2 * sh);
  const uint32_t d = (sum + ((1u << sh) >> 1)) >> sh;
  const uint32_t an = __umul24(av, q.n), dd = __umul24(d, d);
  const uint32_t p = an > dd ? an - dd : 0;
  const uint32_t z = p > q.p_lim ? 255u : min((p * q.s + (1u << 19)) >> 20, 255u);
  const uint32_t a2 = a2t[z];
  const uint32_t b2 = ((256u - a2) * __umul24(sum, q.one_over_n) + 2048u) >> 12;
  return (a2 << 20) | b2;
}

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
  const int maxv = (1 << bd) - 1;

  if (type == RVPF_RESTORE_WIENER) {
    __syncthreads();
    /* wiener_coefficient (postfilter.rs:994-1004) + rounding_variables
     * (predict.rs:1153-1166); symmetric taps */
    const int round0 = bd == 12 ? 5 : 3, round1 = bd == 12 ? 9 : 11;
    const int v0 = u->wiener[0][0], v1 = u->wiener[0][1], v2 = u->wiener[0][2];
    const int h0 = u->wiener[1][0], h1 = u->wiener[1][1], h2 = u->wiener[1][2];
    const int v3 = 128 - 2 * (v0 + v1 + v2), h3 = 128 - 2 * (h0 + h1 + h2);
    const int offset = 1 << (bd + 7 - round0 - 1);
    const int limit = (1 << (bd + 1 + 7 - round0)) - 1;
    if (col < w)
      for (int r = band; r < wh; r += NB) {
        const uint16_t *s = &sm.win[r * WS + WOFF + col];
        const int sum = __mul24(h0, s[0] + s[6]) + __mul24(h1, s[1] + s[5]) + __mul24(h2, s[2] + s[4]) +
                        __mul24(h3, (int)s[3]);
        sm.hbuf[r * MAXW + col] = (int16_t)min(max(round2(sum, round0), -offset), limit - offset);
      }
    __syncthreads();
    if (col < w) {
      int lo, hi;
      band_rows(h, NB, band, lo, hi);
      const int16_t *hb = &sm.hbuf[lo * MAXW + col];
      int r0 = hb[0], r1 = hb[MAXW], r2 = hb[2 * MAXW], r3 = hb[3 * MAXW], r4 = hb[4 * MAXW], r5 = hb[5 * MAXW];
      T *o = out + (int64_t)(y0 + lo) * os + x0 + col;
      for (int i = lo; i < hi; i++, o += os) {
        const int r6 = sm.hbuf[(i + 6) * MAXW + col];
        const int sum = __mul24(v0, r0 + r6) + __mul24(v1, r1 + r5) + __mul24(v2, r2 + r4) + __mul24(v3, r3);
        *o = (T)min(max(round2(sum, round1), 0), maxv);
        r0 = r1;
        r1 = r2;
        r2 = r3;
        r3 = r4;
        r4 = r5;
        r5 = r6;
      }
    }
    return;
  }

  /* self-guided: box_filter_rect (postfilter.rs:901-991) for both passes */
  for (int i = threadIdx.x; i < 256; i += NT) sm.sgr.a2[i] = SGR_A2.v[i];
  __syncthreads();
  const int set = u->sgr_set & 15;
  const int rad0 = AV1_SGR_PARAMS[set][0], rad1 = AV1_SGR_PARAMS[set][2];
  const SgrPass q0 = sgr_pass(2, rad0 ? AV1_SGR_PARAMS[set][1] : 1);
  const SgrPass q1 = sgr_pass(1, rad1 ? AV1_SGR_PARAMS[set][3] : 1);
  const uint16_t *a2t = sm.sgr.a2;
  for (int r0 = 0; r0 < h; r0 += SGR_PART) {
    /* output rows [r0, r1) of this part */
    const int r1 = min(h, r0 + SGR_PART);
    if (r0) __syncthreads(); /* the previous part's weighting has read A/B */
    /* A/B at positions (i, j), i in [r0 - 1, r1], j in [-1, w]; the box of
     * radius 2 covers window rows i+1..i+5 and columns j+1..j+5. The walk
     * of thread (band, col) covers column j = col - 1 (the columns past
     * SW - 2 are done directly below). */
    {
      const int j = col - 1;
      const int np = r1 - r0 + 2;
      int lo, hi;
      band_rows(np, NB, band, lo, hi);
      lo += r0 - 1;
      hi += r0 - 1;
      if (j <= w) {
        uint32_t h5a = 0, h5b = 0, h5c = 0, h5d = 0, q5a = 0, q5b = 0, q5c = 0, q5d = 0;
        uint32_t h3a = 0, h3b = 0, q3a = 0, q3b = 0;
        const uint16_t *wp = &sm.win[(lo + 1) * WS + WOFF + j + 1];
        for (int R = lo + 1; R <= hi + 4; R++, wp += WS) {
          const uint32_t w1 = wp[0], w2 = wp[1], w3 = wp[2], w4 = wp[3], w5 = wp[4];
          const uint32_t h3n = w2 + w3 + w4, h5n = h3n + w1 + w5;
          const uint32_t q3n = __umul24(w2, w2) + __umul24(w3, w3) + __umul24(w4, w4);
          const uint32_t q5n = q3n + __umul24(w1, w1) + __umul24(w5, w5);
          const int i5 = R - 5, i3 = R - 4;
          if (rad0 && i5 >= lo && ((y0 + i5) & 1))
            sm.sgr.ab0[((i5 - r0 + 1) >> 1) * ABW + j + 1] =
                sgr_ab(h5a + h5b + h5c + h5d + h5n, q5a + q5b + q5c + q5d + q5n, bd, q0, a2t);
          if (rad1 && i3 >= lo && i3 < hi) sm.sgr.ab1[(i3 - r0 + 1) * ABW + j + 1] = sgr_ab(h3a + h3b + h3n, q3a + q3b + q3n, bd, q1, a2t);
          h5a = h5b;
          h5b = h5c;
          h5c = h5d;
          h5d = h5n;
          q5a = q5b;
          q5b = q5c;
          q5c = q5d;
          q5d = q5n;
          h3a = h3b;
          h3b = h3n;
          q3a = q3b;
          q3b = q3n;
        }
      }
      /* columns SW - 1 .. w, every row, directly */
      const int nl = w - SW + 2;
      for (int t = threadIdx.x; t < nl * np; t += NT) {
        const int jj = SW - 1 + t / np, i = r0 - 1 + t % np;
        uint32_t s5 = 0, sq5 = 0, s3 = 0, sq3 = 0;
        for (int dy = 0; dy < 5; dy++)
          for (int dx = 0; dx < 5; dx++) {
            const uint32_t v = sm.win[(i + 1 + dy) * WS + WOFF + jj + 1 + dx];
            s5 += v;
            sq5 += __umul24(v, v);
            if (dy >= 1 && dy <= 3 && dx >= 1 && dx <= 3) {
              s3 += v;
              sq3 += __umul24(v, v);
            }
          }
        if (rad0 && ((y0 + i) & 1)) sm.sgr.ab0[((i - r0 + 1) >> 1) * ABW + jj + 1] = sgr_ab(s5, sq5, bd, q0, a2t);
        if (rad1) sm.sgr.ab1[(i - r0 + 1) * ABW + jj + 1] = sgr_ab(s3, sq3, bd, q1, a2t);
      }
    }
    __syncthreads();
    /* 3x3 weighting of both passes and the projection (postfilter.rs:823-851) */
    if (col < w) {
      const int j = col;
      const int w0 = u->sgr_xqd[0], w1 = u->sgr_xqd[1];
      const int w2 = 128 - w0 - w1;
      int lo, hi;
      band_rows(r1 - r0, NB, band, lo, hi);
      lo += r0;
      hi += r0;
      /* per A/B row: E = A(j-1) + A(j+1), C = A(j) (and the same for B) */
      const uint32_t *p1 = &sm.sgr.ab1[(lo - r0) * ABW + j];
      int ea_m = 0, ca_m = 0, eb_m = 0, cb_m = 0, ea_c = 0, ca_c = 0, eb_c = 0, cb_c = 0;
      if (rad1) {
        uint32_t l = p1[0], c = p1[1], r = p1[2];
        ea_m = (int)((l >> 20) + (r >> 20));
        ca_m = (int)(c >> 20);
        eb_m = (int)((l & 0xFFFFF) + (r & 0xFFFFF));
        cb_m = (int)(c & 0xFFFFF);
        l = p1[ABW];
        c = p1[ABW + 1];
        r = p1[ABW + 2];
        ea_c = (int)((l >> 20) + (r >> 20));
        ca_c = (int)(c >> 20);
        eb_c = (int)((l & 0xFFFFF) + (r & 0xFFFFF));
        cb_c = (int)(c & 0xFFFFF);
      }
      /* pass 0: 6*C + 5*E of the last odd row at or above i */
      int ra_p = 0, rb_p = 0;
      if (rad0) {
        const int io = ((y0 + lo) & 1) ? lo : lo - 1;
        const uint32_t *p0 = &sm.sgr.ab0[((io - r0 + 1) >> 1) * ABW + j];
        const uint32_t l = p0[0], c = p0[1], r = p0[2];
        ra_p = 6 * (int)(c >> 20) + 5 * (int)((l >> 20) + (r >> 20));
        rb_p = 6 * (int)(c & 0xFFFFF) + 5 * (int)((l & 0xFFFFF) + (r & 0xFFFFF));
      }
      T *o = out + (int64_t)(y0 + lo) * os + x0 + j;
      for (int i = lo; i < hi; i++, o += os) {
        const int uu = sm.win[(i + 3) * WS + WOFF + j + 3];
        const int ue = uu << 4;
        int f0 = ue, f1 = ue;
        if (rad0) {
          if ((y0 + i) & 1) {
            f0 = round2(__mul24(ra_p, uu) + rb_p, 8);
          } else {
            const uint32_t *p0 = &sm.sgr.ab0[((i + 2 - r0) >> 1) * ABW + j];
            const uint32_t l = p0[0], c = p0[1], r = p0[2];
            const int ra_n = 6 * (int)(c >> 20) + 5 * (int)((l >> 20) + (r >> 20));
            const int rb_n = 6 * (int)(c & 0xFFFFF) + 5 * (int)((l & 0xFFFFF) + (r & 0xFFFFF));
            f0 = round2(__mul24(ra_p + ra_n, uu) + rb_p + rb_n, 9);
            ra_p = ra_n;
            rb_p = rb_n;
          }
        }
        if (rad1) {
          const uint32_t *pn = &sm.sgr.ab1[(i + 2 - r0) * ABW + j];
          const uint32_t l = pn[0], c = pn[1], r = pn[2];
          const int ea_n = (int)((l >> 20) + (r >> 20)), ca_n = (int)(c >> 20);
          const int eb_n = (int)((l & 0xFFFFF) + (r & 0xFFFFF)), cb_n = (int)(c & 0xFFFFF);
          /* weights 3 4 3 / 4 4 4 / 3 4 3 */
          const int av = 3 * (ea_m + ea_n) + 4 * (ca_m + ca_n + ea_c + ca_c);
          const int bv = 3 * (eb_m + eb_n) + 4 * (cb_m + cb_n + eb_c + cb_c);
          f1 = round2(__mul24(av, uu) + bv, 9);
          ea_m = ea_c;
          ca_m = ca_c;
          eb_m = eb_c;
          cb_m = cb_c;
          ea_c = ea_n;
          ca_c = ca_n;
          eb_c = eb_n;
          cb_c = cb_n;
        }
        const int acc = __mul24(w1, ue) + __mul24(w0, f0) + __mul24(w2, f1);
        *o = (T)min(max(round2(acc, 11), 0), maxv);
      }
    }
  }
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
