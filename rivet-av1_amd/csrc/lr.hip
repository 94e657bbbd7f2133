/*
 * lr.hip — loop restoration (AV1 7.17) for gfx950.
 *
 * Follows rivet-av1's loop_restoration / lr_stripe / lr_unit_cols /
 * restore_stripe / LrCtx::sample / restore_rect_from_window /
 * box_filter_rect / wiener_coefficient (src/decoder/postfilter.rs:572-1004).
 *
 * One workgroup (512 threads) per (plane, stripe, slice): a stripe is 64
 * luma rows offset by -8, a slice is min(64, unit size) columns, so a slice
 * lies in one restoration unit (the last unit of a row absorbs the
 * remainder: its slices map to it by clamping the unit column). The
 * workgroup stages its window (slice + 3-sample border) in LDS with the
 * stripe rule of 7.17.6 (rows above / below the stripe come from the
 * deblocked frame, at most 2 rows out), then runs the unit's filter:
 *   Wiener: horizontal 7-tap pass into an int16 LDS buffer (the clip3
 *           range fits 16 bits at 8/10/12 bits), then the vertical pass;
 *   SGR:    for each pass with r > 0, A and B per position from the
 *           (2r+1)^2 box sums (B's product in 64 bits, z saturated to 255
 *           through a 32-bit bound), the 3x3 weighting (6/5 on odd rows
 *           for r = 2, 4/3 for r = 1), then the projection.
 * Samples outside restored units, beyond the frame but inside the
 * mode-info area, are copied from the CDEF output.
 */
#include "av1_tables.h"
#include "rvpf_internal.h"

namespace rvpf {

namespace {

constexpr int SW = 64;            /* max slice width */
constexpr int WW = SW + 6;        /* window stride */
constexpr int WH = 64 + 6;        /* window rows */
constexpr int AW = SW + 2;        /* A/B stride */
constexpr int NT = 512;           /* threads per workgroup */
constexpr int PX_PER_THREAD = 64 * 64 / NT;

struct LrLaunch {
  Geom g;
  LrHdr h;
  PlanePtr deb[3], cdef[3], out[3];
  int plane_id[3];
  int blk_start[4]; /* padded to multiples of 8 per plane */
  int nseg;
  int nslices[3];
  int nstripes;
};

template <typename T>
__device__ __forceinline__ const T *cpp(const PlanePtr &p) {
  return gptr(reinterpret_cast<const T *>(p.data));
}

__device__ __forceinline__ int round2(int x, int n) { return (x + ((1 << n) >> 1)) >> n; }

/* SGR_A2 (postfilter.rs:884-898): 256 at 255 and above, 1 at 0, else
 * ((z << 8) + z / 2) / (z + 1). */
struct A2Table {
  int16_t v[256];
};
constexpr A2Table make_a2() {
  A2Table t{};
  for (int z = 0; z < 256; z++) t.v[z] = (int16_t)(z >= 255 ? 256 : (z == 0 ? 1 : ((z << 8) + z / 2) / (z + 1)));
  return t;
}
constexpr A2Table SGR_A2 = make_a2();

/* LDS of one workgroup. */
struct LrShared {
  uint16_t win[WH * WW];
  int16_t inter[WH * SW];      /* Wiener horizontal output; SGR: A */
  int32_t bbuf[(64 + 2) * AW]; /* SGR: B */
};

/* Slice lb of grid plane gp (stripe-major). */
template <typename T>
__device__ __forceinline__ void lr_slice(const LrLaunch &a, int gp, int lb, LrShared &sm) {
  uint16_t *win = sm.win;
  int16_t *inter = sm.inter;
  int32_t *bbuf = sm.bbuf;
  const Geom &g = a.g;
  const int plane = a.plane_id[gp];
  const int st = lb / a.nslices[gp], sl = lb - st * a.nslices[gp];
  const int ssx = plane ? g.ssx : 0, ssy = plane ? g.ssy : 0;
  const int plane_end_x = ((g.width + ssx) >> ssx) - 1;
  const int plane_end_y = ((g.height + ssy) >> ssy) - 1;
  const int area_w = (g.mi_cols * 4) >> ssx, area_h = (g.mi_rows * 4) >> ssy;
  const int unit_size = a.h.unit_size[plane];
  const int lg = unit_size >= SW ? 6 : 5; /* slice width 64 or 32 */
  const int sw = 1 << lg;
  const int tx = threadIdx.x & (sw - 1), ty = threadIdx.x >> lg, ny = NT >> lg;
  /* lr_stripe, postfilter.rs:640-661 */
  const int stripe_start = (-8 + st * 64) >> ssy;
  const int stripe_end = stripe_start + (64 >> ssy) - 1;
  const int y0 = max(stripe_start, 0);
  const int y1 = min(stripe_end, plane_end_y);
  const int y1c = st == a.nstripes - 1 ? area_h - 1 : min(stripe_end, area_h - 1); /* rows written */
  const int x0 = sl * sw;
  const int x1c = min(x0 + sw, area_w) - 1; /* columns written */
  const int xr = min(x1c, plane_end_x);     /* columns restored */
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

  /* copy what is not restored */
  {
    const int wc = x1c - x0 + 1, hc = y1c - y0 + 1;
    const int x = x0 + tx;
    if (tx < wc) {
      T *o = out + (int64_t)(y0 + ty) * os + x;
      const T *c = cdef + (int64_t)(y0 + ty) * cs_ + x;
      const int64_t ostep = (int64_t)ny * os, cstep = (int64_t)ny * cs_;
      for (int r = ty; r < hc; r += ny, o += ostep, c += cstep) {
        const int y = y0 + r;
        if (type != 0 && x <= xr && y <= y1) continue;
        *o = *c;
      }
    }
  }
  if (type == 0) return;

  const int w = xr - x0 + 1, h = y1 - y0 + 1;
  const int ww = w + 6, wh = h + 6;
  /* LrCtx::sample (postfilter.rs:584-596) over the window */
  {
    const T *deb = cpp<T>(a.deb[plane]);
    const int64_t ds = a.deb[plane].stride;
    for (int r = ty; r < wh; r += ny) {
      int y = min(max(y0 + r - 3, 0), plane_end_y);
      const T *src;
      if (y < stripe_start) {
        y = max(stripe_start - 2, y);
        src = deb + (int64_t)y * ds;
      } else if (y > stripe_end) {
        y = min(stripe_end + 2, y);
        src = deb + (int64_t)y * ds;
      } else {
        src = cdef + (int64_t)y * cs_;
      }
      for (int c = tx; c < ww; c += sw) win[r * WW + c] = src[min(max(x0 + c - 3, 0), plane_end_x)];
    }
  }
  __syncthreads();
  const int bd = g.bd;
  const int maxv = (1 << bd) - 1;

  if (type == RVPF_RESTORE_WIENER) {
    /* wiener_coefficient (postfilter.rs:994-1004) + rounding_variables
     * (predict.rs:1153-1166); symmetric taps */
    const int round0 = bd == 12 ? 5 : 3, round1 = bd == 12 ? 9 : 11;
    const int v0 = u->wiener[0][0], v1 = u->wiener[0][1], v2 = u->wiener[0][2];
    const int h0 = u->wiener[1][0], h1 = u->wiener[1][1], h2 = u->wiener[1][2];
    const int v3 = 128 - 2 * (v0 + v1 + v2), h3 = 128 - 2 * (h0 + h1 + h2);
    const int offset = 1 << (bd + 7 - round0 - 1);
    const int limit = (1 << (bd + 1 + 7 - round0)) - 1;
    if (tx < w)
      for (int r = ty; r < wh; r += ny) {
        const uint16_t *s = &win[r * WW + tx];
        const int sum = __mul24(h0, s[0] + s[6]) + __mul24(h1, s[1] + s[5]) + __mul24(h2, s[2] + s[4]) +
                        __mul24(h3, (int)s[3]);
        inter[r * SW + tx] = (int16_t)min(max(round2(sum, round0), -offset), limit - offset);
      }
    __syncthreads();
    if (tx < w) {
      T *o = out + (int64_t)(y0 + ty) * os + x0 + tx;
      const int64_t ostep = (int64_t)ny * os;
      for (int r = ty; r < h; r += ny, o += ostep) {
        const int16_t *s = &inter[r * SW + tx];
        const int sum = __mul24(v0, s[0] + s[6 * SW]) + __mul24(v1, s[SW] + s[5 * SW]) +
                        __mul24(v2, s[2 * SW] + s[4 * SW]) + __mul24(v3, (int)s[3 * SW]);
        *o = (T)min(max(round2(sum, round1), 0), maxv);
      }
    }
    return;
  }

  /* self-guided: box_filter_rect (postfilter.rs:901-991) for both passes;
   * pixel (r, tx) of row iteration k keeps pass 0's value in flt0[k] */
  const int set = u->sgr_set & 15;
  const int r0 = AV1_SGR_PARAMS[set][0], r1 = AV1_SGR_PARAMS[set][2];
  const int w0 = u->sgr_xqd[0], w1 = u->sgr_xqd[1];
  const int w2 = 128 - w0 - w1;
  int flt0[PX_PER_THREAD];
  int16_t *A = inter;
#pragma unroll
  for (int pass = 0; pass < 2; pass++) {
    const int r = pass == 0 ? r0 : r1;
    if (r == 0) continue;
    const int eps = AV1_SGR_PARAMS[set][pass * 2 + 1];
    const int n = (2 * r + 1) * (2 * r + 1);
    const int n2e = n * n * eps;
    const uint32_t s = ((1u << 20) + (uint32_t)(n2e / 2)) / (uint32_t)n2e;
    const uint32_t one_over_n = ((1u << 12) + (uint32_t)(n / 2)) / (uint32_t)n;
    /* p * s fits in 32 bits up to this bound; beyond it z saturates */
    const uint32_t p_lim = (0xFFFFFFFFu - (1u << 19)) / s;
    __syncthreads();
    /* A and B at rectangle positions (ii - 1, jj - 1) = window (ii + 2, jj + 2) */
    for (int ii = ty; ii < h + 2; ii += ny) {
      if (pass == 0 && ((y0 + ii - 1) & 1) == 0) continue;
      for (int jj = tx; jj < w + 2; jj += sw) {
        uint32_t sum = 0, sq = 0;
        if (r == 2) {
#pragma unroll
          for (int dy = -2; dy <= 2; dy++) {
            const uint16_t *row = &win[(ii + 2 + dy) * WW + jj + 2];
#pragma unroll
            for (int dx = -2; dx <= 2; dx++) {
              const uint32_t v = row[dx];
              sum += v;
              sq += __umul24(v, v);
            }
          }
        } else {
#pragma unroll
          for (int dy = -1; dy <= 1; dy++) {
            const uint16_t *row = &win[(ii + 2 + dy) * WW + jj + 2];
#pragma unroll
            for (int dx = -1; dx <= 1; dx++) {
              const uint32_t v = row[dx];
              sum += v;
              sq += __umul24(v, v);
            }
          }
        }
        const uint32_t av = (uint32_t)round2((int)sq, 2 * (bd - 8));
        const uint32_t d = (uint32_t)round2((int)sum, bd - 8);
        const uint32_t an = __umul24(av, (uint32_t)n), dd = __umul24(d, d);
        const uint32_t p = an > dd ? an - dd : 0;
        const uint32_t z = p > p_lim ? 255u : min((p * s + (1u << 19)) >> 20, 255u);
        const int a2 = SGR_A2.v[z];
        const uint64_t b2 = (uint64_t)(uint32_t)(256 - a2) * (uint64_t)__umul24(sum, one_over_n);
        A[ii * AW + jj] = (int16_t)a2;
        bbuf[ii * AW + jj] = (int32_t)((b2 + 2048) >> 12);
      }
    }
    __syncthreads();
    if (tx < w) {
#pragma unroll
      for (int k = 0; k < PX_PER_THREAD; k++) {
        const int ii = ty + k * ny;
        if (ii >= h) break;
        const int yabs = y0 + ii;
        const int c = (ii + 1) * AW + tx + 1;
        int av, bv, shift;
        if (pass == 0) {
          if (yabs & 1) {
            /* odd row: this row's A/B only, weights 6 (centre) / 5 */
            av = 6 * A[c] + 5 * (A[c - 1] + A[c + 1]);
            bv = 6 * bbuf[c] + 5 * (bbuf[c - 1] + bbuf[c + 1]);
            shift = 4;
          } else {
            /* even row: the rows above and below */
            av = 6 * (A[c - AW] + A[c + AW]) + 5 * (A[c - AW - 1] + A[c - AW + 1] + A[c + AW - 1] + A[c + AW + 1]);
            bv = 6 * (bbuf[c - AW] + bbuf[c + AW]) +
                 5 * (bbuf[c - AW - 1] + bbuf[c - AW + 1] + bbuf[c + AW - 1] + bbuf[c + AW + 1]);
            shift = 5;
          }
        } else {
          av = 4 * (A[c] + A[c - 1] + A[c + 1] + A[c - AW] + A[c + AW]) +
               3 * (A[c - AW - 1] + A[c - AW + 1] + A[c + AW - 1] + A[c + AW + 1]);
          bv = 4 * (bbuf[c] + bbuf[c - 1] + bbuf[c + 1] + bbuf[c - AW] + bbuf[c + AW]) +
               3 * (bbuf[c - AW - 1] + bbuf[c - AW + 1] + bbuf[c + AW - 1] + bbuf[c + AW + 1]);
          shift = 5;
        }
        const int uu = win[(ii + 3) * WW + tx + 3];
        const int v = round2(__mul24(av, uu) + bv, 4 + shift);
        if (pass == 0 && r1 != 0) {
          flt0[k] = v;
        } else {
          /* projection (postfilter.rs:823-851); here the last pass ran */
          const int ue = uu << 4;
          const int f0 = pass == 0 ? v : (r0 ? flt0[k] : ue);
          const int f1 = pass == 1 ? v : ue;
          const int acc = __mul24(w1, ue) + __mul24(w0, f0) + __mul24(w2, f1);
          out[(int64_t)(y0 + ii) * os + x0 + tx] = (T)min(max(round2(acc, 11), 0), maxv);
        }
      }
    }
  }
}

template <typename T>
__global__ __launch_bounds__(NT) void lr_kernel(LrLaunch a) {
  __shared__ LrShared sm;
  int lb;
  const int gp = xcd_segment(blockIdx.x, a.blk_start, a.nseg, &lb);
  if (gp >= a.nseg || lb >= a.nslices[gp] * a.nstripes) return; /* padding */
  lr_slice<T>(a, gp, lb, sm);
}

/* Batched: item j of a frame is slice j of its (unpadded) grid. */
template <typename T>
__global__ __launch_bounds__(NT) void lr_batch(BatchWork w) {
  __shared__ LrShared sm;
  batch_loop(w, [&](int f, int j) {
    const LrLaunch &a = batch_table<LrLaunch>(w.rec, f);
    int gp = 0;
    while (gp + 1 < a.nseg && j >= a.blk_start[gp + 1]) gp++;
    lr_slice<T>(a, gp, j - a.blk_start[gp], sm);
  });
}

int plan(const Geom &g, const LrHdr &h, const PlanePtr deblocked[3], const PlanePtr cdef[3],
         const PlanePtr out[3], bool pad8, LrLaunch &a) {
  a.g = g;
  a.h = h;
  for (int p = 0; p < 3; p++) {
    a.deb[p] = deblocked[p];
    a.cdef[p] = cdef[p];
    a.out[p] = out[p];
  }
  a.nstripes = (g.height + 8 + 63) / 64;
  int n = 0, total = 0;
  for (int p = 0; p < g.nplanes; p++) {
    if (h.type[p] == 0) continue;
    const int ssx = p ? g.ssx : 0;
    const int area_w = (g.mi_cols * 4) >> ssx;
    const int sw = h.unit_size[p] < SW ? h.unit_size[p] : SW;
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
            const PlanePtr out[3], void *rec) {
  return plan(g, h, deblocked, cdef, out, false, *static_cast<LrLaunch *>(rec));
}

int launch_lr_batch(const BatchWork &w, const StagePlan &sp, int bps, hipStream_t st) {
  if (w.total == 0) return 0;
  if (bps == 1)
    lr_batch<uint8_t><<<sp.grid, NT, 0, st>>>(w);
  else
    lr_batch<uint16_t><<<sp.grid, NT, 0, st>>>(w);
  return (int)hipGetLastError();
}

int launch_lr(const Geom &g, const LrHdr &h, const PlanePtr deblocked[3], const PlanePtr cdef[3],
              PlanePtr out[3], int bps, hipStream_t st) {
  LrLaunch a;
  const int total = plan(g, h, deblocked, cdef, out, true, a);
  if (total == 0) return 0;
  if (bps == 1)
    lr_kernel<uint8_t><<<total, NT, 0, st>>>(a);
  else
    lr_kernel<uint16_t><<<total, NT, 0, st>>>(a);
  return (int)hipGetLastError();
}

}  // namespace rvpf
