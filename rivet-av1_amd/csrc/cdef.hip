/*
 * cdef.hip — CDEF (AV1 7.15) for gfx950.
 *
 * Follows rivet-av1's cdef / cdef_block / cdef_direction /
 * cdef_filter_block (src/decoder/postfilter.rs:273-513) and
 * dsp::cdef::filter_scalar (src/dsp/cdef.rs:48-90).
 *
 * One workgroup (256 lanes, 4 waves) per 64x64 filter block (one
 * cdef_idx):
 *   1. the luma tile plus a 2-sample border is staged in LDS; samples at or
 *      beyond the mode-info edge are the sentinel NA_S (unavailable);
 *   2. four lanes per 8x8 block run the direction search, each through one
 *      symmetry of the block (no divergence), and one shuffle gathers the
 *      eight costs;
 *   3. lanes filter one column each, every fourth row, from LDS (coalesced
 *      row stores); then each chroma plane is staged and filtered the same
 *      way.
 * Unfiltered blocks (cdef_idx -1, all four 4x4 skipped, CDEF off) are
 * copied, so `out` receives the whole mode-info area.
 *
 * Unavailable taps: NA_S = 0x7000 exceeds every sample by more than 24576,
 * so constrain() of it is 0 for every legal strength/damping (the
 * magnitude test |d| >> adj > strength holds for adj <= 10), it never wins
 * min(), and (v & 0xfff) maps it to 0 for max(); the reference skips it.
 * The clamp to [min, max] only matters when both strengths are non-zero:
 * with one kind of tap the total weight is 12/16, so the rounded update
 * never leaves [min, max]; the kernel clamps only then.
 */
#include "av1_tables.h"
#include "rvpf_internal.h"

namespace rvpf {

namespace {

constexpr int NA_S = 0x7000;
constexpr int TS = 68; /* luma tile stride: 64 + 2 * 2 */

struct CdefLaunch {
  Geom g;
  CdefHdr h;
  const rvpf_mi *mi;
  int mi_stride;
  const int8_t *cdef_idx;
  int cdef_stride;
  PlanePtr in[3], out[3];
  int enabled;
  int fb_cols;
};

template <typename T>
__device__ __forceinline__ T *pp(const PlanePtr &p) {
  return reinterpret_cast<T *>(p.data);
}

/* constrain() of 7.15.3: sign(d) * min(|d|, max(0, t - (|d| >> adj))),
 * i.e. d clamped to [-m, m] with m = max(0, t - (|d| >> adj)). */
__device__ __forceinline__ int constrain(int diff, int threshold, int adj) {
  const int m = max(threshold - (abs(diff) >> adj), 0);
  return min(max(diff, -m), m);
}

/* Stages rows [y0-2, y0+h+2) x cols [x0-2, x0+w+2) of a plane into t
 * (stride TS); outside [0, xlim) x [0, ylim) -> NA_S. w <= 64. */
template <typename T>
__device__ void stage_tile(uint16_t *t, const T *src, int64_t stride, int x0, int y0, int w, int h,
                           int xlim, int ylim) {
  const int tw = w + 4, th = h + 4;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  for (int r = ty; r < th; r += 4) {
    const int y = y0 + r - 2;
    const bool yok = y >= 0 && y < ylim;
    const T *row = src + (int64_t)y * stride;
    for (int c = tx; c < tw; c += 64) {
      const int x = x0 + c - 2;
      t[r * TS + c] = (uint16_t)((yok && x >= 0 && x < xlim) ? (int)row[x] : NA_S);
    }
  }
}

/* Product of two values known to fit in 16 bits (one 24-bit multiply). */
__device__ __forceinline__ int mul16(int a, int b) { return (int)(int16_t)a * (int)(int16_t)b; }

/* One filtered sample (dsp::cdef::filter_scalar, src/dsp/cdef.rs:61-88). */
struct Taps {
  int pri, sec, pri_adj, sec_adj, pt0, pt1;
  int po0, po1, so00, so01, so10, so11; /* tile offsets */
  bool clamp;
};

__device__ __forceinline__ int filter_px(const uint16_t *t, int ci, const Taps &k) {
  const int x = t[ci];
  int sum = 0, mx = x, mn = x;
  if (k.pri) {
    const int a = t[ci + k.po0], b = t[ci - k.po0], c = t[ci + k.po1], d = t[ci - k.po1];
    sum += mul16(k.pt0, constrain(a - x, k.pri, k.pri_adj) + constrain(b - x, k.pri, k.pri_adj));
    sum += mul16(k.pt1, constrain(c - x, k.pri, k.pri_adj) + constrain(d - x, k.pri, k.pri_adj));
    if (k.clamp) {
      mn = min(min(mn, a), min(b, min(c, d)));
      mx = max(max(mx, a & 0xfff), max(b & 0xfff, max(c & 0xfff, d & 0xfff)));
    }
  }
  if (k.sec) {
    const int a = t[ci + k.so00], b = t[ci - k.so00], c = t[ci + k.so01], d = t[ci - k.so01];
    const int e = t[ci + k.so10], f = t[ci - k.so10], gg = t[ci + k.so11], hh = t[ci - k.so11];
    /* Cdef_Sec_Taps is {2, 1} in both rows */
    sum += 2 * (constrain(a - x, k.sec, k.sec_adj) + constrain(b - x, k.sec, k.sec_adj) +
                constrain(c - x, k.sec, k.sec_adj) + constrain(d - x, k.sec, k.sec_adj));
    sum += constrain(e - x, k.sec, k.sec_adj) + constrain(f - x, k.sec, k.sec_adj) +
           constrain(gg - x, k.sec, k.sec_adj) + constrain(hh - x, k.sec, k.sec_adj);
    if (k.clamp) {
      mn = min(min(min(mn, a), min(b, c)), min(min(d, e), min(f, min(gg, hh))));
      mx = max(max(max(mx, a & 0xfff), max(b & 0xfff, c & 0xfff)),
               max(max(d & 0xfff, e & 0xfff), max(f & 0xfff, max(gg & 0xfff, hh & 0xfff))));
    }
  }
  int v = x + ((8 + sum - (sum < 0)) >> 4);
  if (k.clamp) v = min(max(v, mn), mx);
  return v;
}

__device__ __forceinline__ int floor_log2(int x) { return 31 - __clz(x); }

__device__ __forceinline__ int tap_off(int d, int k) {
  return AV1_CDEF_DIRECTIONS[d][k][0] * TS + AV1_CDEF_DIRECTIONS[d][k][1];
}

__device__ __forceinline__ Taps make_taps(int pri, int sec, int damping, int dir, int coeff_shift) {
  Taps k;
  k.pri = pri;
  k.sec = sec;
  k.pri_adj = pri ? max(damping - floor_log2(pri), 0) : 0;
  k.sec_adj = sec ? max(damping - floor_log2(sec), 0) : 0;
  const int sel = (pri >> coeff_shift) & 1;
  k.pt0 = AV1_CDEF_PRI_TAPS[sel][0];
  k.pt1 = AV1_CDEF_PRI_TAPS[sel][1];
  static_assert(AV1_CDEF_SEC_TAPS[0][0] == 2 && AV1_CDEF_SEC_TAPS[0][1] == 1 && AV1_CDEF_SEC_TAPS[1][0] == 2 &&
                    AV1_CDEF_SEC_TAPS[1][1] == 1,
                "secondary taps are 2, 1");
  k.po0 = tap_off(dir, 0);
  k.po1 = tap_off(dir, 1);
  const int dlo = (dir + 6) & 7, dhi = (dir + 2) & 7;
  k.so00 = tap_off(dlo, 0);
  k.so01 = tap_off(dhi, 0);
  k.so10 = tap_off(dlo, 1);
  k.so11 = tap_off(dhi, 1);
  k.clamp = pri != 0 && sec != 0;
  return k;
}

/* Three of the direction costs of 7.15.2 (postfilter.rs:345-390) for an
 * 8x8 block read through a symmetry of the square: pixel (i, j) of the
 * view is at tile offset b0 + i * rs + j * cs. On the view the lane sums
 * lines i + j (direction 0's lines), i + j / 2 (direction 1's) and i
 * (direction 2's). Under a horizontal flip these are directions 4, 3, 2;
 * under a transpose 0, 7, 6; under transpose + flip 4, 5, 6 with the line
 * order reversed, which leaves the costs unchanged (both cost formulas are
 * symmetric in the line index). */
__device__ void dir_costs3(const uint16_t *t, int b0, int rs, int cs, int sh, int32_t *c0, int32_t *c1,
                           int32_t *c2) {
  int32_t p0[15], p1[11], p2[8];
#pragma unroll
  for (int i = 0; i < 15; i++) p0[i] = 0;
#pragma unroll
  for (int i = 0; i < 11; i++) p1[i] = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    int row = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int x = (t[b0 + i * rs + j * cs] >> sh) - 128;
      p0[i + j] += x;
      p1[i + j / 2] += x;
      row += x;
    }
    p2[i] = row;
  }
  /* |partial| <= 1024: squares and the 840/k weights stay within 24-bit
   * multiplies, their products within 32 bits */
  int32_t a = __mul24(mul16(p0[7], p0[7]), AV1_DIV_TABLE[8]);
#pragma unroll
  for (int i = 0; i < 7; i++)
    a += __mul24(mul16(p0[i], p0[i]) + mul16(p0[14 - i], p0[14 - i]), AV1_DIV_TABLE[i + 1]);
  int32_t b = 0;
#pragma unroll
  for (int j = 0; j < 5; j++) b += mul16(p1[3 + j], p1[3 + j]);
  b = __mul24(b, AV1_DIV_TABLE[8]);
#pragma unroll
  for (int j = 0; j < 3; j++)
    b += __mul24(mul16(p1[j], p1[j]) + mul16(p1[10 - j], p1[10 - j]), AV1_DIV_TABLE[2 * j + 2]);
  int32_t c = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) c += mul16(p2[i], p2[i]);
  *c0 = a;
  *c1 = b;
  *c2 = __mul24(c, AV1_DIV_TABLE[8]);
}

template <typename T>
__global__ __launch_bounds__(256) void cdef_kernel(CdefLaunch a) {
  __shared__ uint16_t tile[TS * TS];
  __shared__ int8_t blk_dir[64];  /* luma direction (yDir); -1: block not filtered */
  __shared__ int16_t blk_pri[64]; /* variance-adjusted luma primary strength */
  const Geom &g = a.g;
  const int tid = threadIdx.x;
  const int tx = tid & 63, ty = tid >> 6;
  const int fbx = blockIdx.x % a.fb_cols, fby = blockIdx.x / a.fb_cols;
  const int cs = g.bd - 8;
  int idx = -1;
  if (a.enabled) idx = a.cdef_idx[(int64_t)fby * a.cdef_stride + fbx];
  if (idx >= 0) idx &= 7; /* cdef_bits <= 3 */
  const int xlim = g.mi_cols * 4, ylim = g.mi_rows * 4;
  const int x0 = fbx * 64, y0 = fby * 64;
  const int bw = min(64, xlim - x0), bh = min(64, ylim - y0); /* multiples of 8 */

  if (idx < 0) {
    /* not filtered: copy the block area of every plane */
    for (int p = 0; p < g.nplanes; p++) {
      const int ssx = p ? g.ssx : 0, ssy = p ? g.ssy : 0;
      const int w = bw >> ssx, h = bh >> ssy, px = x0 >> ssx, py = y0 >> ssy;
      const T *src = pp<T>(a.in[p]);
      T *dst = pp<T>(a.out[p]);
      if (tx < w)
        for (int r = ty; r < h; r += 4)
          dst[(int64_t)(py + r) * a.out[p].stride + px + tx] = src[(int64_t)(py + r) * a.in[p].stride + px + tx];
    }
    return;
  }

  stage_tile<T>(tile, pp<T>(a.in[0]), a.in[0].stride, x0, y0, 64, 64, xlim, ylim);
  __syncthreads();

  /* cdef_block (postfilter.rs:304-328): 4 lanes per 8x8 block, each reads
   * the block through one symmetry and yields three direction costs */
  {
    const int b = tid >> 2, q = tid & 3;
    const int by = b >> 3, bx = b & 7;
    const bool inside = by * 8 < bh && bx * 8 < bw;
    bool skip = true;
    int32_t c0 = 0, c1 = 0, c2 = 0;
    if (inside) {
      const int r = (y0 >> 2) + by * 2, c = (x0 >> 2) + bx * 2;
      const int ms = a.mi_stride;
      skip = a.mi[(int64_t)r * ms + c].seg_skip & a.mi[(int64_t)(r + 1) * ms + c].seg_skip &
             a.mi[(int64_t)r * ms + c + 1].seg_skip & a.mi[(int64_t)(r + 1) * ms + c + 1].seg_skip & 1;
      if (!skip) {
        const int base = (by * 8 + 2) * TS + bx * 8 + 2;
        /* q: 0 identity, 1 horizontal flip, 2 transpose, 3 transpose + flip */
        const int rs = q == 0 || q == 1 ? TS : (q == 2 ? 1 : -1);
        const int csx = q == 0 ? 1 : (q == 1 ? -1 : TS);
        const int b0 = base + (q == 1 ? 7 : 0) + (q == 3 ? 7 : 0);
        dir_costs3(tile, b0, rs, csx, cs, &c0, &c1, &c2);
      }
    }
    /* lane q of the block: q0 -> costs 0,1,2; q1 -> 4,3,2; q2 -> 0,7,6; q3 -> 4,5,6 */
    const int lb = (tid & 63) & ~3;
    const int32_t k0 = __shfl(c0, lb + 0, 64), k1 = __shfl(c1, lb + 0, 64), k2 = __shfl(c2, lb + 0, 64);
    const int32_t k4 = __shfl(c0, lb + 1, 64), k3 = __shfl(c1, lb + 1, 64);
    const int32_t k7 = __shfl(c1, lb + 2, 64), k6 = __shfl(c2, lb + 2, 64);
    const int32_t k5 = __shfl(c1, lb + 3, 64);
    if (q == 0) {
      int dir = -1, pri = 0;
      if (inside && !skip) {
        const int32_t cost[8] = {k0, k1, k2, k3, k4, k5, k6, k7};
        int32_t best = 0;
        dir = 0;
#pragma unroll
        for (int k = 0; k < 8; k++)
          if (cost[k] > best) {
            best = cost[k];
            dir = k;
          }
        const int32_t var = (best - cost[(dir + 4) & 7]) >> 10;
        pri = a.h.y_pri[idx] << cs;
        const int var_str = (var >> 6) != 0 ? min(floor_log2(var >> 6), 12) : 0;
        pri = var != 0 ? (pri * (4 + var_str) + 8) >> 4 : 0;
      }
      blk_dir[b] = (int8_t)dir;
      blk_pri[b] = (int16_t)pri;
    }
  }
  __syncthreads();

  /* luma: thread = column tx, rows ty, ty + 4, ... */
  {
    const int sec = a.h.y_sec[idx] << cs;
    const int damping = a.h.damping + cs;
    const bool use_dir = a.h.y_pri[idx] != 0;
    T *dst = pp<T>(a.out[0]) + (int64_t)(y0 + ty) * a.out[0].stride + x0 + tx;
    const int64_t dstep = 4 * a.out[0].stride;
    if (tx < bw) {
#pragma unroll 2
      for (int r = ty; r < bh; r += 4, dst += dstep) {
        const int b = (r >> 3) * 8 + (tx >> 3);
        const int d = blk_dir[b];
        const int ci = (r + 2) * TS + tx + 2;
        int v = tile[ci];
        if (d >= 0) {
          const int pri = blk_pri[b];
          if (pri | sec) v = filter_px(tile, ci, make_taps(pri, sec, damping, use_dir ? d : 0, cs));
        }
        *dst = (T)v;
      }
    }
  }
  if (g.nplanes == 1) return;

  /* chroma planes (postfilter.rs:332-341) */
  const int ssx = g.ssx, ssy = g.ssy;
  const int cw = 64 >> ssx, ch = 64 >> ssy;
  const int pri = a.h.uv_pri[idx] << cs, sec = a.h.uv_sec[idx] << cs;
  const int damping = a.h.damping + cs - 1;
  const int w = bw >> ssx, h = bh >> ssy;
  /* thread -> (row, column) of the chroma block: cw columns per row */
  const int lgw = 6 - ssx;
  const int cx = tid & (cw - 1), cy0 = tid >> lgw, cny = 256 >> lgw;
  for (int p = 1; p < 3; p++) {
    __syncthreads();
    stage_tile<T>(tile, pp<T>(a.in[p]), a.in[p].stride, x0 >> ssx, y0 >> ssy, cw, ch, xlim >> ssx,
                  ylim >> ssy);
    __syncthreads();
    if (!(pri | sec) || cx >= w) {
      if (cx < w) {
        T *dst = pp<T>(a.out[p]);
        for (int r = cy0; r < h; r += cny)
          dst[(int64_t)((y0 >> ssy) + r) * a.out[p].stride + (x0 >> ssx) + cx] = (T)tile[(r + 2) * TS + cx + 2];
      }
      continue;
    }
    T *dst = pp<T>(a.out[p]) + (int64_t)((y0 >> ssy) + cy0) * a.out[p].stride + (x0 >> ssx) + cx;
    const int64_t dstep = (int64_t)cny * a.out[p].stride;
    for (int r = cy0; r < h; r += cny, dst += dstep) {
      const int b = ((r << ssy) >> 3) * 8 + ((cx << ssx) >> 3);
      const int d = blk_dir[b];
      const int ci = (r + 2) * TS + cx + 2;
      int v = tile[ci];
      if (d >= 0) {
        const int dir = pri == 0 ? 0 : AV1_CDEF_UV_DIR[ssx][ssy][d];
        v = filter_px(tile, ci, make_taps(pri, sec, damping, dir, cs));
      }
      *dst = (T)v;
    }
  }
}

}  // namespace

int launch_cdef(const Geom &g, const CdefHdr &h, const rvpf_mi *mi, int mi_stride,
                const int8_t *cdef_idx, int cdef_stride, int enabled, const PlanePtr in[3],
                PlanePtr out[3], int bps, hipStream_t st) {
  CdefLaunch a;
  a.g = g;
  a.h = h;
  a.mi = mi;
  a.mi_stride = mi_stride;
  a.cdef_idx = cdef_idx;
  a.cdef_stride = cdef_stride;
  for (int p = 0; p < 3; p++) {
    a.in[p] = in[p];
    a.out[p] = out[p];
  }
  a.enabled = enabled;
  a.fb_cols = (g.mi_cols + 15) / 16;
  const int fb_rows = (g.mi_rows + 15) / 16;
  const int n = a.fb_cols * fb_rows;
  if (bps == 1)
    cdef_kernel<uint8_t><<<n, 256, 0, st>>>(a);
  else
    cdef_kernel<uint16_t><<<n, 256, 0, st>>>(a);
  return (int)hipGetLastError();
}

}  // namespace rvpf
