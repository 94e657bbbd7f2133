/*
 * cdef.hip — CDEF (AV1 7.15) for gfx950.
 *
 * Follows rivet-av1's cdef / cdef_block / cdef_direction /
 * cdef_filter_block (src/decoder/postfilter.rs:273-513) and
 * dsp::cdef::filter_scalar (src/dsp/cdef.rs:48-90).
 *
 * One single-wave workgroup (64 lanes) per 64x64 filter block (one
 * cdef_idx):
 *   1. the luma tile plus a 2-sample border is staged in LDS; samples at or
 *      beyond the mode-info edge are the sentinel NA_S (unavailable);
 *   2. each lane runs the direction search of one 8x8 block;
 *   3. lanes filter one column each, row by row, from LDS (coalesced row
 *      stores); then each chroma plane is staged and filtered the same way.
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

__device__ __forceinline__ int constrain(int diff, int threshold, int adj) {
  const int a = abs(diff);
  const int mag = min(a, max(threshold - (a >> adj), 0));
  return diff < 0 ? -mag : mag;
}

/* Stages rows [y0-2, y0+h+2) x cols [x0-2, x0+w+2) of a plane into t
 * (stride TS); outside [0, xlim) x [0, ylim) -> NA_S. */
template <typename T>
__device__ void stage_tile(uint16_t *t, const T *src, int64_t stride, int x0, int y0, int w, int h,
                           int xlim, int ylim) {
  const int tw = w + 4, th = h + 4;
  for (int i = threadIdx.x; i < tw * th; i += blockDim.x) {
    const int r = i / tw, c = i - r * tw;
    const int y = y0 + r - 2, x = x0 + c - 2;
    int v = NA_S;
    if (y >= 0 && y < ylim && x >= 0 && x < xlim) v = src[(int64_t)y * stride + x];
    t[r * TS + c] = (uint16_t)v;
  }
}

/* One filtered sample (dsp::cdef::filter_scalar, src/dsp/cdef.rs:61-88). */
struct Taps {
  int pri, sec, pri_adj, sec_adj, pt0, pt1, st0, st1;
  int po0, po1, so00, so01, so10, so11; /* tile offsets */
  bool clamp;
};

__device__ __forceinline__ int filter_px(const uint16_t *t, int ci, const Taps &k) {
  const int x = t[ci];
  int sum = 0, mx = x, mn = x;
  if (k.pri) {
    const int a = t[ci + k.po0], b = t[ci - k.po0], c = t[ci + k.po1], d = t[ci - k.po1];
    sum += k.pt0 * (constrain(a - x, k.pri, k.pri_adj) + constrain(b - x, k.pri, k.pri_adj));
    sum += k.pt1 * (constrain(c - x, k.pri, k.pri_adj) + constrain(d - x, k.pri, k.pri_adj));
    if (k.clamp) {
      mn = min(min(mn, a), min(b, min(c, d)));
      mx = max(max(mx, a & 0xfff), max(b & 0xfff, max(c & 0xfff, d & 0xfff)));
    }
  }
  if (k.sec) {
    const int a = t[ci + k.so00], b = t[ci - k.so00], c = t[ci + k.so01], d = t[ci - k.so01];
    const int e = t[ci + k.so10], f = t[ci - k.so10], gg = t[ci + k.so11], hh = t[ci - k.so11];
    sum += k.st0 * (constrain(a - x, k.sec, k.sec_adj) + constrain(b - x, k.sec, k.sec_adj) +
                    constrain(c - x, k.sec, k.sec_adj) + constrain(d - x, k.sec, k.sec_adj));
    sum += k.st1 * (constrain(e - x, k.sec, k.sec_adj) + constrain(f - x, k.sec, k.sec_adj) +
                    constrain(gg - x, k.sec, k.sec_adj) + constrain(hh - x, k.sec, k.sec_adj));
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
  k.st0 = AV1_CDEF_SEC_TAPS[sel][0];
  k.st1 = AV1_CDEF_SEC_TAPS[sel][1];
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

/* The direction search of 7.15.2 (postfilter.rs:345-401) for the 8x8
 * block at tile offset b0: returns the direction, *var gets var. */
__device__ int cdef_direction(const uint16_t *t, int b0, int sh, int32_t *var) {
  int32_t part[8][15];
#pragma unroll
  for (int d = 0; d < 8; d++)
#pragma unroll
    for (int i = 0; i < 15; i++) part[d][i] = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int x = (t[b0 + i * TS + j] >> sh) - 128;
      part[0][i + j] += x;
      part[1][i + j / 2] += x;
      part[2][i] += x;
      part[3][3 + i - j / 2] += x;
      part[4][7 + i - j] += x;
      part[5][3 - i / 2 + j] += x;
      part[6][j] += x;
      part[7][i / 2 + j] += x;
    }
  }
  int32_t cost[8];
  cost[2] = cost[6] = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    cost[2] += part[2][i] * part[2][i];
    cost[6] += part[6][i] * part[6][i];
  }
  cost[2] *= AV1_DIV_TABLE[8];
  cost[6] *= AV1_DIV_TABLE[8];
  cost[0] = part[0][7] * part[0][7] * AV1_DIV_TABLE[8];
  cost[4] = part[4][7] * part[4][7] * AV1_DIV_TABLE[8];
#pragma unroll
  for (int i = 0; i < 7; i++) {
    cost[0] += (part[0][i] * part[0][i] + part[0][14 - i] * part[0][14 - i]) * AV1_DIV_TABLE[i + 1];
    cost[4] += (part[4][i] * part[4][i] + part[4][14 - i] * part[4][14 - i]) * AV1_DIV_TABLE[i + 1];
  }
#pragma unroll
  for (int d = 1; d < 8; d += 2) {
    int32_t c = 0;
#pragma unroll
    for (int j = 0; j < 5; j++) c += part[d][3 + j] * part[d][3 + j];
    c *= AV1_DIV_TABLE[8];
#pragma unroll
    for (int j = 0; j < 3; j++)
      c += (part[d][j] * part[d][j] + part[d][10 - j] * part[d][10 - j]) * AV1_DIV_TABLE[2 * j + 2];
    cost[d] = c;
  }
  int32_t best = 0;
  int dir = 0;
#pragma unroll
  for (int k = 0; k < 8; k++)
    if (cost[k] > best) {
      best = cost[k];
      dir = k;
    }
  *var = (best - cost[(dir + 4) & 7]) >> 10;
  return dir;
}

template <typename T>
__global__ __launch_bounds__(64) void cdef_kernel(CdefLaunch a) {
  __shared__ uint16_t tile[TS * TS];
  __shared__ int8_t blk_dir[64];  /* luma direction (yDir); -1: block not filtered */
  __shared__ int16_t blk_pri[64]; /* variance-adjusted luma primary strength */
  const Geom &g = a.g;
  const int lane = threadIdx.x;
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
      for (int i = lane; i < w * h; i += 64) {
        const int r = i / w, c = i - r * w;
        dst[(int64_t)(py + r) * a.out[p].stride + px + c] = src[(int64_t)(py + r) * a.in[p].stride + px + c];
      }
    }
    return;
  }

  stage_tile<T>(tile, pp<T>(a.in[0]), a.in[0].stride, x0, y0, 64, 64, xlim, ylim);
  __syncthreads();

  /* cdef_block (postfilter.rs:304-328): one lane per 8x8 block */
  {
    const int by = lane >> 3, bx = lane & 7;
    int dir = -1, pri = 0;
    if (by * 8 < bh && bx * 8 < bw) {
      const int r = (y0 >> 2) + by * 2, c = (x0 >> 2) + bx * 2;
      const int ms = a.mi_stride;
      const bool skip = a.mi[(int64_t)r * ms + c].seg_skip & a.mi[(int64_t)(r + 1) * ms + c].seg_skip &
                        a.mi[(int64_t)r * ms + c + 1].seg_skip & a.mi[(int64_t)(r + 1) * ms + c + 1].seg_skip & 1;
      if (!skip) {
        int32_t var;
        dir = cdef_direction(tile, (by * 8 + 2) * TS + bx * 8 + 2, cs, &var);
        pri = a.h.y_pri[idx] << cs;
        const int var_str = (var >> 6) != 0 ? min(floor_log2(var >> 6), 12) : 0;
        pri = var != 0 ? (pri * (4 + var_str) + 8) >> 4 : 0;
      }
    }
    blk_dir[lane] = (int8_t)dir;
    blk_pri[lane] = (int16_t)pri;
  }
  __syncthreads();

  /* luma: lane = column, loop over rows */
  {
    const int sec = a.h.y_sec[idx] << cs;
    const int damping = a.h.damping + cs;
    const bool use_dir = a.h.y_pri[idx] != 0;
    T *dst = pp<T>(a.out[0]);
    const int c = lane;
    if (c < bw) {
#pragma unroll 1
      for (int r = 0; r < bh; r++) {
        const int b = (r >> 3) * 8 + (c >> 3);
        const int d = blk_dir[b];
        const int ci = (r + 2) * TS + c + 2;
        int v = tile[ci];
        if (d >= 0) {
          const int pri = blk_pri[b];
          if (pri | sec) v = filter_px(tile, ci, make_taps(pri, sec, damping, use_dir ? d : 0, cs));
        }
        dst[(int64_t)(y0 + r) * a.out[0].stride + x0 + c] = (T)v;
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
  const int bws = 8 >> ssx, bhs = 8 >> ssy; /* 8x8-luma block size in this plane */
  for (int p = 1; p < 3; p++) {
    __syncthreads();
    stage_tile<T>(tile, pp<T>(a.in[p]), a.in[p].stride, x0 >> ssx, y0 >> ssy, cw, ch, xlim >> ssx,
                  ylim >> ssy);
    __syncthreads();
    T *dst = pp<T>(a.out[p]);
#pragma unroll 1
    for (int i = lane; i < w * h; i += 64) {
      const int r = i / w, c = i - r * w;
      const int b = (r / bhs) * 8 + c / bws;
      const int d = blk_dir[b];
      const int ci = (r + 2) * TS + c + 2;
      int v = tile[ci];
      if (d >= 0 && (pri | sec)) {
        const int dir = pri == 0 ? 0 : AV1_CDEF_UV_DIR[ssx][ssy][d];
        v = filter_px(tile, ci, make_taps(pri, sec, damping, dir, cs));
      }
      dst[(int64_t)((y0 >> ssy) + r) * a.out[p].stride + (x0 >> ssx) + c] = (T)v;
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
    cdef_kernel<uint8_t><<<n, 64, 0, st>>>(a);
  else
    cdef_kernel<uint16_t><<<n, 64, 0, st>>>(a);
  return (int)hipGetLastError();
}

}  // namespace rvpf
