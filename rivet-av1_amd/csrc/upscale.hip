/*
 * upscale.hip — super-resolution upscaling (AV1 7.16) for gfx950.
 *
 * Follows rivet-av1's upscale (src/decoder/postfilter.rs:521-566). Each
 * output sample x of a row reads 8 source samples around
 * src_x = -2^14 + initial_subpel_x + x * step_x (14-bit fixed point),
 * clamped to the plane's mode-info width, and applies the 8-tap filter of
 * phase (src_x & (2^14 - 1)) >> 8; the result is rounded by 7 bits and
 * clipped to the sample range.
 *
 * Columns past the upscaled width (the caller may ask for them, to fill a
 * padded plane) continue the same formula, reading the clamped last column.
 *
 * The phase and source position of a column do not depend on the row, so a
 * workgroup takes a tile of UR rows x 1024 output columns: each lane owns 4
 * consecutive columns, works out their source offsets and loads their 32
 * taps once (from a copy of the filter table in LDS), then filters them on
 * every row of the tile. The source span
 * of the tile's columns (at most 918 samples, since the ratio is at most
 * 8/9) is staged in LDS for all UR rows first, all loads in flight at once,
 * twice: as loaded and one sample on, so that every column's 8 samples are
 * four aligned 16-bit pairs in one of the copies, multiplied with the
 * packed taps by 16-bit dot products. Four outputs go
 * out as one 8-byte (4-byte for 8-bit planes) store when the output rows
 * are aligned for it.
 *
 * One launch serves up to six plane jobs (three planes of the CDEF output
 * and three of the deblocked frame in the chain).
 */
#include "av1_tables.h"
#include "rvpf_internal.h"

namespace rvpf {

namespace {

constexpr int SCALE_BITS = 14, EXTRA_BITS = 8;
constexpr int NT = 256;       /* threads per workgroup */
constexpr int CW = NT * 4;    /* output columns per tile */
constexpr int UR = 4;         /* rows per tile */
constexpr int SPS = 4;        /* staged source samples per lane and row */
constexpr int SPAN = SPS * NT; /* staged source samples per row */
constexpr int MAXJ = 6;

typedef short short2_t __attribute__((ext_vector_type(2)));

struct UpJob {
  PlanePtr in, out;
  int cols, rows; /* output columns and rows written */
  int max_x;      /* last readable source column */
  int step, init; /* step_x and initial_subpel_x */
  int blk_x;      /* tiles across */
};

struct UpLaunch {
  UpJob j[MAXJ];
  int blk_start[MAXJ + 1];
  int njobs;
  int bd;
};

/* Four consecutive outputs as one store (T = u16: 8 bytes, u8: 4 bytes). */
template <typename T>
__device__ __forceinline__ void store4(T *p, const int v[4]) {
  if constexpr (sizeof(T) == 2) {
    uint2 w;
    w.x = (uint32_t)v[0] | ((uint32_t)v[1] << 16);
    w.y = (uint32_t)v[2] | ((uint32_t)v[3] << 16);
    *reinterpret_cast<uint2 *>(p) = w;
  } else {
    *reinterpret_cast<uint32_t *>(p) =
        (uint32_t)v[0] | ((uint32_t)v[1] << 8) | ((uint32_t)v[2] << 16) | ((uint32_t)v[3] << 24);
  }
}

template <typename T, bool WIDE>
__global__ __launch_bounds__(NT) void upscale_kernel(UpLaunch a) {
  /* the staged rows twice: as loaded (ev) and one sample on (od), so the 8
   * taps of any column start on a 4-byte boundary in one of them */
  __shared__ __align__(16) uint16_t stg[2 * UR * SPAN]; /* ev rows, then od rows */
  __shared__ __align__(16) int16_t filt[64 * 8];
  const int b = blockIdx.x;
  int ji = 0;
  while (ji + 1 < a.njobs && b >= a.blk_start[ji + 1]) ji++;
  const UpJob &q = a.j[ji];
  const int lb = b - a.blk_start[ji];
  const int ty = lb / q.blk_x;
  const int y0 = ty * UR, x0 = (lb - ty * q.blk_x) * CW;
  const int nr = min(UR, q.rows - y0);
  /* source columns of the tile: from px(x0) - 3 on */
  const int sx0 = (-(1 << SCALE_BITS) + q.init + x0 * q.step) >> SCALE_BITS;
  const T *in = gptr(reinterpret_cast<const T *>(q.in.data));
  static_assert(64 * 8 == 2 * NT, "two filter table entries per lane");
  const int16_t f0 = AV1_UPSCALE_FILTER[threadIdx.x >> 3][threadIdx.x & 7];
  const int16_t f1 = AV1_UPSCALE_FILTER[(threadIdx.x + NT) >> 3][threadIdx.x & 7];
  {
    /* every load of the tile issued before the first LDS write (rows past
     * the plane's last re-read it; columns are clamped to the readable ones) */
    T v[UR][SPS];
#pragma unroll
    for (int r = 0; r < UR; r++) {
      const T *row = in + (int64_t)min(y0 + r, q.rows - 1) * q.in.stride;
#pragma unroll
      for (int i = 0; i < SPS; i++) v[r][i] = row[min(max(sx0 - 3 + i * NT + (int)threadIdx.x, 0), q.max_x)];
    }
    filt[threadIdx.x] = f0;
    filt[threadIdx.x + NT] = f1;
#pragma unroll
    for (int r = 0; r < UR; r++)
#pragma unroll
      for (int i = 0; i < SPS; i++) {
        const int c = i * NT + threadIdx.x;
        stg[r * SPAN + c] = v[r][i];                                 /* ev: sample c at c */
        stg[(UR + r) * SPAN + (c == 0 ? SPAN - 1 : c - 1)] = v[r][i]; /* od: at c - 1; od[SPAN - 1] is never read */
      }
  }
  __syncthreads();
  /* this lane's columns: where their taps start (a 4-byte aligned row
   * offset into ev or od) and their taps as packed 16-bit pairs */
  const int xl = x0 + (int)threadIdx.x * 4;
  int base[4];
  short2_t tap[4][4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int sxq = -(1 << SCALE_BITS) + q.init + (xl + j) * q.step;
    const int off = (sxq >> SCALE_BITS) - sx0;
    base[j] = (off & 1) * UR * SPAN + (off & ~1);
    const int ph = (sxq & ((1 << SCALE_BITS) - 1)) >> EXTRA_BITS;
    const uint4 w = *reinterpret_cast<const uint4 *>(&filt[ph * 8]);
    tap[j][0] = __builtin_bit_cast(short2_t, w.x);
    tap[j][1] = __builtin_bit_cast(short2_t, w.y);
    tap[j][2] = __builtin_bit_cast(short2_t, w.z);
    tap[j][3] = __builtin_bit_cast(short2_t, w.w);
  }
  if (xl >= q.cols) return;
  const int maxv = (1 << a.bd) - 1;
  T *out = gptr(reinterpret_cast<T *>(q.out.data)) + (int64_t)y0 * q.out.stride + xl;
  const bool full = xl + 4 <= q.cols;
  for (int r = 0; r < nr; r++) {
    int v[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      /* samples k, k+1 of the window are one aligned word; samples are
       * below 2^12, so a signed 16-bit dot product holds them */
      const uint32_t *s = reinterpret_cast<const uint32_t *>(&stg[base[j] + r * SPAN]);
      int sum = 0;
#pragma unroll
      for (int k = 0; k < 4; k++) sum = __builtin_amdgcn_sdot2(__builtin_bit_cast(short2_t, s[k]), tap[j][k], sum, false);
      v[j] = min(max((sum + 64) >> 7, 0), maxv);
    }
    T *o = out + (int64_t)r * q.out.stride;
    if (WIDE && full) {
      store4<T>(o, v);
    } else {
#pragma unroll
      for (int j = 0; j < 4; j++)
        if (xl + j < q.cols) o[j] = (T)v[j];
    }
  }
}

int round2i(int x, int n) { return (x + ((1 << n) >> 1)) >> n; }

/* upscale, postfilter.rs:527-545: step_x and initial_subpel_x of a plane */
void plane_params(int fw, int upw, int sx, int64_t *step, int64_t *init) {
  const int64_t dw = round2i(fw, sx), uw = round2i(upw, sx);
  *step = ((dw << SCALE_BITS) + uw / 2) / uw;
  const int64_t err = uw * *step - (dw << SCALE_BITS);
  int64_t i = (-((uw - dw) << (SCALE_BITS - 1)) + uw / 2) / uw + (1 << (EXTRA_BITS - 1)) - err / 2;
  *init = i & ((1 << SCALE_BITS) - 1);
}

/* rows of `q` start on 4-sample boundaries */
bool rows_aligned4(const PlanePtr &q, int bps) {
  return q.stride % 4 == 0 && ((uintptr_t)q.data % (4 * bps)) == 0;
}

}  // namespace

int launch_upscale(const Geom &g, int frame_width, int upscaled_width, const UpscaleJob *jobs, int njobs, int bps,
                   hipStream_t st) {
  if (njobs > MAXJ) return (int)hipErrorInvalidValue;
  UpLaunch a;
  a.bd = g.bd;
  a.njobs = 0;
  int total = 0;
  bool wide = true;
  for (int i = 0; i < njobs; i++) {
    const UpscaleJob &u = jobs[i];
    if (u.cols <= 0 || u.rows <= 0) continue;
    const int sx = u.plane ? g.ssx : 0;
    int64_t step, init;
    plane_params(frame_width, upscaled_width, sx, &step, &init);
    UpJob &q = a.j[a.njobs];
    q.in = u.in;
    q.out = u.out;
    q.cols = u.cols;
    q.rows = u.rows;
    q.max_x = (g.mi_cols >> sx) * 4 - 1;
    q.step = (int)step;
    q.init = (int)init;
    q.blk_x = (u.cols + CW - 1) / CW;
    a.blk_start[a.njobs] = total;
    total += q.blk_x * ((u.rows + UR - 1) / UR);
    wide &= rows_aligned4(u.out, bps);
    a.njobs++;
  }
  for (int i = a.njobs; i <= MAXJ; i++) a.blk_start[i] = total;
  if (total == 0) return 0;
  if (bps == 1) {
    if (wide)
      upscale_kernel<uint8_t, true><<<total, NT, 0, st>>>(a);
    else
      upscale_kernel<uint8_t, false><<<total, NT, 0, st>>>(a);
  } else {
    if (wide)
      upscale_kernel<uint16_t, true><<<total, NT, 0, st>>>(a);
    else
      upscale_kernel<uint16_t, false><<<total, NT, 0, st>>>(a);
  }
  return (int)hipGetLastError();
}

}  // namespace rvpf
