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
 * One workgroup per (plane, row, 256 output columns). The source span of
 * the workgroup's columns (at most 256 + 8 samples, since the upscaling
 * never shrinks) is staged in LDS with one load per lane, then every lane
 * filters its sample from LDS.
 */
#include "av1_tables.h"
#include "rvpf_internal.h"

namespace rvpf {

namespace {

constexpr int SCALE_BITS = 14, EXTRA_BITS = 8;
constexpr int UW = 256; /* output columns per workgroup */

struct UpPlane {
  PlanePtr in, out;
  int cols, rows;   /* output columns and rows written */
  int max_x;        /* last readable source column */
  int step, init;   /* step_x and initial_subpel_x */
  int blk_x;        /* workgroups across */
};

struct UpLaunch {
  UpPlane p[3];
  int blk_start[4];
  int bd;
};

template <typename T>
__global__ __launch_bounds__(UW) void upscale_kernel(UpLaunch a) {
  __shared__ int16_t filt[64 * 8];
  __shared__ uint16_t src[UW + 16];
  const int b = blockIdx.x;
  int pl = 0;
  if (b >= a.blk_start[1]) pl = 1;
  if (b >= a.blk_start[2]) pl = 2;
  const UpPlane &q = a.p[pl];
  const int lb = b - a.blk_start[pl];
  const int y = lb / q.blk_x, x0 = (lb - y * q.blk_x) * UW;
  for (int i = threadIdx.x; i < 64 * 8; i += UW) filt[i] = AV1_UPSCALE_FILTER[i >> 3][i & 7];
  /* source columns of this block: from px(x0) - 3 to px(x0 + UW - 1) + 4 */
  const int sx0 = (-(1 << SCALE_BITS) + q.init + x0 * q.step) >> SCALE_BITS;
  const T *row = gptr(reinterpret_cast<const T *>(q.in.data)) + (int64_t)y * q.in.stride;
  for (int i = threadIdx.x; i < UW + 16; i += UW) src[i] = row[min(max(sx0 - 3 + i, 0), q.max_x)];
  __syncthreads();
  const int x = x0 + threadIdx.x;
  if (x >= q.cols) return;
  const int sxq = -(1 << SCALE_BITS) + q.init + x * q.step;
  const int px = sxq >> SCALE_BITS;
  const int sub = (sxq & ((1 << SCALE_BITS) - 1)) >> EXTRA_BITS;
  /* sample px + k - 3 of the row (clamped) is src[px - sx0 + k] */
  const uint16_t *s = src + (px - sx0);
  const int16_t *f = filt + sub * 8;
  int sum = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) sum += __mul24((int)s[k], (int)f[k]);
  const int v = min(max((sum + 64) >> 7, 0), (1 << a.bd) - 1);
  gptr(reinterpret_cast<T *>(q.out.data))[(int64_t)y * q.out.stride + x] = (T)v;
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

}  // namespace

int launch_upscale(const Geom &g, int frame_width, int upscaled_width, const PlanePtr in[3], PlanePtr out[3],
                   const int cols[3], const int rows[3], int bps, hipStream_t st) {
  UpLaunch a;
  a.bd = g.bd;
  int total = 0;
  for (int p = 0; p < 3; p++) {
    UpPlane &q = a.p[p];
    q = UpPlane{};
    a.blk_start[p] = total;
    if (p >= g.nplanes || cols[p] <= 0 || rows[p] <= 0) continue;
    const int sx = p ? g.ssx : 0;
    int64_t step, init;
    plane_params(frame_width, upscaled_width, sx, &step, &init);
    q.in = in[p];
    q.out = out[p];
    q.cols = cols[p];
    q.rows = rows[p];
    q.max_x = (g.mi_cols >> sx) * 4 - 1;
    q.step = (int)step;
    q.init = (int)init;
    q.blk_x = (q.cols + UW - 1) / UW;
    total += q.blk_x * q.rows;
  }
  a.blk_start[3] = total;
  if (total == 0) return 0;
  if (bps == 1)
    upscale_kernel<uint8_t><<<total, UW, 0, st>>>(a);
  else
    upscale_kernel<uint16_t><<<total, UW, 0, st>>>(a);
  return (int)hipGetLastError();
}

}  // namespace rvpf
