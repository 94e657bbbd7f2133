/*
 * copy.hip — plane copies the chain needs where a stage is off (the
 * reference's FrameBuf clones, src/decoder/postfilter.rs:274, 617) and the
 * display crop without film grain (Decoder::output, src/decoder/mod.rs:893-915).
 */
#include "rvpf_internal.h"

namespace rvpf {

namespace {

struct CopyLaunch {
  PlanePtr in[3], out[3];
  int w[3], h[3];
  int blk_x[3];
  int blk_start[4];
};

/* Block b: 64 x 4 samples; rows of 64 lanes -> coalesced. Samples T are
 * stored as TO (8-bit output from 16-bit planes: the values fit). */
template <typename T, typename TO = T>
__device__ __forceinline__ void copy_block(const CopyLaunch &a, int b) {
  int p = 0;
  if (b >= a.blk_start[1]) p = 1;
  if (b >= a.blk_start[2]) p = 2;
  const int lb = b - a.blk_start[p];
  const int x = (lb % a.blk_x[p]) * 64 + (threadIdx.x & 63);
  const int y = (lb / a.blk_x[p]) * 4 + (threadIdx.x >> 6);
  if (x >= a.w[p] || y >= a.h[p]) return;
  gptr(reinterpret_cast<TO *>(a.out[p].data))[(int64_t)y * a.out[p].stride + x] =
      (TO)gptr(reinterpret_cast<const T *>(a.in[p].data))[(int64_t)y * a.in[p].stride + x];
}

template <typename T, typename TO>
__global__ __launch_bounds__(256) void copy_kernel(CopyLaunch a) {
  copy_block<T, TO>(a, blockIdx.x);
}

template <typename T, bool DIRECT>
__global__ __launch_bounds__(256) void copy_batch(BatchWork w) {
  batch_loop<DIRECT>(w, [&](int f, int j) { copy_block<T>(batch_table<CopyLaunch>(w.rec, f), j); });
}

int plan(const Geom &g, const PlanePtr in[3], const PlanePtr out[3], const int w[3], const int h[3],
         CopyLaunch &a) {
  int total = 0;
  for (int p = 0; p < 3; p++) {
    const bool on = p < g.nplanes && w[p] > 0 && h[p] > 0;
    a.in[p] = in[p];
    a.out[p] = out[p];
    a.w[p] = on ? w[p] : 0;
    a.h[p] = on ? h[p] : 0;
    a.blk_x[p] = on ? (w[p] + 63) / 64 : 1;
    a.blk_start[p] = total;
    if (on) total += a.blk_x[p] * ((h[p] + 3) / 4);
  }
  a.blk_start[3] = total;
  return total;
}

}  // namespace

size_t copy_record_size() { return sizeof(CopyLaunch); }

int copy_plan(const Geom &g, const PlanePtr in[3], const PlanePtr out[3], const int w[3], const int h[3],
              int, void *rec) {
  return plan(g, in, out, w, h, *static_cast<CopyLaunch *>(rec));
}

int launch_copy_batch(const BatchWork &w, const StagePlan &sp, int bps, hipStream_t st) {
  if (w.total == 0) return 0;
  if (bps == 1)
    if (batch_direct(w, sp.grid))
      copy_batch<uint8_t, true><<<sp.grid, 256, 0, st>>>(w);
    else
      copy_batch<uint8_t, false><<<sp.grid, 256, 0, st>>>(w);
  else
    if (batch_direct(w, sp.grid))
      copy_batch<uint16_t, true><<<sp.grid, 256, 0, st>>>(w);
    else
      copy_batch<uint16_t, false><<<sp.grid, 256, 0, st>>>(w);
  return (int)hipGetLastError();
}

int launch_copy_planes(const Geom &g, const PlanePtr in[3], PlanePtr out[3], const int w[3],
                       const int h[3], int bps, hipStream_t st, int out_bps) {
  if (out_bps == 0) out_bps = bps;
  if (out_bps > bps) return (int)hipErrorInvalidValue;
  CopyLaunch a;
  const int total = plan(g, in, out, w, h, a);
  if (total == 0) return 0;
  if (bps == 1)
    copy_kernel<uint8_t, uint8_t><<<total, 256, 0, st>>>(a);
  else if (out_bps == 1)
    copy_kernel<uint16_t, uint8_t><<<total, 256, 0, st>>>(a);
  else
    copy_kernel<uint16_t, uint16_t><<<total, 256, 0, st>>>(a);
  return (int)hipGetLastError();
}

}  // namespace rvpf
