/*
 * rvpf_internal.h — shared definitions of the HIP post-filter library.
 *
 * Device-side argument blocks (plain structs passed by value as kernel
 * arguments), the constant tables, and the launch entry points each stage
 * file provides to the host layer (rvpf_host.cpp).
 */
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/rvpf.h"

namespace rvpf {

/* Samples of one plane: base pointer + stride, typed at the launch site. */
struct PlanePtr {
  void *data;
  int64_t stride; /* samples */
};

/* Frame geometry every kernel needs. */
struct Geom {
  int bd;           /* bit depth */
  int ssx, ssy;     /* chroma subsampling */
  int nplanes;      /* 1 or 3 */
  int width, height; /* frame (= upscaled) width and height */
  int mi_rows, mi_cols;
};

/* Loop filter header data (7.14.4/7.14.5 inputs). */
struct LfHdr {
  int level[4];
  int sharpness;
  int delta_enabled;
  int ref_deltas[8];
  int mode_deltas[2];
  int delta_lf_multi;
  int seg_enabled;
  int seg_lf_enabled[8]; /* bit i-1 set: SEG_LVL_ALT_LF_Y_V + (i-1) enabled, i = 1..4 */
  int seg_lf_data[8][4];
};

struct LfSide {
  const rvpf_mi *mi;
  const uint8_t *tx[3];
  int mi_stride;
};

struct CdefHdr {
  int damping;
  int y_pri[8], y_sec[8], uv_pri[8], uv_sec[8];
};

struct LrHdr {
  int type[3];      /* frame_restoration_type per plane (0 = off) */
  int unit_size[3];
  int unit_rows[3], unit_cols[3];
  const rvpf_lr_unit *units[3];
};

/* Film grain parameters in the form the kernels read. */
struct GrainHdr {
  int bd, ssx, ssy, nplanes;
  int w, h; /* display size */
  uint32_t seed;
  int num_y, num_cb, num_cr, cfl; /* cfl = chroma_scaling_from_luma */
  int scaling_shift;
  int ar_lag, ar_shift, grain_scale_shift;
  int ar_y[24], ar_cb[25], ar_cr[25]; /* coefficients, minus 128 */
  int cb_mult, cb_luma_mult, cb_offset, cr_mult, cr_luma_mult, cr_offset;
  int overlap;
  int min_value, max_luma, max_chroma;
  int points[3][2][16]; /* [plane][value/scaling][i] */
  int npoints[3];
};

/* Device scratch for the grain stage. */
struct GrainScratch {
  int16_t *tmpl;    /* [3][73][82] */
  int16_t *lut;     /* [3][256] */
  uint8_t *offsets; /* [stripes][blocks]: (offset_x << 4) | offset_y */
  int max_stripes, max_blocks;
};

/* Workgroups are dispatched round-robin over the 8 XCDs (each with its
 * own L2): workgroup b runs on XCD b % 8. Renumbering so that XCD k
 * receives one contiguous range of the n tiles keeps neighbouring tiles,
 * and the halo lines they share, in one L2. A bijection of [0, n). */
__device__ __forceinline__ int xcd_remap(int b, int n) {
  constexpr int NX = 8;
  const int q = n / NX, r = n % NX;
  const int k = b % NX, j = b / NX;
  return k * q + min(k, r) + j;
}

/* xcd_remap for a grid made of segments (one per plane) of unequal cost:
 * each segment's block count is padded to a multiple of 8 (pad[s] =
 * first block of segment s, pad[nseg] = grid size), and XCD k gets the
 * k-th contiguous eighth of every segment, so every XCD sees the same mix.
 * Returns the segment of block b and sets *local to its index there. */
__device__ __forceinline__ int xcd_segment(int b, const int *pad, int nseg, int *local) {
  const int k = b & 7, j = b >> 3;
  int acc = 0;
  for (int s = 0; s < nseg; s++) {
    const int n8 = (pad[s + 1] - pad[s]) >> 3;
    if (j < acc + n8) {
      *local = k * n8 + (j - acc);
      return s;
    }
    acc += n8;
  }
  *local = 0;
  return nseg; /* not reached for b < pad[nseg] */
}

/* Launch entry points (each returns a hipError_t cast to int). */
int launch_deblock(const Geom &g, const LfHdr &h, const LfSide &s, PlanePtr planes[3], int bps,
                   hipStream_t st);
int launch_cdef(const Geom &g, const CdefHdr &h, const rvpf_mi *mi, int mi_stride,
                const int8_t *cdef_idx, int cdef_stride, int enabled, const PlanePtr in[3],
                PlanePtr out[3], int bps, hipStream_t st);
int launch_lr(const Geom &g, const LrHdr &h, const PlanePtr deblocked[3], const PlanePtr cdef[3],
              PlanePtr out[3], int bps, hipStream_t st);
int launch_copy_planes(const Geom &g, const PlanePtr in[3], PlanePtr out[3], const int w[3],
                       const int h[3], int bps, hipStream_t st);
int launch_grain_prepare(const GrainHdr &h, const GrainScratch &s, hipStream_t st);
int launch_grain_apply(const GrainHdr &h, const GrainScratch &s, const PlanePtr in[3],
                       PlanePtr out[3], int bps, hipStream_t st);

/* Constant-table upload (once per device). */
int upload_tables();

}  // namespace rvpf
