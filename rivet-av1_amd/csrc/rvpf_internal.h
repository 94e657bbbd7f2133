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

/* Film grain templates in device scratch: per plane, 8 copies of the
 * 73 x 82 template, copy s shifted left by s samples, rows padded to 88,
 * so that any 8 consecutive template samples are one aligned 16-byte
 * load (copy 0 is the template itself). */
constexpr int GRAIN_TH = 73, GRAIN_TW = 82, GRAIN_TWS = 88;
constexpr int GRAIN_TCOPY = GRAIN_TH * GRAIN_TWS;  /* samples per shifted copy */
constexpr int GRAIN_TPLANE = 8 * GRAIN_TCOPY;      /* samples per plane */
constexpr size_t GRAIN_TMPL_BYTES = 3 * (size_t)GRAIN_TPLANE * sizeof(int16_t);
/* Scaling function entries per plane: one per sample value up to 12 bits. */
constexpr int GRAIN_LUT_N = 4096;
constexpr size_t GRAIN_LUT_BYTES = 3 * (size_t)GRAIN_LUT_N * sizeof(int16_t);

/* Device scratch for the grain stage. */
struct GrainScratch {
  int16_t *tmpl;    /* [3][8][73][88], see GRAIN_TPLANE */
  int16_t *lut;     /* [3][GRAIN_LUT_N]: the scaling function at every sample value */
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

/* --- batched launches ------------------------------------------------
 * One launch of a stage over many frames. Every frame contributes `items`
 * work items (tiles) described by its own launch record; the records sit
 * in a device array and start[f] is the first item of frame f (prefix
 * sums, start[n] = total). A fixed grid of workgroups takes `chunk` items
 * at a time from a device counter until all are taken. */
struct BatchWork {
  const void *rec;  /* device: n launch records of the stage's type */
  const int *start; /* device: n + 1 prefix sums of work items */
  int n, total;
  int chunk;        /* items per grab */
  int *queue;       /* device, 2 ints: next item, workgroups finished */
};

/* A batch table is read-only while a batched kernel runs; reading it
 * through the constant address space lets the compiler use scalar loads
 * for the (wave-uniform) record fields, and keeps the plane pointers
 * stored in them known as global memory. */
template <class R>
__device__ __forceinline__ const R &batch_table(const void *p, int i) {
  typedef const R __attribute__((address_space(4))) ConstR;
  return *(const R *)((ConstR *)p + i);
}

/* A pointer known to address global memory (plane and side-info pointers
 * read from a record): the round trip through the global address space
 * lets the compiler emit global rather than flat memory instructions. */
template <class T>
__device__ __forceinline__ T *gptr(T *p) {
  typedef T __attribute__((address_space(1))) GlobalT;
  return (T *)((GlobalT *)p);
}

/* Frame of item i: the last f with start[f] <= i. */
__device__ __forceinline__ int batch_frame(const int *start, int n, int i) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (batch_table<int>(start, mid) <= i)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

/* The grab loop of a batched kernel: body(f, j) runs item j of frame f
 * with the whole workgroup (f and j are wave-uniform). A grid with a
 * workgroup for every item (DIRECT, batch_direct() on the host) needs no
 * queue and no loop: workgroup b runs item xcd_remap(b) (neighbouring
 * tiles on one XCD, as the per-frame kernels map theirs). Without a loop
 * around the tile code the compiler allocates registers as for the
 * per-frame kernel (inside a loop it hoists each tile's per-thread set-up
 * out of the body and takes about twice as many VGPRs). A smaller
 * (persistent) grid, or chunks of several items, takes chunks from the
 * queue, the next grab in flight while the current chunk runs, so the
 * atomic's round trip is not waited for between chunks; a grabbed chunk
 * below the total is always run. The queue is re-armed by the last
 * workgroup to finish, so the next launch of the same work (a graph
 * replay) starts from zero without a memset. Every workgroup leaves once
 * its grab passes the total. */
template <class Body>
__device__ __forceinline__ void batch_run(const BatchWork &w, int i0, int i1, Body &body) {
  for (int i = i0; i < i1; i++) {
    const int f = __builtin_amdgcn_readfirstlane(batch_frame(w.start, w.n, i));
    const int j = __builtin_amdgcn_readfirstlane(i - batch_table<int>(w.start, f));
    body(f, j);
    __syncthreads(); /* the next item reuses the LDS */
  }
}

template <bool DIRECT, class Body>
__device__ __forceinline__ void batch_loop(const BatchWork &w, Body &&body) {
  if (DIRECT) {
    const int i = xcd_remap(blockIdx.x, gridDim.x);
    if (i < w.total) {
      const int f = __builtin_amdgcn_readfirstlane(batch_frame(w.start, w.n, i));
      const int j = __builtin_amdgcn_readfirstlane(i - batch_table<int>(w.start, f));
      body(f, j);
    }
    return;
  }
  __shared__ int s_item;
  int next = 0;
  if (threadIdx.x == 0) next = atomicAdd(&w.queue[0], w.chunk);
  for (;;) {
    if (threadIdx.x == 0) s_item = next;
    __syncthreads();
    const int i0 = s_item;
    __syncthreads();
    if (i0 >= w.total) break;
    if (threadIdx.x == 0) next = atomicAdd(&w.queue[0], w.chunk);
    batch_run(w, i0, min(i0 + w.chunk, w.total), body);
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&w.queue[1], 1) == (int)gridDim.x - 1) {
      atomicExch(&w.queue[0], 0);
      atomicExch(&w.queue[1], 0);
    }
  }
}

/* Whether a launch of `grid` workgroups over w has one per item. */
inline bool batch_direct(const BatchWork &w, int grid) { return w.chunk == 1 && grid >= w.total; }

/* Tile shapes a batched stage can run with (the autotuner's choices). */
struct StagePlan {
  int shape;       /* stage-specific tile shape index */
  int grid;        /* workgroups of the launch */
  int chunk;       /* items per grab */
};

/* Per-stage batch interface: record size, planning one frame into a
 * record (returns its work items), and the batched launch. */
size_t lf_record_size();
int lf_plan(const Geom &g, const LfHdr &h, const LfSide &s, const PlanePtr planes[3], int shape, void *rec);
int launch_deblock_batch(const BatchWork &wv, const BatchWork &wh, const StagePlan &sp, int bps, hipStream_t st);

size_t cdef_record_size();
int cdef_plan(const Geom &g, const CdefHdr &h, const rvpf_mi *mi, int mi_stride, const int8_t *cdef_idx,
              int cdef_stride, int enabled, const PlanePtr in[3], const PlanePtr out[3], void *rec);
/* 1 when the record's planes take CDEF's chunked staging */
int cdef_record_vec(const void *rec);
int launch_cdef_batch(const BatchWork &w, const StagePlan &sp, int bps, int vec, hipStream_t st);

size_t lr_record_size();
int lr_plan(const Geom &g, const LrHdr &h, const PlanePtr deblocked[3], const PlanePtr cdef[3],
            const PlanePtr out[3], int bps, void *rec);
int launch_lr_batch(const BatchWork &w, const StagePlan &sp, int bps, hipStream_t st);

size_t copy_record_size();
int copy_plan(const Geom &g, const PlanePtr in[3], const PlanePtr out[3], const int w[3], const int h[3],
              int shape, void *rec);
int launch_copy_batch(const BatchWork &w, const StagePlan &sp, int bps, hipStream_t st);

size_t grain_prep_record_size();
int grain_prep_plan(const GrainHdr &h, const GrainScratch &s, void *rec);
int launch_grain_prepare_batch(const BatchWork &w, const StagePlan &sp, hipStream_t st);

size_t grain_apply_record_size();
int grain_apply_plan(const GrainHdr &h, const GrainScratch &s, const PlanePtr in[3], const PlanePtr out[3],
                     int bps, void *rec);
int launch_grain_apply_batch(const BatchWork &w, const StagePlan &sp, int bps, hipStream_t st);

/* Number of tile shapes of each shape-tunable stage. */
constexpr int LF_SHAPES = 3;    /* 64x4, 32x8, 16x16 edges per workgroup */
constexpr int GRAIN_SHAPES = 1; /* one: a noise stripe's 64 luma columns (128 at 4:2:0) per workgroup */
constexpr int COPY_SHAPES = 1;

/* Differentiable loop restoration (lrdiff.hip). */
int launch_lrd_gather(const PlanePtr deb[3], const PlanePtr cdef[3], const PlanePtr src[3], int bps,
                      const int pex[3], const int pey[3], const rvpf_lrd_item *items, int n, int nunits,
                      float *win, float *u, float *s, hipStream_t st);
int launch_lrd_wiener(int backward, const rvpf_lrd_item *items, int n, int nunits, const float *win,
                      const float *src, const float *taps, float *resid, double *out, hipStream_t st);
int launch_lrd_sgr_filter(const rvpf_lrd_item *items, int n, int nunits, const float *win, const float *u,
                          const int32_t *sets, int bd, float *flt, hipStream_t st);
int launch_lrd_sgr(int backward, const rvpf_lrd_item *items, int n, int nunits, const float *flt,
                   const float *u, const float *src, const float *w, float *resid, double *out, hipStream_t st);
int launch_lrd_rect_sse(const rvpf_lrd_item *items, int n, int nunits, const int pex[3], const int pey[3],
                        const PlanePtr a[3], const PlanePtr b[3], int bps, double *out, hipStream_t st);

/* Launch entry points (each returns a hipError_t cast to int). */
int launch_deblock(const Geom &g, const LfHdr &h, const LfSide &s, PlanePtr planes[3], int bps,
                   hipStream_t st);
/* Both passes in one launch from `in` into `out` (the mode-info area of
 * each filtered plane; `in` keeps its samples). */
int launch_deblock_fused(const Geom &g, const LfHdr &h, const LfSide &s, const PlanePtr in[3], const PlanePtr out[3],
                         int bps, hipStream_t st);
int launch_cdef(const Geom &g, const CdefHdr &h, const rvpf_mi *mi, int mi_stride,
                const int8_t *cdef_idx, int cdef_stride, int enabled, const PlanePtr in[3],
                PlanePtr out[3], int bps, hipStream_t st);
int launch_lr(const Geom &g, const LrHdr &h, const PlanePtr deblocked[3], const PlanePtr cdef[3],
              PlanePtr out[3], int bps, hipStream_t st);
/* out_bps: bytes per output sample when it differs from the input's
 * (1 from 2: 8-bit frames written as bytes); 0 = the same. */
int launch_copy_planes(const Geom &g, const PlanePtr in[3], PlanePtr out[3], const int w[3],
                       const int h[3], int bps, hipStream_t st, int out_bps = 0);
int launch_grain_prepare(const GrainHdr &h, const GrainScratch &s, hipStream_t st);
int launch_grain_apply(const GrainHdr &h, const GrainScratch &s, const PlanePtr in[3],
                       PlanePtr out[3], int bps, hipStream_t st, int out_bps = 0);

/* The encoder's CDEF search (cdef.hip): for every 64x64 block that codes
 * an index (cdef_idx >= 0), the squared error against src of each luma
 * candidate (ypri/ysec strengths, ny) and each chroma candidate (nuv),
 * into ysse[block * ny + k] / uvsse[block * nuv + k]; coded[block] = 0/1. */
int launch_cdef_search(const Geom &g, const rvpf_mi *mi, int mi_stride, const int8_t *cdef_idx, int cdef_stride,
                       const PlanePtr in[3], const PlanePtr src[3], int damping, const int8_t *ypri,
                       const int8_t *ysec, int ny, const int8_t *uvpri, const int8_t *uvsec, int nuv, uint64_t *ysse,
                       uint64_t *uvsse, int8_t *coded, int bps, hipStream_t st);

/* One plane of a super-resolution upscaling launch: rows 0..rows-1 and
 * columns 0..cols-1 of `out` from `in` (plane index 0..2). */
struct UpscaleJob {
  PlanePtr in, out;
  int plane;
  int cols, rows;
};

/* Super-resolution upscaling (upscale.hip) of up to six plane jobs in one
 * launch; g is the coded (downscaled) frame's geometry, whose mode-info
 * width bounds the columns read. */
int launch_upscale(const Geom &g, int frame_width, int upscaled_width, const UpscaleJob *jobs, int njobs, int bps,
                   hipStream_t st);

/* Constant-table upload (once per device). */
int upload_tables();

}  // namespace rvpf
