/*
 * synth.h — seeded synthetic workloads for the post-filter chain.
 *
 * Produces a source picture, a blockwise-quantised "reconstruction" and
 * spec-legal side info (partitions to 4x4 with 64 or 128 superblocks,
 * aligned transform grids, skip / intra / inter mix, loop filter levels,
 * CDEF strengths and indices, restoration units, film grain parameters),
 * in the boundary layout of include/rvpf.h. Host memory only.
 *
 * The source follows the benchmark's description: a sum of 4 random 2D
 * sinusoids (periods 8..64 px), 3 moving rectangles of random level and
 * Gaussian noise (sigma 6 at 8 bits), clipped, scaled by 2^(bd-8).
 */
#ifndef RVPF_SYNTH_H
#define RVPF_SYNTH_H

#include <stdint.h>

#include "../../include/rvpf.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct synth_cfg {
  int32_t width, height;
  int32_t bit_depth;
  int32_t ssx, ssy, mono;
  int32_t sb_size;   /* 64 or 128 */
  int32_t qindex;    /* 0..255 */
  uint64_t seed;
  int32_t frame_index;   /* frame number in the GOP (motion, key/inter) */
  double pan_px;         /* global pan per frame, pixels */
  int32_t lf;            /* deblocking on */
  int32_t lf_extras;     /* sharpness, delta LF, segmentation features */
  int32_t cdef;          /* CDEF on */
  int32_t cdef_bits;     /* 0..3 */
  int32_t lr;            /* 0 off, 1 Wiener on luma, 2 Wiener/SGR/none mixed on all planes */
  int32_t lr_unit_size;  /* luma unit size 64, 128 or 256 (else 64); 4:2:0 chroma halves it */
  int32_t grain;         /* film grain on */
  int32_t grain_lag;     /* ar_coeff_lag */
  int32_t grain_y_points;
  int32_t mc_identity;   /* matrix_coefficients = 0 */
  int32_t superres_denom; /* 0: off; 9..16: coded width = (width * 8 + denom / 2) / denom, upscaled to width */
} synth_cfg;

typedef struct synth_geom {
  int32_t mi_rows, mi_cols;
  int32_t mi_stride, mi_alloc_rows; /* 4x4 arrays: mi_alloc_rows x mi_stride */
  int32_t cdef_stride, cdef_rows;
  int32_t plane_w[3], plane_h[3]; /* allocated plane width (= stride) and rows */
  int32_t crop_w[3], crop_h[3];   /* display plane sizes */
  int32_t lr_unit_rows[3], lr_unit_cols[3];
  int32_t num_planes;
  int32_t coded_width;            /* frame_width: width, or the super-resolution coded width */
  int32_t ref_w[3];               /* allocated width of a reference-frame (upscaled) plane */
} synth_geom;

int synth_geometry(const synth_cfg *c, synth_geom *g);

/* Fills everything (arrays sized per synth_geometry). Any output pointer
 * may be NULL to skip it. Returns 0 or -1 on a bad configuration. */
int synth_frame(const synth_cfg *c, rvpf_frame_params *p, rvpf_mi *mi, uint8_t *tx0,
                uint8_t *tx1, uint8_t *tx2, int8_t *cdef_idx, rvpf_lr_unit *lr0,
                rvpf_lr_unit *lr1, rvpf_lr_unit *lr2, uint16_t *src0, uint16_t *src1,
                uint16_t *src2, uint16_t *rec0, uint16_t *rec1, uint16_t *rec2);

/* Struct sizes, for binding checks: frame_params, side_info, mi, lr_unit,
 * grain_params, planes. */
void synth_struct_sizes(int32_t out[11]);

#ifdef __cplusplus
}
#endif

#endif
