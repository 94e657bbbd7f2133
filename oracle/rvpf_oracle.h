/*
 * rvpf_oracle.h — CPU oracle for the post-filter chain.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of rivet-av1's
 * CPU post-filters (src/decoder/postfilter.rs, src/dsp/lf.rs,
 * src/dsp/cdef.rs, src/decoder/grain.rs), used as the parity checker for
 * the HIP path and as the CPU baseline. The product library never links,
 * loads or calls it; only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg do.
 *
 * All pointers here are HOST pointers; the structs are the ones of
 * include/rvpf.h (an rvpf_side_info's arrays are host arrays here).
 * Planes are uint16_t with a row stride in samples.
 *
 * Deviation from the reference outside the parity region: CDEF samples
 * whose own position lies beyond the mode-info edge (the right/bottom
 * columns of an 8x8 block when mi_cols/mi_rows is odd) keep their input
 * value here; the reference computes them from an "unavailable" marker with
 * wrapping i32 arithmetic. No later stage and no output reads them.
 */
#ifndef RVPF_ORACLE_H
#define RVPF_ORACLE_H

#include <stdint.h>

#include "../include/rvpf.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_planes {
  uint16_t *data[3];
  int64_t stride[3];
} orc_planes;

/* Super-resolution upscaling (7.16, src/decoder/postfilter.rs:521-566):
 * in (frame_width wide) -> out (upscaled_width wide), rows of the frame
 * height; the rest of out is left as it is. */
int orc_upscale(const rvpf_frame_params *p, const orc_planes *in, orc_planes *out);

/* 7.14.6 on four lanes, v[k][lane], q0 at k = 7 (src/dsp/lf.rs:58-211). */
void orc_lf_filter_lanes(int32_t v[14][4], int limit, int blimit, int thresh,
                         int filter_size, int chroma, int bit_depth);

/* 7.14.4/5: lvl, limit, blimit, thresh of the block at (row, col)
 * (src/decoder/postfilter.rs:207-264). */
void orc_filter_strength(const rvpf_frame_params *p, const rvpf_side_info *s, int row, int col,
                         int plane, int pass, int out[4]);

/* The loop filter, in place (src/decoder/postfilter.rs:67-203). */
int orc_loop_filter(const rvpf_frame_params *p, const rvpf_side_info *s, orc_planes *frame,
                    int threads);

/* 7.15.2 direction search on the 8x8 luma block at pixel (x0, y0)
 * (src/decoder/postfilter.rs:345-401). Returns the direction; *var gets var. */
int orc_cdef_direction(const uint16_t *luma, int64_t stride, int x0, int y0, int bit_depth,
                       int32_t *var);

/* 7.15.3 filter of the w x h block at position (2, 2) of a 12x12 window
 * (INT32_MIN = unavailable) into out (row stride w); src/dsp/cdef.rs:54-90. */
void orc_cdef_filter_window(const int32_t *win, int w, int h, int pri_str, int sec_str, int damping,
                            int dir, int coeff_shift, int32_t *out);

/* CDEF of a frame: out = in, then every filtered 8x8 block
 * (src/decoder/postfilter.rs:273-513, src/dsp/cdef.rs:48-90). */
int orc_cdef(const rvpf_frame_params *p, const rvpf_side_info *s, const orc_planes *in,
             orc_planes *out, int threads);

/* Loop restoration: out = cdef, then every restored stripe x unit
 * (src/decoder/postfilter.rs:608-1004). */
int orc_lr(const rvpf_frame_params *p, const rvpf_side_info *s, const orc_planes *deblocked,
           const orc_planes *cdef, orc_planes *out, int threads);

/* Encoder loop restoration search (src/encoder/lr.rs:64-430), for one
 * unit (plane, unit row ur, unit column uc) of a frame whose deblocked,
 * CDEF and source planes are given.
 * fit_wiener (lr.rs:240-325): quantised taps out[0] vertical, out[1]
 * horizontal, and (fit != NULL) the unquantised fractions. */
int orc_fit_wiener(const rvpf_frame_params *f, const orc_planes *deb, const orc_planes *cdef,
                   const orc_planes *src, int plane, int ur, int uc, int step, int32_t out[2][3],
                   double fit[2][3]);
/* fit_sgr (lr.rs:332-430): quantised xqd, the unquantised (w0, w2) and the
 * exact error with xqd. */
int orc_fit_sgr(const rvpf_frame_params *f, const orc_planes *deb, const orc_planes *cdef, const orc_planes *src,
                int plane, int ur, int uc, int set, int step, int32_t xqd[2], double wfit[2], uint64_t *sse);
/* Exact error of the unit restored with u (u NULL or type none: the CDEF
 * output) against the source (lr.rs:128-153). */
uint64_t orc_unit_sse(const rvpf_frame_params *f, const orc_planes *deb, const orc_planes *cdef,
                      const orc_planes *src, int plane, int ur, int uc, const rvpf_lr_unit *u);
/* The unit's choice (lr.rs:153-186) among none, the fitted Wiener filter
 * and the fitted self-guided filter of each set in sets[]. */
int orc_lr_search_unit(const rvpf_frame_params *f, const orc_planes *deb, const orc_planes *cdef,
                       const orc_planes *src, int plane, int ur, int uc, double lambda, const int *sets,
                       int nsets, int step, rvpf_lr_unit *out, double *cost_out);

/* The encoder's CDEF search candidates (cdef.rs:60-84): luma (pri, sec)
 * and chroma (pri, sec), secondary as coded (3 = strength 4). */
int orc_cdef_candidates(int thorough, int32_t ypri[44], int32_t ysec[44], int32_t uvpri[24], int32_t uvsec[24],
                        int *ny, int *nuv);
/* cdef::search (src/encoder/cdef.rs:52-253) on the deblocked frame `deb`
 * against `src`: the chosen parameters, the cdef_idx table (layout of
 * s->cdef_idx, fb_rows x cdef_stride) and, when not NULL, the per-block
 * error tables (coded blocks in raster order, ny / nuv per block) and the
 * number of coded blocks. */
int orc_cdef_search(const rvpf_frame_params *p, const rvpf_side_info *s, const orc_planes *deb,
                    const orc_planes *src, int base_q_idx, double lambda, int thorough, int threads,
                    rvpf_cdef_search_result *out, int8_t *table, uint64_t *ysse_out, uint64_t *uvsse_out,
                    int *nb_out);
/* search_lf (src/encoder/mod.rs:1460-1490): levels from the start levels
 * p->loop_filter_level; sse[scale][plane] when not NULL. recon is not
 * modified. */
int orc_lf_search(const rvpf_frame_params *p, const rvpf_side_info *s, const orc_planes *recon,
                  const orc_planes *src, int threads, int32_t out[4], uint64_t sse[6][3]);

/* Film grain templates (7.18.3.3), grain.rs:40-135: luma 73x82 and
 * cb/cr (rows x 82 used up to the chroma template size), row stride 82. */
void orc_grain_templates(const rvpf_frame_params *p, int32_t luma[73 * 82], int32_t cb[73 * 82],
                         int32_t cr[73 * 82]);

/* Scaling LUTs (7.18.3.4), grain.rs:137-165: 3 x 256. */
void orc_grain_scaling_luts(const rvpf_frame_params *p, int32_t lut[3][256]);

/* Film grain on tight planes (luma w x h, chroma rounded up), in place
 * (src/decoder/grain.rs:24-354). */
int orc_grain_apply(const rvpf_frame_params *p, int w, int h, uint16_t *planes[3]);

/* The chain (src/decoder/mod.rs:747-917): deblock in place on recon,
 * CDEF, LR into ref_out, then crop (+ grain) into display (tight planes,
 * may be NULL). */
int orc_postfilter_frame(const rvpf_frame_params *p, const rvpf_side_info *s, orc_planes *recon,
                         orc_planes *ref_out, uint16_t *display[3], int threads);

/* LFSR of 7.18.3.2 (grain.rs:8-19): one step; returns the new 16-bit state. */
uint32_t orc_grain_rng_step(uint32_t state);

#ifdef __cplusplus
}
#endif

#endif
