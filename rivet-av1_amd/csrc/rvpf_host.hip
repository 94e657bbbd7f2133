/*
 * rvpf_host.hip — the C ABI of include/rvpf.h: validation, device scratch,
 * parameter packing and the stage order of the chain.
 *
 * The chain order is Decoder::decode_frame_wrapup's (src/decoder/mod.rs:747-816):
 * loop filter only if loop_filter_level[0] or [1] is non-zero, CDEF
 * (identity when coded_lossless, allow_intrabc or !enable_cdef), loop
 * restoration (identity when no plane uses it), then Decoder::output's crop
 * and film grain (src/decoder/mod.rs:846-917). With super-resolution the
 * CDEF output and the deblocked frame are upscaled (mod.rs:752-757) and loop
 * restoration runs on the upscaled frames.
 */
#include <stdio.h>
#include <string.h>

#include <new>
#include <vector>

#include "av1_tables.h"
#include "rvpf_host.h"

using namespace rvpf;

namespace {
thread_local char g_err[256] = "";
}  // namespace

namespace rvpf {
namespace host {

int fail(rvpf_ctx *ctx, int code, const char *msg) {
  char *dst = ctx ? ctx->err : g_err;
  snprintf(dst, 256, "%s", msg);
  return code;
}

int hip_fail(rvpf_ctx *ctx, hipError_t e, const char *where) {
  char buf[256];
  snprintf(buf, sizeof buf, "%s: %s", where, hipGetErrorString(e));
  return fail(ctx, RVPF_E_DEVICE, buf);
}

int count_units(int unit, int size) {
  int n = (size + (unit >> 1)) / unit;
  return n < 1 ? 1 : n;
}

bool increasing(const int32_t *v, int n) {
  for (int i = 1; i < n; i++)
    if (v[i] <= v[i - 1]) return false;
  return true;
}

bool in_range(const int32_t *v, int n, int lo, int hi) {
  for (int i = 0; i < n; i++)
    if (v[i] < lo || v[i] > hi) return false;
  return true;
}

int nplanes_of(const rvpf_frame_params *p) { return p->mono_chrome ? 1 : 3; }

bool lf_on(const rvpf_frame_params *p) { return p->loop_filter_level[0] != 0 || p->loop_filter_level[1] != 0; }

bool cdef_on(const rvpf_frame_params *p) {
  return !(p->coded_lossless || p->allow_intrabc || !p->enable_cdef);
}

bool lr_plane_on(const rvpf_frame_params *p, int pl) { return p->frame_restoration_type[pl] != 0; }

bool lr_on(const rvpf_frame_params *p) {
  for (int pl = 0; pl < nplanes_of(p); pl++)
    if (lr_plane_on(p, pl)) return true;
  return false;
}

bool grain_on(const rvpf_frame_params *p) { return p->film_grain_params_present && p->grain.apply_grain; }

void plane_ss(const rvpf_frame_params *p, int pl, int *sx, int *sy) {
  *sx = pl ? p->ssx : 0;
  *sy = pl ? p->ssy : 0;
}

/* the mode-info area of a plane */
void area(const rvpf_frame_params *p, int pl, int *w, int *h) {
  int sx, sy;
  plane_ss(p, pl, &sx, &sy);
  *w = (p->mi_cols * 4) >> sx;
  *h = (p->mi_rows * 4) >> sy;
}

void crop(const rvpf_frame_params *p, int pl, int *w, int *h) {
  int sx, sy;
  plane_ss(p, pl, &sx, &sy);
  *w = (p->upscaled_width + sx) >> sx;
  *h = (p->frame_height + sy) >> sy;
}

/* the area loop restoration reads and writes: the mode-info area, or with
 * super-resolution the upscaled frame's (upscaled_width rounded up to 8) */
void lr_area(const rvpf_frame_params *p, int pl, int *w, int *h) {
  area(p, pl, w, h);
  if (p->use_superres) {
    int sx, sy;
    plane_ss(p, pl, &sx, &sy);
    *w = (2 * ((p->upscaled_width + 7) >> 3) * 4) >> sx;
  }
}

int validate_params(const rvpf_frame_params *p, char *msg) {
#define BAD(...)                          \
  do {                                    \
    snprintf(msg, 256, __VA_ARGS__);      \
    return RVPF_E_INVALID;                \
  } while (0)
  if (!p) BAD("params is NULL");
  if (p->use_superres) {
    /* superres_params (7.21): frame_width = (upscaled_width * 8 + d / 2) / d, d in 9..16 */
    bool ok = false;
    for (int d = 9; d <= 16; d++) ok |= p->frame_width == (p->upscaled_width * 8 + d / 2) / d;
    if (!ok) BAD("super-resolution: frame_width %d is not upscaled_width %d * 8 / d for d in 9..16", p->frame_width,
                 p->upscaled_width);
  } else if (p->upscaled_width != p->frame_width) {
    BAD("upscaled_width %d differs from frame_width %d without super-resolution", p->upscaled_width, p->frame_width);
  }
  if (p->bit_depth != 8 && p->bit_depth != 10 && p->bit_depth != 12) BAD("bit_depth %d", p->bit_depth);
  if (p->ssx < 0 || p->ssx > 1 || p->ssy < 0 || p->ssy > 1 || (p->ssx == 0 && p->ssy == 1))
    BAD("subsampling %d,%d", p->ssx, p->ssy);
  if (p->frame_width < 1 || p->frame_height < 1 || p->frame_width > 65536 || p->frame_height > 65536)
    BAD("frame size %dx%d", p->frame_width, p->frame_height);
  if (p->mi_cols != 2 * ((p->frame_width + 7) >> 3) || p->mi_rows != 2 * ((p->frame_height + 7) >> 3))
    BAD("mi_rows/mi_cols %d/%d do not match the frame size", p->mi_rows, p->mi_cols);
  if (!in_range(p->loop_filter_level, 4, 0, 63)) BAD("loop_filter_level out of range");
  if (p->loop_filter_sharpness < 0 || p->loop_filter_sharpness > 7) BAD("loop_filter_sharpness");
  if (!in_range(p->loop_filter_ref_deltas, 8, -64, 63) || !in_range(p->loop_filter_mode_deltas, 2, -64, 63))
    BAD("loop filter deltas out of range");
  for (int s = 0; s < 8; s++)
    if (!in_range(&p->seg_feature_data[s][1], 4, -63, 63)) BAD("segment loop filter data out of range");
  if (cdef_on(p)) {
    if (p->cdef_damping < 3 || p->cdef_damping > 6) BAD("cdef_damping %d", p->cdef_damping);
    for (int i = 0; i < 8; i++) {
      if (p->cdef_y_pri_strength[i] < 0 || p->cdef_y_pri_strength[i] > 15 || p->cdef_uv_pri_strength[i] < 0 ||
          p->cdef_uv_pri_strength[i] > 15)
        BAD("cdef primary strength out of range");
      const int ys = p->cdef_y_sec_strength[i], us = p->cdef_uv_sec_strength[i];
      if (!(ys == 0 || ys == 1 || ys == 2 || ys == 4) || !(us == 0 || us == 1 || us == 2 || us == 4))
        BAD("cdef secondary strength must be 0, 1, 2 or 4");
    }
  }
  for (int pl = 0; pl < nplanes_of(p); pl++) {
    if (p->frame_restoration_type[pl] < 0 || p->frame_restoration_type[pl] > 3) BAD("frame_restoration_type");
    if (!lr_plane_on(p, pl)) continue;
    /* lr_params (src/header.rs:1026-1043): luma 64, 128 or 256; chroma the
     * same or, with 4:2:0, half of it. A 64-row stripe then lies in one unit row. */
    const int us = p->loop_restoration_size[pl];
    const bool half_ok = pl > 0 && p->ssx && p->ssy && us == 32;
    if (us != 64 && us != 128 && us != 256 && !half_ok) BAD("loop_restoration_size[%d] = %d", pl, us);
    int cw, ch;
    crop(p, pl, &cw, &ch);
    if (p->lr_unit_rows[pl] != count_units(us, ch) || p->lr_unit_cols[pl] != count_units(us, cw))
      BAD("lr_unit_rows/cols of plane %d do not match count_units_in_frame", pl);
  }
  if (grain_on(p)) {
    const rvpf_grain_params *g = &p->grain;
    if (g->num_y_points < 0 || g->num_y_points > 14) BAD("num_y_points > 14");
    if (g->num_cb_points < 0 || g->num_cb_points > 10 || g->num_cr_points < 0 || g->num_cr_points > 10)
      BAD("num_cb_points / num_cr_points > 10");
    if (!in_range(g->point_y_value, g->num_y_points, 0, 255) || !in_range(g->point_y_scaling, g->num_y_points, 0, 255) ||
        !in_range(g->point_cb_value, g->num_cb_points, 0, 255) || !in_range(g->point_cb_scaling, g->num_cb_points, 0, 255) ||
        !in_range(g->point_cr_value, g->num_cr_points, 0, 255) || !in_range(g->point_cr_scaling, g->num_cr_points, 0, 255))
      BAD("film grain points out of range");
    if (!increasing(g->point_y_value, g->num_y_points) || !increasing(g->point_cb_value, g->num_cb_points) ||
        !increasing(g->point_cr_value, g->num_cr_points))
      BAD("film grain point values do not increase");
    if (p->mono_chrome && (g->chroma_scaling_from_luma || g->num_cb_points || g->num_cr_points))
      BAD("chroma grain on a monochrome frame");
    if (g->grain_seed > 0xffff) BAD("grain_seed");
    if (g->grain_scaling_minus_8 < 0 || g->grain_scaling_minus_8 > 3 || g->ar_coeff_lag < 0 || g->ar_coeff_lag > 3 ||
        g->ar_coeff_shift_minus_6 < 0 || g->ar_coeff_shift_minus_6 > 3 || g->grain_scale_shift < 0 ||
        g->grain_scale_shift > 3)
      BAD("film grain shifts out of range");
    if (!in_range(g->ar_coeffs_y_plus_128, 24, 0, 255) || !in_range(g->ar_coeffs_cb_plus_128, 25, 0, 255) ||
        !in_range(g->ar_coeffs_cr_plus_128, 25, 0, 255))
      BAD("film grain AR coefficients out of range");
    if (g->cb_mult < 0 || g->cb_mult > 255 || g->cb_luma_mult < 0 || g->cb_luma_mult > 255 || g->cb_offset < 0 ||
        g->cb_offset > 511 || g->cr_mult < 0 || g->cr_mult > 255 || g->cr_luma_mult < 0 || g->cr_luma_mult > 255 ||
        g->cr_offset < 0 || g->cr_offset > 511)
      BAD("film grain chroma multipliers out of range");
  }
  return RVPF_OK;
#undef BAD
}

int validate_side(const rvpf_frame_params *p, const rvpf_side_info *s, char *msg) {
  const bool need_mi = lf_on(p) || cdef_on(p);
  if (need_mi) {
    if (!s || !s->mi) {
      snprintf(msg, 256, "side info: mi is NULL");
      return RVPF_E_INVALID;
    }
    if (s->mi_stride < p->mi_cols + 1) {
      snprintf(msg, 256, "side info: mi_stride %d < mi_cols + 1", s->mi_stride);
      return RVPF_E_INVALID;
    }
  }
  if (lf_on(p))
    for (int pl = 0; pl < nplanes_of(p); pl++)
      if ((pl == 0 || p->loop_filter_level[1 + pl]) && !s->tx_size[pl]) {
        snprintf(msg, 256, "side info: tx_size[%d] is NULL", pl);
        return RVPF_E_INVALID;
      }
  if (cdef_on(p) && (!s->cdef_idx || s->cdef_stride < (p->mi_cols + 15) / 16)) {
    snprintf(msg, 256, "side info: cdef_idx is NULL or cdef_stride too small");
    return RVPF_E_INVALID;
  }
  for (int pl = 0; pl < nplanes_of(p); pl++)
    if (lr_plane_on(p, pl) && (!s || !s->lr_units[pl])) {
      snprintf(msg, 256, "side info: lr_units[%d] is NULL", pl);
      return RVPF_E_INVALID;
    }
  return RVPF_OK;
}

int validate_planes(const rvpf_frame_params *p, const rvpf_planes *pl, const char *name, int use_crop, int margin,
                    char *msg, bool pairs) {
  if (!pl) {
    snprintf(msg, 256, "%s planes are NULL", name);
    return RVPF_E_INVALID;
  }
  if (pl->bytes_per_sample != 2 && pl->bytes_per_sample != 1) {
    snprintf(msg, 256, "%s: bytes_per_sample %d", name, pl->bytes_per_sample);
    return RVPF_E_INVALID;
  }
  if (pl->bytes_per_sample == 1 && p->bit_depth != 8) {
    snprintf(msg, 256, "%s: 8-bit samples above 8 bits per sample", name);
    return RVPF_E_UNSUPPORTED;
  }
  for (int i = 0; i < nplanes_of(p); i++) {
    int w, h;
    if (use_crop == 1)
      crop(p, i, &w, &h);
    else if (use_crop == 2)
      lr_area(p, i, &w, &h);
    else
      area(p, i, &w, &h);
    w += margin;
    h += margin;
    if (!pl->data[i] || pl->stride[i] < pl->width[i] || pl->width[i] < w || pl->height[i] < h) {
      snprintf(msg, 256, "%s plane %d: needs %dx%d samples, has %dx%d (stride %lld)", name, i, w, h, pl->width[i],
               pl->height[i], (long long)pl->stride[i]);
      return RVPF_E_INVALID;
    }
    /* the loop filter moves 4 samples at a time */
    if (margin && (pl->stride[i] % 4 != 0 || ((uintptr_t)pl->data[i] & 7) != 0)) {
      snprintf(msg, 256, "%s plane %d: stride must be a multiple of 4 samples and data 8-byte aligned", name, i);
      return RVPF_E_INVALID;
    }
    if (pairs && (pl->stride[i] % 2 != 0 || ((uintptr_t)pl->data[i] & (2 * pl->bytes_per_sample - 1)) != 0)) {
      snprintf(msg, 256, "%s plane %d: stride must be even and data aligned to two samples", name, i);
      return RVPF_E_INVALID;
    }
    if (((uintptr_t)pl->data[i] & (pl->bytes_per_sample - 1)) != 0) {
      snprintf(msg, 256, "%s plane %d: data not aligned to its sample size", name, i);
      return RVPF_E_INVALID;
    }
  }
  return RVPF_OK;
}

Geom make_geom(const rvpf_frame_params *p) {
  Geom g;
  g.bd = p->bit_depth;
  g.ssx = p->mono_chrome ? 1 : p->ssx;
  g.ssy = p->mono_chrome ? 1 : p->ssy;
  g.nplanes = nplanes_of(p);
  g.width = p->frame_width;
  g.height = p->frame_height;
  g.mi_rows = p->mi_rows;
  g.mi_cols = p->mi_cols;
  return g;
}

/* The geometry loop restoration runs on: the upscaled frame with
 * super-resolution (its width, and a mode-info width covering it). */
Geom make_lr_geom(const rvpf_frame_params *p) {
  Geom g = make_geom(p);
  if (p->use_superres) {
    g.width = p->upscaled_width;
    g.mi_cols = 2 * ((p->upscaled_width + 7) >> 3);
  }
  return g;
}

void to_ptrs(const rvpf_planes *pl, PlanePtr out[3]) {
  for (int i = 0; i < 3; i++) {
    out[i].data = pl ? pl->data[i] : nullptr;
    out[i].stride = pl ? pl->stride[i] : 0;
  }
}

LfHdr make_lf(const rvpf_frame_params *p) {
  LfHdr h;
  memset(&h, 0, sizeof h);
  for (int i = 0; i < 4; i++) h.level[i] = p->loop_filter_level[i];
  h.sharpness = p->loop_filter_sharpness;
  h.delta_enabled = p->loop_filter_delta_enabled;
  for (int i = 0; i < 8; i++) h.ref_deltas[i] = p->loop_filter_ref_deltas[i];
  h.mode_deltas[0] = p->loop_filter_mode_deltas[0];
  h.mode_deltas[1] = p->loop_filter_mode_deltas[1];
  h.delta_lf_multi = p->delta_lf_multi;
  h.seg_enabled = p->segmentation_enabled;
  for (int s = 0; s < 8; s++)
    for (int i = 0; i < 4; i++) {
      if (p->seg_feature_enabled[s][1 + i]) h.seg_lf_enabled[s] |= 1 << i;
      h.seg_lf_data[s][i] = p->seg_feature_data[s][1 + i];
    }
  return h;
}

CdefHdr make_cdef(const rvpf_frame_params *p) {
  CdefHdr h;
  h.damping = p->cdef_damping;
  for (int i = 0; i < 8; i++) {
    h.y_pri[i] = p->cdef_y_pri_strength[i];
    h.y_sec[i] = p->cdef_y_sec_strength[i];
    h.uv_pri[i] = p->cdef_uv_pri_strength[i];
    h.uv_sec[i] = p->cdef_uv_sec_strength[i];
  }
  return h;
}

LrHdr make_lr(const rvpf_frame_params *p, const rvpf_side_info *s) {
  LrHdr h;
  memset(&h, 0, sizeof h);
  for (int pl = 0; pl < nplanes_of(p); pl++) {
    h.type[pl] = p->frame_restoration_type[pl];
    h.unit_size[pl] = p->loop_restoration_size[pl];
    h.unit_rows[pl] = p->lr_unit_rows[pl];
    h.unit_cols[pl] = p->lr_unit_cols[pl];
    h.units[pl] = s ? s->lr_units[pl] : nullptr;
  }
  return h;
}

GrainHdr make_grain(const rvpf_frame_params *p) {
  const rvpf_grain_params *g = &p->grain;
  GrainHdr h;
  memset(&h, 0, sizeof h);
  const int bd = p->bit_depth;
  h.bd = bd;
  h.ssx = p->ssx;
  h.ssy = p->ssy;
  h.nplanes = nplanes_of(p);
  h.w = p->upscaled_width;
  h.h = p->frame_height;
  h.seed = g->grain_seed;
  h.num_y = g->num_y_points;
  h.num_cb = g->num_cb_points;
  h.num_cr = g->num_cr_points;
  h.cfl = g->chroma_scaling_from_luma;
  h.scaling_shift = g->grain_scaling_minus_8 + 8;
  h.ar_lag = g->ar_coeff_lag;
  h.ar_shift = g->ar_coeff_shift_minus_6 + 6;
  h.grain_scale_shift = g->grain_scale_shift;
  for (int i = 0; i < 24; i++) h.ar_y[i] = g->ar_coeffs_y_plus_128[i] - 128;
  for (int i = 0; i < 25; i++) {
    h.ar_cb[i] = g->ar_coeffs_cb_plus_128[i] - 128;
    h.ar_cr[i] = g->ar_coeffs_cr_plus_128[i] - 128;
  }
  h.cb_mult = g->cb_mult;
  h.cb_luma_mult = g->cb_luma_mult;
  h.cb_offset = g->cb_offset;
  h.cr_mult = g->cr_mult;
  h.cr_luma_mult = g->cr_luma_mult;
  h.cr_offset = g->cr_offset;
  h.overlap = g->overlap_flag;
  if (g->clip_to_restricted_range) {
    h.min_value = 16 << (bd - 8);
    h.max_luma = 235 << (bd - 8);
    h.max_chroma = p->matrix_coefficients == 0 ? h.max_luma : 240 << (bd - 8);
  } else {
    h.min_value = 0;
    h.max_luma = h.max_chroma = (256 << (bd - 8)) - 1;
  }
  /* scaling points per plane (grain.rs:140-146) */
  for (int pl = 0; pl < 3; pl++) {
    int n;
    const int32_t *xs, *ys;
    if (pl == 0 || g->chroma_scaling_from_luma) {
      n = g->num_y_points;
      xs = g->point_y_value;
      ys = g->point_y_scaling;
    } else if (pl == 1) {
      n = g->num_cb_points;
      xs = g->point_cb_value;
      ys = g->point_cb_scaling;
    } else {
      n = g->num_cr_points;
      xs = g->point_cr_value;
      ys = g->point_cr_scaling;
    }
    h.npoints[pl] = n;
    for (int i = 0; i < n && i < 16; i++) {
      h.points[pl][0][i] = xs[i];
      h.points[pl][1][i] = ys[i];
    }
  }
  return h;
}

/* grows a scratch buffer of the context to at least `bytes` */
int ensure_buf(rvpf_ctx *ctx, void **buf, size_t *cap, size_t bytes, const char *what) {
  if (*cap >= bytes) return RVPF_OK;
  if (*buf) {
    /* pending work may still use the old buffer */
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return hip_fail(ctx, e, "synchronize");
    (void)hipFree(*buf);
  }
  *buf = nullptr;
  *cap = 0;
  hipError_t e = hipMalloc(buf, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *buf = nullptr;
    char msg[128];
    snprintf(msg, sizeof msg, "cannot allocate the %s", what);
    return fail(ctx, RVPF_E_NOMEM, msg);
  }
  *cap = bytes;
  return RVPF_OK;
}

int scratch_acquire(rvpf_ctx *ctx, ScratchGate *g, hipStream_t st) {
  if (!g->used || g->stream == st) {
    g->used = 1;
    g->stream = st;
    return RVPF_OK;
  }
  hipError_t e = hipSuccess;
  if (!g->multi) {
    /* the first call from another stream: the earlier calls recorded
     * nothing (and their stream may be gone), so wait for the device */
    e = hipDeviceSynchronize();
    g->multi = 1;
  } else if (g->recorded) {
    e = hipStreamWaitEvent(st, g->free, 0);
  }
  g->stream = st;
  return e == hipSuccess ? RVPF_OK : hip_fail(ctx, e, "stream wait (scratch)");
}

int scratch_release(rvpf_ctx *ctx, ScratchGate *g, hipStream_t st) {
  if (!g->multi) return RVPF_OK;
  hipError_t e = hipEventRecord(g->free, st);
  if (e != hipSuccess) return hip_fail(ctx, e, "event record (scratch)");
  g->recorded = 1;
  return RVPF_OK;
}

/* Lays out planes of w[i] x h[i] samples (planes with w[i] == 0 get none)
 * one after another from `base`, rows padded to 64 samples and planes to
 * 256 bytes; returns the bytes needed. With base NULL only counts. */
size_t layout_planes(const int w[3], const int h[3], int bps, void *base, PlanePtr out[3]) {
  size_t bytes = 0;
  for (int i = 0; i < 3; i++) {
    if (w[i] == 0) continue;
    const int64_t stride = (w[i] + 63) & ~63;
    if (base) {
      out[i].data = (char *)base + bytes;
      out[i].stride = stride;
    }
    bytes += (size_t)stride * h[i] * bps;
    bytes = (bytes + 255) & ~(size_t)255;
  }
  return bytes;
}

void grain_offsets_dims(int w, int h, int *stripes, int *blocks) {
  *stripes = ((h + 1) / 2 + 15) / 16 + 1;
  *blocks = ((w + 1) / 2 + 15) / 16 + 1;
}

int ensure_grain_scratch(rvpf_ctx *ctx, int slot, int w, int h) {
  int stripes, blocks;
  grain_offsets_dims(w, h, &stripes, &blocks);
  GrainScratch &g = ctx->grain[slot];
  if (g.tmpl && stripes <= g.max_stripes && blocks <= g.max_blocks) return RVPF_OK;
  if (g.offsets) {
    /* pending work may still read the old offsets */
    hipError_t e = hipDeviceSynchronize();
    if (e != hipSuccess) return hip_fail(ctx, e, "synchronize");
    (void)hipFree(g.offsets);
  }
  g.offsets = nullptr;
  const int ms = stripes > g.max_stripes ? stripes : g.max_stripes;
  const int mb = blocks > g.max_blocks ? blocks : g.max_blocks;
  hipError_t e = hipSuccess;
  if (!g.tmpl) e = hipMalloc(&g.tmpl, GRAIN_TMPL_BYTES);
  if (e == hipSuccess && !g.lut) e = hipMalloc(&g.lut, GRAIN_LUT_BYTES);
  if (e == hipSuccess) e = hipMalloc(&g.offsets, (size_t)ms * mb);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(ctx, RVPF_E_NOMEM, "cannot allocate the film grain scratch");
  }
  g.max_stripes = ms;
  g.max_blocks = mb;
  return RVPF_OK;
}

/* Profiling: a StageMark records a timing event on `st` when built and
 * another when it goes out of scope, bracketing the stage's kernels; it
 * does nothing unless profiling is enabled. */
hipEvent_t take_event(rvpf_ctx *ctx) {
  if (!ctx->free_evs.empty()) {
    hipEvent_t e = ctx->free_evs.back();
    ctx->free_evs.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) {
    (void)hipGetLastError();
    return nullptr;
  }
  return e;
}

struct StageMark {
  rvpf_ctx *ctx;
  int stage;
  hipStream_t st;
  hipEvent_t a;
  StageMark(rvpf_ctx *c, int s, hipStream_t stream) : ctx(c), stage(s), st(stream), a(nullptr) {
    if (!ctx->profiling) return;
    a = take_event(ctx);
    if (a && hipEventRecord(a, st) != hipSuccess) {
      (void)hipGetLastError();
      ctx->free_evs.push_back(a);
      a = nullptr;
    }
  }
  ~StageMark() {
    if (!a) return;
    hipEvent_t b = take_event(ctx);
    if (b && hipEventRecord(b, st) == hipSuccess) {
      ctx->marks.push_back({stage, a, b});
    } else {
      (void)hipGetLastError();
      ctx->free_evs.push_back(a);
      if (b) ctx->free_evs.push_back(b);
    }
  }
};

int check_launch(rvpf_ctx *ctx, int rc, const char *what) {
  if (rc != 0) return hip_fail(ctx, (hipError_t)rc, what);
  return RVPF_OK;
}

/* Builds the film grain templates of `p` into a scratch slot on stream
 * `prep`; the apply (run_output) waits for ev_ready[slot]. */
int grain_prepare(rvpf_ctx *ctx, const rvpf_frame_params *p, hipStream_t prep, GrainTicket *t) {
  const int64_t n = ctx->grain_frames;
  const int slot = (int)(n % GRAIN_SLOTS);
  int rc;
  hipError_t e;
  if ((rc = ensure_grain_scratch(ctx, slot, p->upscaled_width, p->frame_height))) return rc;
  if (n >= GRAIN_SLOTS) {
    /* the slot's last reader is the apply of frame n - GRAIN_SLOTS: wait for
     * the earliest ev_done record after it (one was made within
     * GRAIN_DONE_EVERY frames of it unless a call failed in between) */
    int k = -1;
    for (int i = 0; i < 2; i++)
      if (ctx->done_frame[i] >= n - GRAIN_SLOTS && (k < 0 || ctx->done_frame[i] < ctx->done_frame[k])) k = i;
    if (k >= 0)
      e = hipStreamWaitEvent(prep, ctx->ev_done[k], 0);
    else
      e = hipDeviceSynchronize();
    if (e != hipSuccess) return hip_fail(ctx, e, "stream wait");
  }
  ctx->grain_frames = n + 1;
  {
    StageMark m(ctx, RVPF_STAGE_GRAIN_PREPARE, prep);
    if ((rc = check_launch(ctx, launch_grain_prepare(make_grain(p), ctx->grain[slot], prep), "grain prepare")))
      return rc;
  }
  if ((e = hipEventRecord(ctx->ev_ready[slot], prep)) != hipSuccess) return hip_fail(ctx, e, "event record");
  t->slot = slot;
  t->frame = n;
  t->built_on = prep;
  return RVPF_OK;
}

/* the display planes hold the reference planes' sample size, or bytes
 * from 16-bit planes at 8 bits */
bool display_bps_ok(const rvpf_frame_params *p, int ref_bps, int disp_bps) {
  return disp_bps == ref_bps || (p->bit_depth == 8 && ref_bps == 2 && disp_bps == 1);
}

/* Frame::new (src/frame.rs:161-190): Y, U, V tightly packed, one after another */
void packed_layout(const rvpf_frame_params *p, rvpf_packed_layout *l) {
  memset(l, 0, sizeof *l);
  l->num_planes = nplanes_of(p);
  l->bytes_per_sample = p->bit_depth > 8 ? 2 : 1;
  int64_t off = 0;
  for (int i = 0; i < l->num_planes; i++) {
    crop(p, i, &l->width[i], &l->height[i]);
    l->offset[i] = off;
    off += (int64_t)l->width[i] * l->height[i] * l->bytes_per_sample;
  }
  l->bytes = off;
}

/* display output: crop (+ grain with the templates of `slot`) of `ref` */
int run_output(rvpf_ctx *ctx, const rvpf_frame_params *p, const rvpf_planes *ref, rvpf_planes *disp,
               hipStream_t st, const GrainTicket &t) {
  const int slot = t.slot;
  const Geom g = make_geom(p);
  PlanePtr in[3], out[3];
  to_ptrs(ref, in);
  to_ptrs(disp, out);
  if (grain_on(p)) {
    int rc;
    hipError_t e;
    if (t.built_on != st && (e = hipStreamWaitEvent(st, ctx->ev_ready[slot], 0)) != hipSuccess)
      return hip_fail(ctx, e, "stream wait");
    {
      StageMark m(ctx, RVPF_STAGE_GRAIN_APPLY, st);
      if ((rc = check_launch(ctx,
                             launch_grain_apply(make_grain(p), ctx->grain[slot], in, out, ref->bytes_per_sample, st,
                                                disp->bytes_per_sample),
                             "grain apply")))
        return rc;
    }
    if (t.frame % GRAIN_DONE_EVERY == GRAIN_DONE_EVERY - 1) {
      const int k = (int)((t.frame / GRAIN_DONE_EVERY) & 1);
      if ((e = hipEventRecord(ctx->ev_done[k], st)) != hipSuccess) return hip_fail(ctx, e, "event record");
      ctx->done_frame[k] = t.frame;
    }
    return RVPF_OK;
  }
  int w[3] = {0, 0, 0}, h[3] = {0, 0, 0};
  for (int i = 0; i < g.nplanes; i++) crop(p, i, &w[i], &h[i]);
  StageMark m(ctx, RVPF_STAGE_COPY, st);
  return check_launch(
      ctx, launch_copy_planes(g, in, out, w, h, ref->bytes_per_sample, st, disp->bytes_per_sample), "display copy");
}

/* The chain after deblocking with super-resolution (decode_frame_wrapup,
 * src/decoder/mod.rs:752-758): CDEF of every plane into scratch, the CDEF
 * output upscaled (into ref_out for planes without loop restoration), the
 * deblocked frame upscaled for the restored planes, then loop restoration on
 * the upscaled frames into ref_out. The upscaled planes are written over the
 * whole area loop restoration reads and writes (upscaled width rounded up to
 * 8, mode-info rows), so ref_out holds no stale samples there. */
int superres_tail(rvpf_ctx *ctx, const rvpf_frame_params *p, const rvpf_side_info *s, const PlanePtr rec[3],
                  PlanePtr ref[3], int bps, hipStream_t st) {
  const Geom g = make_geom(p);
  const int np = g.nplanes;
  int rc;
  PlanePtr cdef[3];
  for (int i = 0; i < 3; i++) cdef[i] = rec[i];
  if (cdef_on(p)) {
    int w[3] = {0, 0, 0}, h[3] = {0, 0, 0};
    for (int i = 0; i < np; i++) area(p, i, &w[i], &h[i]);
    if ((rc = ensure_buf(ctx, &ctx->cdef_buf, &ctx->cdef_bytes, layout_planes(w, h, bps, nullptr, nullptr),
                         "CDEF scratch planes")))
      return rc;
    layout_planes(w, h, bps, ctx->cdef_buf, cdef);
    StageMark m(ctx, RVPF_STAGE_CDEF, st);
    if ((rc = check_launch(ctx,
                           launch_cdef(g, make_cdef(p), s->mi, s->mi_stride, s->cdef_idx, s->cdef_stride, 1, rec,
                                       cdef, bps, st),
                           "cdef")))
      return rc;
  }
  /* upscaled planes: the CDEF output, then (distinct only when CDEF ran)
   * the deblocked frame, for the restored planes */
  int uw[3] = {0, 0, 0}, uh[3] = {0, 0, 0}, lw[6] = {0, 0, 0, 0, 0, 0}, lh[6] = {0, 0, 0, 0, 0, 0};
  const bool two = cdef_on(p);
  for (int i = 0; i < np; i++) {
    lr_area(p, i, &uw[i], &uh[i]);
    if (!lr_plane_on(p, i)) continue;
    lw[i] = uw[i];
    lh[i] = uh[i];
    if (two) {
      lw[3 + i] = uw[i];
      lh[3 + i] = uh[i];
    }
  }
  PlanePtr up_cdef[3], up_deb[3];
  for (int i = 0; i < 3; i++) up_cdef[i] = up_deb[i] = ref[i];
  if (lr_on(p)) {
    const size_t b0 = layout_planes(lw, lh, bps, nullptr, nullptr);
    const size_t b1 = layout_planes(lw + 3, lh + 3, bps, nullptr, nullptr);
    if ((rc = ensure_buf(ctx, &ctx->up_buf, &ctx->up_bytes, b0 + b1, "upscaled scratch planes"))) return rc;
    PlanePtr tmp[3];
    layout_planes(lw, lh, bps, ctx->up_buf, tmp);
    for (int i = 0; i < np; i++)
      if (lr_plane_on(p, i)) up_cdef[i] = tmp[i];
    for (int i = 0; i < 3; i++) up_deb[i] = up_cdef[i];
    if (two) layout_planes(lw + 3, lh + 3, bps, (char *)ctx->up_buf + b0, up_deb);
  }
  {
    /* one launch: every plane of the CDEF output, and the deblocked planes
     * the restoration reads */
    UpscaleJob jobs[6];
    int nj = 0;
    for (int i = 0; i < np; i++) jobs[nj++] = UpscaleJob{cdef[i], up_cdef[i], i, uw[i], uh[i]};
    if (two)
      for (int i = 0; i < np; i++)
        if (lw[3 + i]) jobs[nj++] = UpscaleJob{rec[i], up_deb[i], i, lw[3 + i], lh[3 + i]};
    StageMark m(ctx, RVPF_STAGE_UPSCALE, st);
    if ((rc = check_launch(ctx, launch_upscale(g, p->frame_width, p->upscaled_width, jobs, nj, bps, st), "upscale")))
      return rc;
  }
  if (lr_on(p)) {
    StageMark m(ctx, RVPF_STAGE_LR, st);
    if ((rc = check_launch(ctx, launch_lr(make_lr_geom(p), make_lr(p, s), up_deb, up_cdef, ref, bps, st), "lr")))
      return rc;
  }
  return RVPF_OK;
}

}  // namespace host
}  // namespace rvpf

using namespace rvpf::host;

extern "C" {

int rvpf_ctx_create(rvpf_ctx **out) {
  if (!out) return fail(nullptr, RVPF_E_INVALID, "out is NULL");
  *out = nullptr;
  /* value-initialised: every handle and counter starts at zero */
  rvpf_ctx *ctx = new (std::nothrow) rvpf_ctx();
  if (!ctx) return fail(nullptr, RVPF_E_NOMEM, "cannot allocate the context");
  hipError_t e = hipGetDevice(&ctx->device);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&ctx->cus, hipDeviceAttributeMultiprocessorCount, ctx->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking);
  /* ordering events between the device's streams only: no system-scope
   * fence (a cache writeback and invalidation) when they complete */
  const unsigned evf = hipEventDisableTiming | hipEventDisableSystemFence;
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->chain_gate.free, evf);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->search_gate.free, evf);
  for (int i = 0; i < GRAIN_SLOTS; i++) {
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_ready[i], evf);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev_ready[i], ctx->side);
  }
  for (int i = 0; i < 2; i++) {
    ctx->done_frame[i] = -1;
    if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->ev_done[i], evf);
  }
  if (e != hipSuccess) {
    int rc = hip_fail(nullptr, e, "rvpf_ctx_create");
    (void)hipGetLastError();
    rvpf_ctx_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return RVPF_OK;
}

void rvpf_ctx_destroy(rvpf_ctx *ctx) {
  if (!ctx) return;
  if (ctx->side) (void)hipStreamSynchronize(ctx->side);
  if (ctx->deb_buf) (void)hipFree(ctx->deb_buf);
  if (ctx->cdef_buf) (void)hipFree(ctx->cdef_buf);
  if (ctx->up_buf) (void)hipFree(ctx->up_buf);
  if (ctx->search_buf) (void)hipFree(ctx->search_buf);
  (void)hipDeviceSynchronize();
  for (auto &m : ctx->marks) {
    (void)hipEventDestroy(m.a);
    (void)hipEventDestroy(m.b);
  }
  for (auto e : ctx->free_evs) (void)hipEventDestroy(e);
  for (int i = 0; i < GRAIN_SLOTS; i++) {
    if (ctx->grain[i].tmpl) (void)hipFree(ctx->grain[i].tmpl);
    if (ctx->grain[i].lut) (void)hipFree(ctx->grain[i].lut);
    if (ctx->grain[i].offsets) (void)hipFree(ctx->grain[i].offsets);
    if (ctx->ev_ready[i]) (void)hipEventDestroy(ctx->ev_ready[i]);
  }
  for (int i = 0; i < 2; i++)
    if (ctx->ev_done[i]) (void)hipEventDestroy(ctx->ev_done[i]);
  if (ctx->chain_gate.free) (void)hipEventDestroy(ctx->chain_gate.free);
  if (ctx->search_gate.free) (void)hipEventDestroy(ctx->search_gate.free);
  if (ctx->side) (void)hipStreamDestroy(ctx->side);
  delete ctx;
}

int rvpf_ctx_sync(rvpf_ctx *ctx, void *stream) {
  if (!ctx) return fail(nullptr, RVPF_E_INVALID, "ctx is NULL");
  hipError_t e = hipStreamSynchronize((hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "rvpf_ctx_sync");
  return RVPF_OK;
}

const char *rvpf_ctx_last_error(const rvpf_ctx *ctx) { return ctx ? ctx->err : g_err; }

int rvpf_ctx_set_profiling(rvpf_ctx *ctx, int enable) {
  if (!ctx) return fail(nullptr, RVPF_E_INVALID, "ctx is NULL");
  ctx->profiling = enable != 0;
  return RVPF_OK;
}

int rvpf_ctx_read_profile(rvpf_ctx *ctx, rvpf_profile *out) {
  if (!ctx || !out) return fail(ctx, RVPF_E_INVALID, "ctx or out is NULL");
  memset(out, 0, sizeof *out);
  int rc = RVPF_OK;
  for (auto &m : ctx->marks) {
    float ms = 0;
    hipError_t e = hipEventSynchronize(m.b);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, m.a, m.b);
    if (e != hipSuccess) {
      rc = hip_fail(ctx, e, "profile read");
    } else if (m.stage >= 0 && m.stage < RVPF_STAGE_COUNT) {
      out->ms[m.stage] += ms;
      out->launches[m.stage] += 1;
    }
    ctx->free_evs.push_back(m.a);
    ctx->free_evs.push_back(m.b);
  }
  ctx->marks.clear();
  return rc;
}

int rvpf_validate(const rvpf_frame_params *p, const rvpf_side_info *s) {
  int rc = validate_params(p, g_err);
  if (rc) return rc;
  return validate_side(p, s, g_err);
}

int rvpf_deblock(rvpf_ctx *ctx, const rvpf_frame_params *p, const rvpf_side_info *s, rvpf_planes *frame,
                 void *stream) {
  if (!ctx) return fail(nullptr, RVPF_E_INVALID, "ctx is NULL");
  int rc;
  if ((rc = validate_params(p, ctx->err)) || (rc = validate_side(p, s, ctx->err)) ||
      (rc = validate_planes(p, frame, "frame", 0, 8, ctx->err)))
    return rc;
  if (!lf_on(p)) return RVPF_OK;
  const Geom g = make_geom(p);
  PlanePtr pl[3];
  to_ptrs(frame, pl);
  LfSide side;
  side.mi = s->mi;
  side.mi_stride = s->mi_stride;
  for (int i = 0; i < 3; i++) side.tx[i] = s->tx_size[i];
  StageMark m(ctx, RVPF_STAGE_DEBLOCK, (hipStream_t)stream);
  return check_launch(ctx, launch_deblock(g, make_lf(p), side, pl, frame->bytes_per_sample, (hipStream_t)stream),
                      "deblock");
}

int rvpf_cdef(rvpf_ctx *ctx, const rvpf_frame_params *p, const rvpf_side_info *s, const rvpf_planes *deblocked,
              rvpf_planes *out, void *stream) {
  if (!ctx) return fail(nullptr, RVPF_E_INVALID, "ctx is NULL");
  int rc;
  if ((rc = validate_params(p, ctx->err)) || (rc = validate_side(p, s, ctx->err)) ||
      (rc = validate_planes(p, deblocked, "deblocked", 0, 0, ctx->err)) ||
      (rc = validate_planes(p, out, "out", 0, 0, ctx->err, true)))
    return rc;
  if (deblocked->bytes_per_sample != out->bytes_per_sample)
    return fail(ctx, RVPF_E_INVALID, "deblocked and out sample sizes differ");
  const Geom g = make_geom(p);
  PlanePtr in[3], o[3];
  to_ptrs(deblocked, in);
  to_ptrs(out, o);
  if (!cdef_on(p)) {
    int w[3], h[3];
    for (int i = 0; i < g.nplanes; i++) area(p, i, &w[i], &h[i]);
    StageMark m(ctx, RVPF_STAGE_COPY, (hipStream_t)stream);
    return check_launch(ctx, launch_copy_planes(g, in, o, w, h, out->bytes_per_sample, (hipStream_t)stream),
                        "cdef copy");
  }
  StageMark m(ctx, RVPF_STAGE_CDEF, (hipStream_t)stream);
  return check_launch(ctx,
                      launch_cdef(g, make_cdef(p), s->mi, s->mi_stride, s->cdef_idx, s->cdef_stride, 1, in, o,
                                  out->bytes_per_sample, (hipStream_t)stream),
                      "cdef");
}

int rvpf_upscale(rvpf_ctx *ctx, const rvpf_frame_params *p, const rvpf_planes *in, rvpf_planes *out, void *stream) {
  if (!ctx) return fail(nullptr, RVPF_E_INVALID, "ctx is NULL");
  int rc;
  if ((rc = validate_params(p, ctx->err))) return rc;
  if (!p->use_superres) return fail(ctx, RVPF_E_INVALID, "use_superres is not set");
  if ((rc = validate_planes(p, in, "in", 0, 0, ctx->err)) || (rc = validate_planes(p, out, "out", 1, 0, ctx->err)))
    return rc;
  if (in->bytes_per_sample != out->bytes_per_sample) return fail(ctx, RVPF_E_INVALID, "in and out sample sizes differ");
  PlanePtr i_[3], o[3];
  to_ptrs(in, i_);
  to_ptrs(out, o);
  UpscaleJob jobs[3];
  const int np = nplanes_of(p);
  for (int i = 0; i < np; i++) {
    jobs[i] = UpscaleJob{i_[i], o[i], i, 0, 0};
    crop(p, i, &jobs[i].cols, &jobs[i].rows);
  }
  hipStream_t st = (hipStream_t)stream;
  StageMark m(ctx, RVPF_STAGE_UPSCALE, st);
  return check_launch(
      ctx, launch_upscale(make_geom(p), p->frame_width, p->upscaled_width, jobs, np, out->bytes_per_sample, st),
      "upscale");
}

int rvpf_lr(rvpf_ctx *ctx, const rvpf_frame_params *p, const rvpf_side_info *s, const rvpf_planes *deblocked,
            const rvpf_planes *cdef, rvpf_planes *out, void *stream) {
  if (!ctx) return fail(nullptr, RVPF_E_INVALID, "ctx is NULL");
  int rc;
  if ((rc = validate_params(p, ctx->err)) || (rc = validate_side(p, s, ctx->err)) ||
      (rc = validate_planes(p, deblocked, "deblocked", 2, 0, ctx->err)) ||
      (rc = validate_planes(p, cdef, "cdef", 2, 0, ctx->err)) || (rc = validate_planes(p, out, "out", 2, 0, ctx->err)))
    return rc;
  if (deblocked->bytes_per_sample != out->bytes_per_sample || cdef->bytes_per_sample != out->bytes_per_sample)
    return fail(ctx, RVPF_E_INVALID, "plane sample sizes differ");
  const Geom g = make_lr_geom(p);
  PlanePtr d[3], c[3], o[3];
  to_ptrs(deblocked, d);
  to_ptrs(cdef, c);
  to_ptrs(out, o);
  int w[3] = {0, 0, 0}, h[3] = {0, 0, 0};
  for (int i = 0; i < g.nplanes; i++)
    if (!lr_plane_on(p, i)) lr_area(p, i, &w[i], &h[i]);
  hipStream_t st = (hipStream_t)stream;
  {
    StageMark m(ctx, RVPF_STAGE_COPY, st);
    if ((rc = check_launch(ctx, launch_copy_planes(g, c, o, w, h, out->bytes_per_sample, st), "lr copy"))) return rc;
  }
  StageMark m(ctx, RVPF_STAGE_LR, st);
  return check_launch(ctx, launch_lr(g, make_lr(p, s), d, c, o, out->bytes_per_sample, st), "lr");
}

int rvpf_grain(rvpf_ctx *ctx, const rvpf_frame_params *p, const rvpf_planes *ref, rvpf_planes *display,
               void *stream) {
  if (!ctx) return fail(nullptr, RVPF_E_INVALID, "ctx is NULL");
  int rc;
  if ((rc = validate_params(p, ctx->err)) || (rc = validate_planes(p, ref, "ref", 1, 0, ctx->err)) ||
      (rc = validate_planes(p, display, "display", 1, 0, ctx->err)))
    return rc;
  if (!display_bps_ok(p, ref->bytes_per_sample, display->bytes_per_sample))
    return fail(ctx, RVPF_E_INVALID, "ref and display sample sizes differ (bytes from 16-bit samples only at 8 bits)");
  hipStream_t st = (hipStream_t)stream;
  GrainTicket t = {0, 0, st};
  if (grain_on(p) && (rc = grain_prepare(ctx, p, st, &t))) return rc;
  return run_output(ctx, p, ref, display, st, t);
}

int rvpf_packed_layout_of(const rvpf_frame_params *p, rvpf_packed_layout *out) {
  int rc = validate_params(p, g_err);
  if (rc) return rc;
  if (!out) return fail(nullptr, RVPF_E_INVALID, "out is NULL");
  packed_layout(p, out);
  return RVPF_OK;
}

int rvpf_output_packed(rvpf_ctx *ctx, const rvpf_frame_params *p, const rvpf_planes *ref, void *packed,
                       int64_t bytes, void *stream) {
  if (!ctx) return fail(nullptr, RVPF_E_INVALID, "ctx is NULL");
  int rc;
  if ((rc = validate_params(p, ctx->err)) || (rc = validate_planes(p, ref, "ref", 1, 0, ctx->err))) return rc;
  rvpf_packed_layout l;
  packed_layout(p, &l);
  if (!packed || bytes < l.bytes) {
    snprintf(ctx->err, sizeof ctx->err, "packed picture: needs %lld bytes, has %lld", (long long)l.bytes,
             (long long)bytes);
    return RVPF_E_INVALID;
  }
  if (((uintptr_t)packed & (l.bytes_per_sample - 1)) != 0)
    return fail(ctx, RVPF_E_INVALID, "packed picture not aligned to its sample size");
  if (!display_bps_ok(p, ref->bytes_per_sample, l.bytes_per_sample))
    return fail(ctx, RVPF_E_INVALID, "ref and packed sample sizes differ");
  rvpf_planes disp;
  memset(&disp, 0, sizeof disp);
  disp.bytes_per_sample = l.bytes_per_sample;
  for (int i = 0; i < l.num_planes; i++) {
    disp.data[i] = (char *)packed + l.offset[i];
    disp.stride[i] = l.width[i];
    disp.width[i] = l.width[i];
    disp.height[i] = l.height[i];
  }
  hipStream_t st = (hipStream_t)stream;
  GrainTicket t = {0, 0, st};
  if (grain_on(p) && (rc = grain_prepare(ctx, p, ctx->side, &t))) return rc;
  return run_output(ctx, p, ref, &disp, st, t);
}

int rvpf_postfilter_frame(rvpf_ctx *ctx, const rvpf_frame_params *p, const rvpf_side_info *s, rvpf_planes *recon,
                          rvpf_planes *ref_out, rvpf_planes *display, void *stream) {
  if (!ctx) return fail(nullptr, RVPF_E_INVALID, "ctx is NULL");
  int rc;
  if ((rc = validate_params(p, ctx->err)) || (rc = validate_side(p, s, ctx->err)) ||
      (rc = validate_planes(p, recon, "recon", 0, lf_on(p) ? 8 : 0, ctx->err)) ||
      (rc = validate_planes(p, ref_out, "ref_out", 2, 0, ctx->err, true)))
    return rc;
  if (display && (rc = validate_planes(p, display, "display", 1, 0, ctx->err))) return rc;
  const int bps = recon->bytes_per_sample;
  if (ref_out->bytes_per_sample != bps || (display && !display_bps_ok(p, bps, display->bytes_per_sample)))
    return fail(ctx, RVPF_E_INVALID, "plane sample sizes differ");
  hipStream_t st = (hipStream_t)stream;
  const Geom g = make_geom(p);
  const bool show_grain = display && grain_on(p);
  /* film grain templates on the side stream, overlapping the filters */
  GrainTicket t = {0, 0, st};
  if (show_grain && (rc = grain_prepare(ctx, p, ctx->side, &t))) return rc;
  PlanePtr rec[3], ref[3], deb[3];
  to_ptrs(recon, rec);
  to_ptrs(ref_out, ref);
  const bool any_lr = lr_on(p);
  /* context scratch this call writes (the deblocked frame, the CDEF output
   * of the restored planes, the upscaled frames) waits for the call that
   * read it last */
  const bool gate = lf_on(p) || p->use_superres || (cdef_on(p) && any_lr);
  if (gate && (rc = scratch_acquire(ctx, &ctx->chain_gate, st))) return rc;
  /* 1. deblock, both passes in one launch, from recon (not modified) into
   * scratch planes */
  for (int i = 0; i < 3; i++) deb[i] = rec[i];
  if (lf_on(p)) {
    LfSide side;
    side.mi = s->mi;
    side.mi_stride = s->mi_stride;
    for (int i = 0; i < 3; i++) side.tx[i] = s->tx_size[i];
    const LfHdr lh = make_lf(p);
    int w[3] = {0, 0, 0}, h[3] = {0, 0, 0};
    for (int i = 0; i < g.nplanes; i++)
      if (i == 0 || lh.level[1 + i] != 0) area(p, i, &w[i], &h[i]);
    if ((rc = ensure_buf(ctx, &ctx->deb_buf, &ctx->deb_bytes, layout_planes(w, h, bps, nullptr, nullptr),
                         "deblocked scratch planes")))
      return rc;
    PlanePtr tmp[3];
    layout_planes(w, h, bps, ctx->deb_buf, tmp);
    for (int i = 0; i < g.nplanes; i++)
      if (w[i]) deb[i] = tmp[i];
    StageMark m(ctx, RVPF_STAGE_DEBLOCK, st);
    if ((rc = check_launch(ctx, launch_deblock_fused(g, lh, side, rec, deb, bps, st), "deblock"))) return rc;
  }
  if (p->use_superres) {
    if ((rc = superres_tail(ctx, p, s, deb, ref, bps, st))) return rc;
  } else {
    /* 2. CDEF: planes restored next go to scratch, the others straight to ref_out */
    PlanePtr cdef[3];
    if (cdef_on(p)) {
      for (int i = 0; i < 3; i++) cdef[i] = ref[i];
      if (any_lr) {
        int w[3] = {0, 0, 0}, h[3] = {0, 0, 0};
        for (int i = 0; i < g.nplanes; i++)
          if (lr_plane_on(p, i)) area(p, i, &w[i], &h[i]);
        if ((rc = ensure_buf(ctx, &ctx->cdef_buf, &ctx->cdef_bytes, layout_planes(w, h, bps, nullptr, nullptr),
                             "CDEF scratch planes")))
          return rc;
        layout_planes(w, h, bps, ctx->cdef_buf, cdef);
      }
      StageMark m(ctx, RVPF_STAGE_CDEF, st);
      if ((rc = check_launch(ctx,
                             launch_cdef(g, make_cdef(p), s->mi, s->mi_stride, s->cdef_idx, s->cdef_stride, 1, deb,
                                         cdef, bps, st),
                             "cdef")))
        return rc;
    } else {
      /* CDEF is the identity: planes not restored are copied to ref_out */
      int w[3] = {0, 0, 0}, h[3] = {0, 0, 0};
      for (int i = 0; i < g.nplanes; i++) {
        cdef[i] = deb[i];
        if (!lr_plane_on(p, i)) area(p, i, &w[i], &h[i]);
      }
      StageMark m(ctx, RVPF_STAGE_COPY, st);
      if ((rc = check_launch(ctx, launch_copy_planes(g, deb, ref, w, h, bps, st), "copy"))) return rc;
    }
    /* 3. loop restoration of the restored planes into ref_out */
    if (any_lr) {
      StageMark m(ctx, RVPF_STAGE_LR, st);
      if ((rc = check_launch(ctx, launch_lr(g, make_lr(p, s), deb, cdef, ref, bps, st), "lr"))) return rc;
    }
  }
  if (gate && (rc = scratch_release(ctx, &ctx->chain_gate, st))) return rc;
  /* 4. display */
  if (!display) return RVPF_OK;
  return run_output(ctx, p, ref_out, display, st, t);
}

}  // extern "C"
