/*
 * synth.c — seeded synthetic workloads for the post-filter chain (see synth.h).
 *
 * The reference produces its side info with its own encoder; that encoder
 * cannot express 128x128 superblocks, 12 bits, 4:4:4, film grain or mixed
 * restoration at 1080p, so this generator synthesises side info the way a
 * conforming bitstream lays it out: a partition tree (forced splits at the
 * frame edge, as partition syntax 5.11.4 forces them), one transform size
 * per block tiled from the block origin, chroma transform sizes from the
 * block owning each chroma 4x4 (get_plane_residual_size), cdef_idx = -1 on
 * 64x64s with no coded residual, restoration units of every type.
 */
#include "synth.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "av1_tables.h"

/* ------------------------------------------------------------------ rng */
typedef struct {
  uint64_t s;
} rng_t;

static uint64_t rng_u64(rng_t *r) {
  uint64_t z = (r->s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double rng_unit(rng_t *r) { return (double)(rng_u64(r) >> 11) * (1.0 / 9007199254740992.0); }
static int rng_int(rng_t *r, int n) { return n <= 1 ? 0 : (int)(rng_u64(r) % (uint64_t)n); }
static int rng_range(rng_t *r, int lo, int hi) { return lo + rng_int(r, hi - lo + 1); }
static double rng_normal(rng_t *r) {
  double u1 = rng_unit(r), u2 = rng_unit(r);
  if (u1 < 1e-300) u1 = 1e-300;
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}
static rng_t rng_make(uint64_t seed, uint64_t a, uint64_t b) {
  rng_t r = {seed ^ (a * 0xD1B54A32D192ED03ull) ^ (b * 0x8CB92BA72F3D8DD7ull)};
  rng_u64(&r);
  return r;
}

static inline int imin(int a, int b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

static int count_units(int unit, int size) {
  int n = (size + (unit >> 1)) / unit;
  return n < 1 ? 1 : n;
}

int synth_geometry(const synth_cfg *c, synth_geom *g) {
  memset(g, 0, sizeof *g);
  if (c->width < 1 || c->height < 1 || c->width > 16384 || c->height > 16384) return -1;
  if (c->bit_depth != 8 && c->bit_depth != 10 && c->bit_depth != 12) return -1;
  if (c->sb_size != 64 && c->sb_size != 128) return -1;
  if (c->superres_denom && (c->superres_denom < 9 || c->superres_denom > 16)) return -1;
  /* 7.21 (superres_params): the coded width from the upscaled one */
  g->coded_width = c->superres_denom ? (c->width * 8 + c->superres_denom / 2) / c->superres_denom : c->width;
  g->mi_cols = 2 * ((g->coded_width + 7) >> 3);
  g->mi_rows = 2 * ((c->height + 7) >> 3);
  g->mi_stride = g->mi_cols + 32;
  g->mi_alloc_rows = g->mi_rows + 32;
  g->cdef_stride = (g->mi_cols >> 4) + 2;
  g->cdef_rows = (g->mi_rows >> 4) + 2;
  g->num_planes = c->mono ? 1 : 3;
  const int aw = (g->mi_cols * 4 + 127) & ~127, ah = (g->mi_rows * 4 + 127) & ~127;
  const int unit = c->lr_unit_size >= 64 ? c->lr_unit_size : 64; /* luma units: 64, 128, 256 */
  for (int p = 0; p < g->num_planes; p++) {
    const int sx = p ? c->ssx : 0, sy = p ? c->ssy : 0;
    g->plane_w[p] = (aw >> sx) + 32;
    g->plane_h[p] = (ah >> sy) + 32;
    g->crop_w[p] = (c->width + sx) >> sx;
    {
      const int awu = (2 * ((c->width + 7) >> 3) * 4 + 127) & ~127;
      g->ref_w[p] = c->superres_denom ? (awu >> sx) + 32 : g->plane_w[p];
    }
    g->crop_h[p] = (c->height + sy) >> sy;
    const int us = p == 0 ? unit : ((c->ssx && c->ssy && unit > 32) ? unit >> 1 : unit);
    g->lr_unit_rows[p] = count_units(us, (c->height + sy) >> sy);
    g->lr_unit_cols[p] = count_units(us, (c->width + sx) >> sx);
  }
  return 0;
}

void synth_struct_sizes(int32_t out[11]) {
  out[0] = (int32_t)sizeof(rvpf_frame_params);
  out[1] = (int32_t)sizeof(rvpf_side_info);
  out[2] = (int32_t)sizeof(rvpf_mi);
  out[3] = (int32_t)sizeof(rvpf_lr_unit);
  out[4] = (int32_t)sizeof(rvpf_grain_params);
  out[5] = (int32_t)sizeof(rvpf_planes);
  out[6] = (int32_t)sizeof(rvpf_batch_frame);
  out[7] = (int32_t)sizeof(rvpf_tile_config);
  out[8] = (int32_t)sizeof(rvpf_lrd_item);
  out[9] = (int32_t)sizeof(rvpf_packed_layout);
  out[10] = (int32_t)sizeof(rvpf_cdef_search_result);
}

/* ------------------------------------------------------------- source */
typedef struct {
  double amp, kx, ky, phase;
} sinus_t;
typedef struct {
  double x, y, w, h, vx, vy, level;
} rect_t;

static void gen_source(const synth_cfg *c, int plane, uint16_t *dst, int pw, int ph) {
  const int sx = plane ? c->ssx : 0, sy = plane ? c->ssy : 0;
  /* the coded picture (narrower with super-resolution) */
  const int cw = c->superres_denom ? (c->width * 8 + c->superres_denom / 2) / c->superres_denom : c->width;
  rng_t r = rng_make(c->seed, 1000 + (uint64_t)plane, 0);
  sinus_t s[4];
  for (int k = 0; k < 4; k++) {
    const double period = 8.0 + 56.0 * rng_unit(&r);
    const double th = 6.283185307179586 * rng_unit(&r);
    s[k].amp = 8.0 + 28.0 * rng_unit(&r);
    s[k].kx = 6.283185307179586 * cos(th) / period;
    s[k].ky = 6.283185307179586 * sin(th) / period;
    s[k].phase = 6.283185307179586 * rng_unit(&r);
  }
  rect_t rc[3];
  for (int k = 0; k < 3; k++) {
    rc[k].w = (cw >> sx) * (0.1 + 0.25 * rng_unit(&r));
    rc[k].h = (c->height >> sy) * (0.1 + 0.25 * rng_unit(&r));
    rc[k].x = (cw >> sx) * rng_unit(&r);
    rc[k].y = (c->height >> sy) * rng_unit(&r);
    rc[k].vx = (rng_unit(&r) * 8.0 - 4.0) / (1 << sx);
    rc[k].vy = (rng_unit(&r) * 8.0 - 4.0) / (1 << sy);
    rc[k].level = 255.0 * rng_unit(&r);
  }
  const double pan = c->pan_px * c->frame_index / (1 << sx);
  const double scale = (double)(1 << (c->bit_depth - 8));
  const int maxv = (1 << c->bit_depth) - 1;
  /* separable sinusoids: sin(ax + by + p) = sin(ax) cos(by + p) + cos(ax) sin(by + p) */
  double *cx = (double *)malloc(sizeof(double) * 8 * (size_t)pw);
  double *cy = (double *)malloc(sizeof(double) * 8 * (size_t)ph);
  for (int k = 0; k < 4; k++) {
    for (int x = 0; x < pw; x++) {
      cx[(2 * k) * pw + x] = sin(s[k].kx * (x + pan));
      cx[(2 * k + 1) * pw + x] = cos(s[k].kx * (x + pan));
    }
    for (int y = 0; y < ph; y++) {
      cy[(2 * k) * ph + y] = s[k].amp * cos(s[k].ky * y + s[k].phase);
      cy[(2 * k + 1) * ph + y] = s[k].amp * sin(s[k].ky * y + s[k].phase);
    }
  }
  rng_t nr = rng_make(c->seed, 2000 + (uint64_t)plane, (uint64_t)c->frame_index);
  for (int y = 0; y < ph; y++) {
    for (int x = 0; x < pw; x++) {
      double v = 128.0;
      for (int k = 0; k < 4; k++)
        v += cx[(2 * k) * pw + x] * cy[(2 * k) * ph + y] + cx[(2 * k + 1) * pw + x] * cy[(2 * k + 1) * ph + y];
      for (int k = 0; k < 3; k++) {
        const double rx = rc[k].x + rc[k].vx * c->frame_index, ry = rc[k].y + rc[k].vy * c->frame_index;
        if (x >= rx && x < rx + rc[k].w && y >= ry && y < ry + rc[k].h) v = rc[k].level;
      }
      v += 6.0 * rng_normal(&nr);
      if (v < 0) v = 0;
      if (v > 255) v = 255;
      dst[(int64_t)y * pw + x] = (uint16_t)clampi((int)lrint(v * scale), 0, maxv);
    }
  }
  free(cx);
  free(cy);
}

/* ----------------------------------------------------------- partitions */
static int block_enum(int w4, int h4) {
  for (int b = 0; b < 22; b++)
    if (AV1_NUM_4X4_BLOCKS_WIDE[b] == w4 && AV1_NUM_4X4_BLOCKS_HIGH[b] == h4) return b;
  return -1;
}
static int tx_enum(int w, int h) {
  for (int t = 0; t < 19; t++)
    if (AV1_TX_WIDTH[t] == w && AV1_TX_HEIGHT[t] == h) return t;
  return -1;
}

typedef struct {
  const synth_cfg *c;
  const synth_geom *g;
  rvpf_mi *mi;
  uint8_t *tx[3];
  rng_t r;
  int key;
  /* coded blocks */
  int nblk, cap;
  int (*blk)[6]; /* r, c, w4, h4, tx, skip */
} part_ctx;

static int chroma_ok(const synth_cfg *c, int w4, int h4) {
  const int b = block_enum(w4, h4);
  if (b < 0) return 0;
  if (c->mono) return 1;
  return AV1_SUBSAMPLED_SIZE[b][c->ssx][c->ssy] != 22;
}

static void code_block(part_ctx *pc, int r, int c, int w4, int h4) {
  const synth_cfg *cf = pc->c;
  const synth_geom *g = pc->g;
  rng_t *rr = &pc->r;
  const int b = block_enum(w4, h4);
  /* transform size: the largest for the block (64 cap), split 0..2 times */
  int tw = imin(w4 * 4, 64), th = imin(h4 * 4, 64);
  const double q = cf->qindex / 255.0;
  int splits = rng_unit(rr) < 0.55 - 0.35 * q ? 1 : 0;
  if (splits && rng_unit(rr) < 0.3) splits = 2;
  for (int k = 0; k < splits; k++) {
    if (tw == 4 && th == 4) break;
    if (tw > th) tw /= 2;
    else if (th > tw) th /= 2;
    else {
      tw /= 2;
      th /= 2;
    }
  }
  const int tx = tx_enum(tw, th);
  const int skip = rng_unit(rr) < 0.2 + 0.45 * q;
  const int intra = pc->key || rng_unit(rr) < 0.3;
  const int seg = cf->lf_extras ? rng_int(rr, 8) : 0;
  int8_t dlf[4] = {0, 0, 0, 0};
  if (cf->lf_extras)
    for (int i = 0; i < 4; i++) dlf[i] = (int8_t)rng_range(rr, -12, 12);
  const uint8_t ymode = intra ? (uint8_t)rng_int(rr, 13) : (uint8_t)rng_range(rr, 14, 24);
  const int8_t ref0 = intra ? 0 : (int8_t)rng_range(rr, 1, 7);
  for (int i = 0; i < h4; i++)
    for (int j = 0; j < w4; j++) {
      const int rr_ = r + i, cc = c + j;
      if (rr_ >= g->mi_alloc_rows || cc >= g->mi_stride) continue;
      rvpf_mi *m = &pc->mi[(int64_t)rr_ * g->mi_stride + cc];
      m->mi_size = (uint8_t)b;
      m->y_mode = ymode;
      m->ref_frame0 = ref0;
      m->seg_skip = (uint8_t)(skip | (seg << 1));
      memcpy(m->delta_lf, dlf, 4);
      pc->tx[0][(int64_t)rr_ * g->mi_stride + cc] = (uint8_t)tx;
    }
  if (pc->nblk == pc->cap) {
    pc->cap = pc->cap ? pc->cap * 2 : 4096;
    pc->blk = realloc(pc->blk, sizeof(*pc->blk) * (size_t)pc->cap);
  }
  int *e = pc->blk[pc->nblk++];
  e[0] = r;
  e[1] = c;
  e[2] = w4;
  e[3] = h4;
  e[4] = tx;
  e[5] = skip;
}

/* partition(): NONE / HORZ / VERT / SPLIT / HORZ_4 / VERT_4, forced splits
 * where half the block lies outside the frame. */
static void partition(part_ctx *pc, int r, int c, int n4) {
  const synth_geom *g = pc->g;
  if (r >= g->mi_rows || c >= g->mi_cols) return;
  const int half = n4 >> 1;
  const int has_rows = (r + half) < g->mi_rows, has_cols = (c + half) < g->mi_cols;
  if (n4 == 1) {
    code_block(pc, r, c, 1, 1);
    return;
  }
  if (!has_rows || !has_cols) {
    if (n4 == 2 && !chroma_ok(pc->c, 1, 1)) {
      code_block(pc, r, c, 2, 2);
      return;
    }
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) partition(pc, r + i * half, c + j * half, half);
    return;
  }
  const double q = pc->c->qindex / 255.0;
  const double p_split = n4 >= 16 ? 0.85 - 0.3 * q : (n4 >= 8 ? 0.6 - 0.3 * q : (n4 >= 4 ? 0.45 - 0.3 * q : 0.3 - 0.25 * q));
  const double u = rng_unit(&pc->r);
  if (u < p_split) {
    for (int i = 0; i < 2; i++)
      for (int j = 0; j < 2; j++) partition(pc, r + i * half, c + j * half, half);
    return;
  }
  const int kind = rng_int(&pc->r, 6); /* 0 none, 1 horz, 2 vert, 3 horz4, 4 vert4, 5 none */
  if (kind == 1 && chroma_ok(pc->c, n4, half)) {
    code_block(pc, r, c, n4, half);
    if (r + half < g->mi_rows) code_block(pc, r + half, c, n4, half);
  } else if (kind == 2 && chroma_ok(pc->c, half, n4)) {
    code_block(pc, r, c, half, n4);
    if (c + half < g->mi_cols) code_block(pc, r, c + half, half, n4);
  } else if (kind == 3 && n4 >= 4 && n4 <= 16 && chroma_ok(pc->c, n4, n4 / 4)) {
    for (int k = 0; k < 4; k++)
      if (r + k * (n4 / 4) < g->mi_rows) code_block(pc, r + k * (n4 / 4), c, n4, n4 / 4);
  } else if (kind == 4 && n4 >= 4 && n4 <= 16 && chroma_ok(pc->c, n4 / 4, n4)) {
    for (int k = 0; k < 4; k++)
      if (c + k * (n4 / 4) < g->mi_cols) code_block(pc, r, c + k * (n4 / 4), n4 / 4, n4);
  } else {
    code_block(pc, r, c, n4, n4);
  }
}

/* One transform block of a plane: mean (quantised) + quantised residual;
 * a skipped block keeps only its quantised mean. */
static void quantise_rect(const uint16_t *src, uint16_t *rec, int64_t stride, int x0, int y0, int w,
                          int h, int pw, int ph, double qs, int skip, int maxv) {
  if (x0 >= pw || y0 >= ph) return;
  w = imin(w, pw - x0);
  h = imin(h, ph - y0);
  double mean = 0;
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) mean += src[(int64_t)(y0 + y) * stride + x0 + x];
  mean /= (double)(w * h);
  const double qdc = skip ? 2.0 * qs : qs * 0.5;
  const double pred = floor(mean / qdc + 0.5) * qdc;
  for (int y = 0; y < h; y++)
    for (int x = 0; x < w; x++) {
      const int64_t i = (int64_t)(y0 + y) * stride + x0 + x;
      double v = pred;
      if (!skip) v += floor((src[i] - mean) / qs + 0.5) * qs;
      rec[i] = (uint16_t)clampi((int)lrint(v), 0, maxv);
    }
}

static int uv_tx_size(int cw, int ch) {
  return tx_enum(imin(cw, 32), imin(ch, 32));
}

int synth_frame(const synth_cfg *c, rvpf_frame_params *p, rvpf_mi *mi, uint8_t *tx0, uint8_t *tx1,
                uint8_t *tx2, int8_t *cdef_idx, rvpf_lr_unit *lr0, rvpf_lr_unit *lr1,
                rvpf_lr_unit *lr2, uint16_t *src0, uint16_t *src1, uint16_t *src2, uint16_t *rec0,
                uint16_t *rec1, uint16_t *rec2) {
  synth_geom g;
  if (synth_geometry(c, &g)) return -1;
  const int np = g.num_planes;
  const int maxv = (1 << c->bit_depth) - 1;
  rng_t hr = rng_make(c->seed, 3000, (uint64_t)c->frame_index);
  const int key = c->frame_index == 0;

  /* ---------------- frame params ---------------- */
  rvpf_frame_params P;
  memset(&P, 0, sizeof P);
  P.bit_depth = c->bit_depth;
  P.ssx = c->mono ? 1 : c->ssx;
  P.ssy = c->mono ? 1 : c->ssy;
  P.mono_chrome = c->mono;
  P.matrix_coefficients = c->mc_identity ? 0 : 1;
  P.enable_cdef = c->cdef;
  P.film_grain_params_present = c->grain;
  P.frame_width = g.coded_width;
  P.frame_height = c->height;
  P.upscaled_width = c->width;
  P.use_superres = c->superres_denom != 0;
  P.mi_rows = g.mi_rows;
  P.mi_cols = g.mi_cols;
  if (c->lf) {
    const int base = clampi(c->qindex / 4 + rng_range(&hr, -3, 3), 1, 63);
    P.loop_filter_level[0] = clampi(base + rng_range(&hr, -4, 4), 1, 63);
    P.loop_filter_level[1] = clampi(base + rng_range(&hr, -4, 4), 1, 63);
    P.loop_filter_level[2] = rng_unit(&hr) < 0.15 ? 0 : clampi(base / 2 + rng_range(&hr, -4, 4), 0, 63);
    P.loop_filter_level[3] = rng_unit(&hr) < 0.15 ? 0 : clampi(base / 2 + rng_range(&hr, -4, 4), 0, 63);
  }
  static const int default_ref_deltas[8] = {1, 0, 0, 0, -1, 0, -1, -1};
  P.loop_filter_delta_enabled = 1;
  memcpy(P.loop_filter_ref_deltas, default_ref_deltas, sizeof default_ref_deltas);
  if (c->lf_extras) {
    P.loop_filter_sharpness = rng_int(&hr, 8);
    P.loop_filter_delta_enabled = rng_unit(&hr) < 0.8;
    for (int i = 0; i < 8; i++) P.loop_filter_ref_deltas[i] = rng_range(&hr, -10, 10);
    P.loop_filter_mode_deltas[0] = rng_range(&hr, -6, 6);
    P.loop_filter_mode_deltas[1] = rng_range(&hr, -6, 6);
    P.delta_lf_multi = rng_int(&hr, 2);
    P.segmentation_enabled = 1;
    for (int s = 0; s < 8; s++)
      for (int f = 1; f <= 4; f++)
        if (rng_unit(&hr) < 0.4) {
          P.seg_feature_enabled[s][f] = 1;
          P.seg_feature_data[s][f] = rng_range(&hr, -30, 30);
        }
  }
  P.cdef_damping = rng_range(&hr, 3, 6);
  for (int i = 0; i < 8; i++) {
    static const int sec[4] = {0, 1, 2, 4};
    P.cdef_y_pri_strength[i] = rng_int(&hr, 16);
    P.cdef_y_sec_strength[i] = sec[rng_int(&hr, 4)];
    P.cdef_uv_pri_strength[i] = rng_int(&hr, 16);
    P.cdef_uv_sec_strength[i] = sec[rng_int(&hr, 4)];
  }
  const int unit = c->lr_unit_size >= 64 ? c->lr_unit_size : 64; /* luma units: 64, 128, 256 */
  for (int pl = 0; pl < np; pl++) {
    P.frame_restoration_type[pl] = 0;
    if (c->lr == 1 && pl == 0) P.frame_restoration_type[pl] = RVPF_RESTORE_WIENER;
    if (c->lr == 2) P.frame_restoration_type[pl] = 3; /* RESTORE_SWITCHABLE */
    P.loop_restoration_size[pl] = pl == 0 ? unit : ((c->ssx && c->ssy && unit > 32) ? unit >> 1 : unit);
    P.lr_unit_rows[pl] = g.lr_unit_rows[pl];
    P.lr_unit_cols[pl] = g.lr_unit_cols[pl];
  }
  if (c->grain) {
    rvpf_grain_params *G = &P.grain;
    G->apply_grain = 1;
    G->grain_seed = (uint32_t)rng_int(&hr, 65536);
    G->num_y_points = clampi(c->grain_y_points, 0, 14);
    /* increasing point values */
    int used[256] = {0};
    for (int i = 0; i < G->num_y_points; i++) {
      int v;
      do v = rng_int(&hr, 256);
      while (used[v]);
      used[v] = 1;
    }
    for (int v = 0, i = 0; v < 256; v++)
      if (used[v]) G->point_y_value[i++] = v;
    for (int i = 0; i < G->num_y_points; i++) G->point_y_scaling[i] = rng_int(&hr, 256);
    if (!c->mono) {
      G->chroma_scaling_from_luma = rng_unit(&hr) < 0.25;
      if (!G->chroma_scaling_from_luma && !(c->ssx && c->ssy && G->num_y_points == 0)) {
        for (int pl = 1; pl <= 2; pl++) {
          const int n = rng_int(&hr, 11);
          int u2[256] = {0};
          for (int i = 0; i < n; i++) {
            int v;
            do v = rng_int(&hr, 256);
            while (u2[v]);
            u2[v] = 1;
          }
          int32_t *xs = pl == 1 ? G->point_cb_value : G->point_cr_value;
          int32_t *ys = pl == 1 ? G->point_cb_scaling : G->point_cr_scaling;
          for (int v = 0, i = 0; v < 256; v++)
            if (u2[v]) xs[i++] = v;
          for (int i = 0; i < n; i++) ys[i] = rng_int(&hr, 256);
          if (pl == 1) G->num_cb_points = n;
          else G->num_cr_points = n;
        }
      }
    }
    G->grain_scaling_minus_8 = rng_int(&hr, 4);
    G->ar_coeff_lag = clampi(c->grain_lag, 0, 3);
    for (int i = 0; i < 24; i++) G->ar_coeffs_y_plus_128[i] = 128 + rng_range(&hr, -24, 24);
    for (int i = 0; i < 25; i++) G->ar_coeffs_cb_plus_128[i] = 128 + rng_range(&hr, -24, 24);
    for (int i = 0; i < 25; i++) G->ar_coeffs_cr_plus_128[i] = 128 + rng_range(&hr, -24, 24);
    G->ar_coeff_shift_minus_6 = rng_int(&hr, 4);
    G->grain_scale_shift = rng_int(&hr, 2);
    G->cb_mult = rng_int(&hr, 256);
    G->cb_luma_mult = rng_int(&hr, 256);
    G->cb_offset = rng_int(&hr, 512);
    G->cr_mult = rng_int(&hr, 256);
    G->cr_luma_mult = rng_int(&hr, 256);
    G->cr_offset = rng_int(&hr, 512);
    G->overlap_flag = rng_unit(&hr) < 0.7;
    G->clip_to_restricted_range = rng_unit(&hr) < 0.3;
  }
  if (p) *p = P;

  /* ---------------- source ---------------- */
  uint16_t *srcs[3] = {src0, src1, src2};
  uint16_t *recs[3] = {rec0, rec1, rec2};
  uint16_t *own_src[3] = {NULL, NULL, NULL};
  for (int pl = 0; pl < np; pl++) {
    uint16_t *s = srcs[pl];
    if (!s && recs[pl]) s = own_src[pl] = (uint16_t *)malloc(sizeof(uint16_t) * (size_t)g.plane_w[pl] * g.plane_h[pl]);
    if (s) gen_source(c, pl, s, g.plane_w[pl], g.plane_h[pl]);
    srcs[pl] = s;
  }

  /* ---------------- partitions + mode info ---------------- */
  const size_t n4x4 = (size_t)g.mi_stride * g.mi_alloc_rows;
  part_ctx pc;
  memset(&pc, 0, sizeof pc);
  pc.c = c;
  pc.g = &g;
  pc.mi = mi ? mi : (rvpf_mi *)calloc(n4x4, sizeof(rvpf_mi));
  uint8_t *txs[3] = {tx0, tx1, tx2};
  for (int pl = 0; pl < 3; pl++) pc.tx[pl] = txs[pl] ? txs[pl] : (uint8_t *)calloc(n4x4, 1);
  memset(pc.mi, 0, n4x4 * sizeof(rvpf_mi));
  for (int pl = 0; pl < 3; pl++) memset(pc.tx[pl], 0, n4x4);
  pc.r = rng_make(c->seed, 4000, (uint64_t)c->frame_index);
  pc.key = key;
  const int sb4 = c->sb_size / 4;
  for (int r = 0; r < g.mi_rows; r += sb4)
    for (int cc = 0; cc < g.mi_cols; cc += sb4) partition(&pc, r, cc, sb4);

  /* chroma transform sizes from the block owning each chroma 4x4 */
  if (np > 1) {
    const int crow = (g.mi_alloc_rows >> c->ssy), ccol = (g.mi_stride >> c->ssx);
    for (int cy = 0; cy < crow; cy++)
      for (int cx = 0; cx < ccol; cx++) {
        const int r = (cy << c->ssy) | c->ssy, col = (cx << c->ssx) | c->ssx;
        if (r >= g.mi_alloc_rows || col >= g.mi_stride) continue;
        const rvpf_mi *m = &pc.mi[(int64_t)r * g.mi_stride + col];
        const int ps = AV1_SUBSAMPLED_SIZE[m->mi_size][c->ssx][c->ssy];
        const int t = ps == 22 ? 0 : uv_tx_size(4 * AV1_NUM_4X4_BLOCKS_WIDE[ps], 4 * AV1_NUM_4X4_BLOCKS_HIGH[ps]);
        pc.tx[1][(int64_t)cy * g.mi_stride + cx] = (uint8_t)t;
        pc.tx[2][(int64_t)cy * g.mi_stride + cx] = (uint8_t)t;
      }
  }

  /* ---------------- reconstruction ---------------- */
  const double ac = AV1_AC_QLOOKUP[(c->bit_depth - 8) >> 1][clampi(c->qindex, 0, 255)];
  const double qs = fmax(1.0, ac / 16.0);
  for (int pl = 0; pl < np; pl++)
    if (recs[pl]) memcpy(recs[pl], srcs[pl], sizeof(uint16_t) * (size_t)g.plane_w[pl] * g.plane_h[pl]);
  for (int k = 0; k < pc.nblk; k++) {
    const int *e = pc.blk[k];
    const int r = e[0], col = e[1], w4 = e[2], h4 = e[3], tx = e[4], skip = e[5];
    if (recs[0]) {
      const int tw = AV1_TX_WIDTH[tx], th = AV1_TX_HEIGHT[tx];
      for (int y = 0; y < h4 * 4; y += th)
        for (int x = 0; x < w4 * 4; x += tw)
          quantise_rect(srcs[0], recs[0], g.plane_w[0], col * 4 + x, r * 4 + y, tw, th, g.plane_w[0],
                        g.plane_h[0], qs, skip, maxv);
    }
    if (np > 1) {
      /* the block owns chroma if it contains an MI at (odd row | ssy, odd col | ssx) */
      int owns = 0;
      for (int i = 0; i < h4 && !owns; i++)
        for (int j = 0; j < w4 && !owns; j++)
          owns = (((r + i) & c->ssy) == c->ssy) && (((col + j) & c->ssx) == c->ssx);
      if (!owns) continue;
      const int b = block_enum(w4, h4);
      const int ps = AV1_SUBSAMPLED_SIZE[b][c->ssx][c->ssy];
      if (ps == 22) continue;
      const int cw = 4 * AV1_NUM_4X4_BLOCKS_WIDE[ps], ch = 4 * AV1_NUM_4X4_BLOCKS_HIGH[ps];
      const int cx0 = ((col * 4) >> c->ssx) & ~(cw - 1), cy0 = ((r * 4) >> c->ssy) & ~(ch - 1);
      const int t = uv_tx_size(cw, ch);
      const int tw = AV1_TX_WIDTH[t], th = AV1_TX_HEIGHT[t];
      for (int pl = 1; pl < 3; pl++) {
        if (!recs[pl]) continue;
        for (int y = 0; y < ch; y += th)
          for (int x = 0; x < cw; x += tw)
            quantise_rect(srcs[pl], recs[pl], g.plane_w[pl], cx0 + x, cy0 + y, tw, th, g.plane_w[pl],
                          g.plane_h[pl], qs, skip, maxv);
      }
    }
  }

  /* ---------------- cdef_idx ---------------- */
  if (cdef_idx) {
    rng_t cr = rng_make(c->seed, 5000, (uint64_t)c->frame_index);
    for (int i = 0; i < g.cdef_stride * g.cdef_rows; i++) cdef_idx[i] = -1;
    for (int r64 = 0; r64 * 16 < g.mi_rows; r64++)
      for (int c64 = 0; c64 * 16 < g.mi_cols; c64++) {
        int any = 0;
        for (int i = 0; i < 16 && !any; i++)
          for (int j = 0; j < 16 && !any; j++) {
            const int r = r64 * 16 + i, col = c64 * 16 + j;
            if (r < g.mi_rows && col < g.mi_cols) any = !(pc.mi[(int64_t)r * g.mi_stride + col].seg_skip & 1);
          }
        cdef_idx[r64 * g.cdef_stride + c64] = any ? (int8_t)rng_int(&cr, 1 << clampi(c->cdef_bits, 0, 3)) : -1;
      }
  }

  /* ---------------- restoration units ---------------- */
  rvpf_lr_unit *lrs[3] = {lr0, lr1, lr2};
  rng_t lr_r = rng_make(c->seed, 6000, (uint64_t)c->frame_index);
  for (int pl = 0; pl < np; pl++) {
    if (!lrs[pl]) continue;
    const int n = g.lr_unit_rows[pl] * g.lr_unit_cols[pl];
    memset(lrs[pl], 0, sizeof(rvpf_lr_unit) * (size_t)n);
    if (P.frame_restoration_type[pl] == 0) continue;
    for (int u = 0; u < n; u++) {
      rvpf_lr_unit *U = &lrs[pl][u];
      int type;
      if (P.frame_restoration_type[pl] == RVPF_RESTORE_WIENER) type = RVPF_RESTORE_WIENER;
      else {
        const double v = rng_unit(&lr_r);
        type = v < 0.2 ? RVPF_RESTORE_NONE : (v < 0.6 ? RVPF_RESTORE_WIENER : RVPF_RESTORE_SGRPROJ);
      }
      U->type = (uint8_t)type;
      for (int pass = 0; pass < 2; pass++)
        for (int i = 0; i < 3; i++)
          U->wiener[pass][i] =
              (int8_t)((pl > 0 && i == 0) ? 0 : rng_range(&lr_r, AV1_WIENER_TAPS_MIN[i], AV1_WIENER_TAPS_MAX[i]));
      const int set = rng_int(&lr_r, 16);
      U->sgr_set = (uint8_t)set;
      int x0 = rng_range(&lr_r, AV1_SGRPROJ_XQD_MIN[0], AV1_SGRPROJ_XQD_MAX[0]);
      int x1 = rng_range(&lr_r, AV1_SGRPROJ_XQD_MIN[1], AV1_SGRPROJ_XQD_MAX[1]);
      if (AV1_SGR_PARAMS[set][0] == 0) x0 = 0;
      if (AV1_SGR_PARAMS[set][2] == 0) x1 = clampi((1 << 7) - x0, AV1_SGRPROJ_XQD_MIN[1], AV1_SGRPROJ_XQD_MAX[1]);
      U->sgr_xqd[0] = (int8_t)x0;
      U->sgr_xqd[1] = (int8_t)x1;
    }
  }

  if (!mi) free(pc.mi);
  for (int pl = 0; pl < 3; pl++)
    if (!txs[pl]) free(pc.tx[pl]);
  for (int pl = 0; pl < 3; pl++) free(own_src[pl]);
  free(pc.blk);
  return 0;
}
