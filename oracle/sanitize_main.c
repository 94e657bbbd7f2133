/*
 * sanitize_main.c — drives the oracle and the synthetic workload generator
 * over small frames of every layout, for a host build under AddressSanitizer
 * and UndefinedBehaviorSanitizer (test infrastructure: built and run by
 * tests/test_sanitizers.py). Every buffer is allocated at exactly the size
 * the layout of include/rvpf.h gives it, so an access past any of them is
 * reported. Prints one line per frame and exits non-zero on a failure.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../rivet-av1_amd/csrc/synth.h"
#include "rvpf_oracle.h"

typedef struct {
  synth_geom g;
  rvpf_frame_params p;
  rvpf_mi *mi;
  uint8_t *tx[3];
  int8_t *cdef_idx;
  rvpf_lr_unit *lr[3];
  uint16_t *src[3], *rec[3];
} frame_t;

static void *xmalloc(size_t n) {
  void *p = malloc(n ? n : 1);
  if (!p) {
    fprintf(stderr, "out of memory\n");
    exit(2);
  }
  return p;
}

static int make_frame(const synth_cfg *c, frame_t *f) {
  memset(f, 0, sizeof *f);
  if (synth_geometry(c, &f->g)) return -1;
  const synth_geom *g = &f->g;
  const size_t n4 = (size_t)g->mi_stride * g->mi_alloc_rows;
  f->mi = xmalloc(n4 * sizeof(rvpf_mi));
  for (int i = 0; i < 3; i++) f->tx[i] = xmalloc(n4);
  f->cdef_idx = xmalloc((size_t)g->cdef_stride * g->cdef_rows);
  for (int i = 0; i < 3; i++) {
    const size_t nu = (size_t)g->lr_unit_rows[i] * g->lr_unit_cols[i];
    f->lr[i] = xmalloc((nu ? nu : 1) * sizeof(rvpf_lr_unit));
    if (i < g->num_planes) {
      f->src[i] = xmalloc((size_t)g->plane_w[i] * g->plane_h[i] * 2);
      f->rec[i] = xmalloc((size_t)g->plane_w[i] * g->plane_h[i] * 2);
    }
  }
  return synth_frame(c, &f->p, f->mi, f->tx[0], f->tx[1], f->tx[2], f->cdef_idx, f->lr[0], f->lr[1], f->lr[2],
                     f->src[0], f->src[1], f->src[2], f->rec[0], f->rec[1], f->rec[2]);
}

static void free_frame(frame_t *f) {
  free(f->mi);
  free(f->cdef_idx);
  for (int i = 0; i < 3; i++) {
    free(f->tx[i]);
    free(f->lr[i]);
    free(f->src[i]);
    free(f->rec[i]);
  }
}

static rvpf_side_info side_of(const frame_t *f) {
  rvpf_side_info s;
  memset(&s, 0, sizeof s);
  s.mi = f->mi;
  s.mi_stride = f->g.mi_stride;
  for (int i = 0; i < 3; i++) {
    s.tx_size[i] = f->tx[i];
    s.lr_units[i] = f->lr[i];
  }
  s.cdef_idx = f->cdef_idx;
  s.cdef_stride = f->g.cdef_stride;
  return s;
}

/* planes of rows x width[i] (stride = width) copied from `from` or zeroed */
static orc_planes new_planes(const frame_t *f, const int32_t *width, uint16_t *const *from) {
  orc_planes o;
  memset(&o, 0, sizeof o);
  for (int i = 0; i < f->g.num_planes; i++) {
    const size_t n = (size_t)width[i] * f->g.plane_h[i];
    o.data[i] = xmalloc(n * 2);
    o.stride[i] = width[i];
    if (from)
      memcpy(o.data[i], from[i], n * 2);
    else
      memset(o.data[i], 0, n * 2);
  }
  return o;
}

static void free_planes(orc_planes *o) {
  for (int i = 0; i < 3; i++) free(o->data[i]);
}

static int run(const synth_cfg *c) {
  frame_t f;
  if (make_frame(c, &f)) {
    fprintf(stderr, "synth failed for %dx%d\n", c->width, c->height);
    return 1;
  }
  const synth_geom *g = &f.g;
  const rvpf_side_info s = side_of(&f);
  /* the chain: deblocked recon, reference and display frames */
  orc_planes rec = new_planes(&f, g->plane_w, f.rec);
  orc_planes ref = new_planes(&f, g->ref_w, NULL);
  uint16_t *disp[3] = {NULL, NULL, NULL};
  for (int i = 0; i < g->num_planes; i++) disp[i] = xmalloc((size_t)g->crop_w[i] * g->crop_h[i] * 2);
  int rc = orc_postfilter_frame(&f.p, &s, &rec, &ref, disp, 2);
  /* stages on their own, and the encoder searches */
  orc_planes deb = new_planes(&f, g->plane_w, f.rec);
  orc_planes cd = new_planes(&f, g->plane_w, NULL);
  orc_planes src = new_planes(&f, g->plane_w, f.src);
  rc |= orc_loop_filter(&f.p, &s, &deb, 2);
  rc |= orc_cdef(&f.p, &s, &deb, &cd, 2);
  if (f.p.use_superres) {
    orc_planes up = new_planes(&f, g->ref_w, NULL);
    rc |= orc_upscale(&f.p, &cd, &up);
    free_planes(&up);
  } else {
    orc_planes lr = new_planes(&f, g->plane_w, NULL);
    rc |= orc_lr(&f.p, &s, &deb, &cd, &lr, 2);
    free_planes(&lr);
    const int sets[6] = {0, 3, 6, 10, 13, 14};
    for (int pl = 0; pl < g->num_planes; pl++)
      for (int ur = 0; ur < f.p.lr_unit_rows[pl]; ur++)
        for (int uc = 0; uc < f.p.lr_unit_cols[pl]; uc++) {
          rvpf_lr_unit u;
          double cost;
          rc |= orc_lr_search_unit(&f.p, &deb, &cd, &src, pl, ur, uc, 1000.0, sets, 6, 1 + (uc & 1), &u, &cost);
        }
    rvpf_cdef_search_result res;
    int8_t *table = xmalloc((size_t)g->cdef_stride * g->cdef_rows);
    memcpy(table, f.cdef_idx, (size_t)g->cdef_stride * g->cdef_rows);
    int nb = 0;
    rc |= orc_cdef_search(&f.p, &s, &deb, &src, c->qindex, 3000.0, c->seed & 1, 2, &res, table, NULL, NULL, &nb);
    free(table);
    int32_t lv[4];
    uint64_t sse[6][3];
    orc_planes rec2 = new_planes(&f, g->plane_w, f.rec);
    rc |= orc_lf_search(&f.p, &s, &rec2, &src, 2, lv, sse);
    free_planes(&rec2);
  }
  printf("%dx%d %d-bit ss %d,%d mono %d sr %d: rc %d\n", c->width, c->height, c->bit_depth, c->ssx, c->ssy,
         c->mono, c->superres_denom, rc);
  for (int i = 0; i < 3; i++) free(disp[i]);
  free_planes(&rec);
  free_planes(&ref);
  free_planes(&deb);
  free_planes(&cd);
  free_planes(&src);
  free_frame(&f);
  return rc != 0;
}

int main(void) {
  /* w, h, bd, ssx, ssy, mono, sb, q, lr, unit, grain lag, lf_extras, superres */
  static const int cases[][13] = {
      {66, 66, 8, 1, 1, 0, 64, 100, 2, 64, 3, 0, 0},     {226, 130, 10, 1, 1, 0, 128, 160, 2, 64, 2, 1, 0},
      {130, 70, 12, 0, 0, 1, 64, 120, 2, 128, 1, 0, 0},  {150, 100, 8, 1, 0, 0, 64, 180, 2, 64, 0, 1, 0},
      {170, 90, 10, 0, 0, 0, 64, 64, 2, 64, 3, 0, 0},    {257, 129, 10, 1, 1, 0, 128, 90, 2, 256, 3, 1, 0},
      {17, 9, 8, 1, 1, 0, 64, 200, 2, 64, 3, 0, 0},      {200, 96, 8, 1, 1, 0, 64, 120, 2, 64, 3, 0, 16},
      {170, 90, 12, 0, 0, 0, 64, 64, 2, 64, 2, 0, 12},   {130, 70, 8, 1, 1, 1, 64, 120, 2, 64, 3, 0, 11},
      {1, 1, 8, 1, 1, 0, 64, 60, 2, 64, 3, 1, 0},        {7, 5, 10, 0, 0, 0, 64, 60, 2, 64, 3, 1, 0},
      {16, 200, 12, 1, 0, 0, 64, 60, 2, 64, 3, 1, 0},    {9, 1, 8, 1, 1, 1, 64, 60, 2, 64, 3, 1, 0},
  };
  int bad = 0;
  for (size_t k = 0; k < sizeof cases / sizeof cases[0]; k++) {
    const int *v = cases[k];
    synth_cfg c;
    memset(&c, 0, sizeof c);
    c.width = v[0];
    c.height = v[1];
    c.bit_depth = v[2];
    c.ssx = v[3];
    c.ssy = v[4];
    c.mono = v[5];
    if (c.mono) c.ssx = c.ssy = 1;
    c.sb_size = v[6];
    c.qindex = v[7];
    c.seed = 100 + k;
    c.frame_index = (int)(k % 3);
    c.pan_px = 1.5;
    c.lf = 1;
    c.lf_extras = v[11];
    c.cdef = 1;
    c.cdef_bits = 3;
    c.lr = v[8];
    c.lr_unit_size = v[9];
    c.grain = 1;
    c.grain_lag = v[10];
    c.grain_y_points = 8;
    c.superres_denom = v[12];
    bad |= run(&c);
  }
  return bad;
}
