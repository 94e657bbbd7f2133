/*
 * lf.hip — deblocking loop filter (AV1 7.14) for gfx950.
 *
 * Follows rivet-av1's loop_filter_threads / edge_loop_filter /
 * filter_strength (src/decoder/postfilter.rs:67-264) and the sample filter
 * dsp::lf::filter_lanes (src/dsp/lf.rs:58-211).
 *
 * One launch per pass (vertical edges, then horizontal edges), all planes
 * flattened into one grid, in place. A workgroup takes a block of 256
 * 4-sample edges (64 x 4 in the per-frame launch): each thread decides one
 * edge (7.14.2-7.14.5; the level tables sit in LDS), the filtered edges
 * are gathered into one LDS list per filter class (4, chroma 6-tap, luma
 * 8-tap, 16), and each list is then filtered with every lane on the same
 * branch-free class filter, one line of an edge per lane (a row of the
 * edge in the vertical pass, a column in the horizontal one).
 *
 * The reference walks the edges of a row in order and writes back p2..q2
 * (p5..q5) whatever filter ran; a parallel pass must not, since a 4-wide
 * edge's neighbour 4 samples on modifies those samples. Each line
 * therefore stores only its filter's footprint:
 *   filter size 4            -> p1..q1
 *   filter size 8 (6/8-tap)  -> p3..q3 (p3, q3, and for chroma p2, q2, are
 *                               written back unchanged)
 *   filter size 16           -> p7..q7 (p7, p6, q6, q7 unchanged)
 * For side info whose transform blocks are aligned to their size (every
 * AV1 stream), a size-s edge at x has no other edge in (x - s, x + s), so
 * these footprints never overlap another line's reads or writes and the
 * result equals the sequential one sample for sample, whatever order the
 * lists run in.
 */
#include <string.h>

#include "av1_tables.h"
#include "rvpf_internal.h"

namespace rvpf {

namespace {

__device__ __forceinline__ int clip3(int lo, int hi, int x) { return min(max(x, lo), hi); }

/* Everything the edge decision looks up per lane, built on the host and
 * staged in LDS once per workgroup: log2 of the transform width / height
 * of each TX_SIZE, log2 of the width / height of each block size as seen
 * by luma [0] and chroma [1] (0xff where the subsampled size is invalid,
 * i.e. every position is a block edge), and the level parameters. */
struct LfTables {
  uint8_t tx_log2[2][20];   /* [pass][tx size] */
  uint8_t blk_log2[2][2][22]; /* [luma / chroma][pass][mi_size] */
  int8_t ref_deltas[8];
  int8_t mode_deltas[2];
  uint8_t seg_lf_enabled[8];
  int8_t seg_lf_data[8][4];
  uint8_t level[4];
  uint8_t pad[2];
};
static_assert(sizeof(LfTables) % 4 == 0, "LfTables is staged in words");

/* filter_strength, src/decoder/postfilter.rs:207-264: lvl of the block
 * (header values from the LDS tables t). */
__device__ __forceinline__ int lf_level(const LfHdr &h, const LfTables &t, const rvpf_mi &m, int i) {
  const int seg = (m.seg_skip >> 1) & 7;
  const int rf = m.ref_frame0;
  const int mode = m.y_mode;
  const int mode_type = (mode >= 14 && mode != 16 && mode != 24); /* >= NEARESTMV, not GLOBALMV / GLOBAL_GLOBALMV */
  /* delta_lf[i] (or [0]) by a shift of the four packed bytes, so the
   * block's mode info stays in registers */
  const uint32_t dw = (uint32_t)(uint8_t)m.delta_lf[0] | ((uint32_t)(uint8_t)m.delta_lf[1] << 8) |
                      ((uint32_t)(uint8_t)m.delta_lf[2] << 16) | ((uint32_t)(uint8_t)m.delta_lf[3] << 24);
  const int dlf = (int)(int8_t)(dw >> (8 * (h.delta_lf_multi ? i : 0)));
  int lvl = clip3(0, 63, dlf + h.level[i]);
  if (h.seg_enabled && ((t.seg_lf_enabled[seg] >> i) & 1)) lvl = clip3(0, 63, t.seg_lf_data[seg][i] + lvl);
  if (h.delta_enabled) {
    const int ns = 1 << (lvl >> 5);
    if (rf == 0)
      lvl += t.ref_deltas[0] * ns;
    else if (rf > 0)
      lvl += (t.ref_deltas[rf & 7] + t.mode_deltas[mode_type]) * ns;
    lvl = clip3(0, 63, lvl);
  }
  return lvl;
}

/* 7.14.6 on one line of samples across an edge, for one filter class:
 *   C = 0: filter size 4 (narrow filter only),      v = p1 p0 | q0 q1
 *   C = 1: chroma size 8 (6-tap or narrow),           v = p2 .. p0 | q0 .. q2
 *   C = 2: luma size 8 (8-tap or narrow),             v = p3 .. p0 | q0 .. q3
 *   C = 3: size 16 (16-wide, 8-tap or narrow),        v = p6 .. p0 | q0 .. q6
 * Every candidate output is computed and the one the masks select is kept,
 * so all lanes of a class run the same instructions. Returns whether the
 * line was filtered (v modified in place). Same arithmetic as
 * dsp::lf::filter_lanes (src/dsp/lf.rs:58-211). */
template <int C>
struct LfClass {
  static constexpr int N = C == 0 ? 4 : (C == 1 ? 6 : (C == 2 ? 8 : 14));
};

template <int C>
__device__ __forceinline__ bool filter_class(int *v, int limit, int blimit, int thresh, int bd) {
  constexpr int N = LfClass<C>::N, H = N / 2;
  const int p0 = v[H - 1], p1 = v[H - 2], q0 = v[H], q1 = v[H + 1];
  bool mask = abs(p0 - q0) * 2 + (abs(p1 - q1) >> 1) > blimit || abs(p1 - p0) > limit || abs(q1 - q0) > limit;
  int p2 = 0, q2 = 0, p3 = 0, q3 = 0;
  if (C >= 1) {
    p2 = v[H - 3];
    q2 = v[H + 2];
    mask = mask || abs(p2 - p1) > limit || abs(q2 - q1) > limit;
  }
  if (C >= 2) {
    p3 = v[H - 4];
    q3 = v[H + 3];
    mask = mask || abs(p3 - p2) > limit || abs(q3 - q2) > limit;
  }
  if (mask) return false;
  const int sh = bd - 8;
  const int t = 1 << sh;
  bool flat = false;
  if (C >= 1) {
    flat = abs(p1 - p0) <= t && abs(q1 - q0) <= t && abs(p2 - p0) <= t && abs(q2 - q0) <= t;
    if (C >= 2) flat = flat && abs(p3 - p0) <= t && abs(q3 - q0) <= t;
  }
  /* narrow filter (7.14.6.3) */
  const bool hev = abs(p1 - p0) > thresh || abs(q1 - q0) > thresh;
  const int lo = -(1 << (bd - 1)), hi = (1 << (bd - 1)) - 1;
  const int off = 0x80 << sh;
  const int ps1 = p1 - off, ps0 = p0 - off, qs0 = q0 - off, qs1 = q1 - off;
  int fl = hev ? clip3(lo, hi, ps1 - qs1) : 0;
  fl = clip3(lo, hi, fl + 3 * (qs0 - ps0));
  const int filter1 = clip3(lo, hi, fl + 4) >> 3;
  const int filter2 = clip3(lo, hi, fl + 3) >> 3;
  const int f = (filter1 + 1) >> 1;
  const int n_q0 = clip3(lo, hi, qs0 - filter1) + off;
  const int n_p0 = clip3(lo, hi, ps0 + filter2) + off;
  const int n_q1 = hev ? q1 : clip3(lo, hi, qs1 - f) + off;
  const int n_p1 = hev ? p1 : clip3(lo, hi, ps1 + f) + off;
  if (C == 0 || !flat) {
    v[H - 2] = n_p1;
    v[H - 1] = n_p0;
    v[H] = n_q0;
    v[H + 1] = n_q1;
    return true;
  }
  if (C == 1) {
    /* log2 size 3, n = 2, taps |j| <= 1 doubled: 6-tap chroma filter */
    v[H - 2] = (3 * p2 + 2 * p1 + 2 * p0 + q0 + 4) >> 3;
    v[H - 1] = (p2 + 2 * p1 + 2 * p0 + 2 * q0 + q1 + 4) >> 3;
    v[H] = (p1 + 2 * p0 + 2 * q0 + 2 * q1 + q2 + 4) >> 3;
    v[H + 1] = (p0 + 2 * q0 + 2 * q1 + 3 * q2 + 4) >> 3;
    return true;
  }
  bool flat2 = false;
  if (C == 3) {
    flat2 = abs(v[0] - p0) <= t && abs(v[13] - q0) <= t && abs(v[1] - p0) <= t && abs(v[12] - q0) <= t &&
            abs(v[2] - p0) <= t && abs(v[11] - q0) <= t;
  }
  if (C == 2 || !flat2) {
    /* log2 size 3, n = 3, centre tap doubled: 8-tap flat filter */
    v[H - 3] = (3 * p3 + 2 * p2 + p1 + p0 + q0 + 4) >> 3;
    v[H - 2] = (2 * p3 + p2 + 2 * p1 + p0 + q0 + q1 + 4) >> 3;
    v[H - 1] = (p3 + p2 + p1 + 2 * p0 + q0 + q1 + q2 + 4) >> 3;
    v[H] = (p2 + p1 + p0 + 2 * q0 + q1 + q2 + q3 + 4) >> 3;
    v[H + 1] = (p1 + p0 + q0 + 2 * q1 + q2 + 2 * q3 + 4) >> 3;
    v[H + 2] = (p0 + q0 + q1 + 2 * q2 + 3 * q3 + 4) >> 3;
    return true;
  }
  /* log2 size 4, n = 6, taps |j| <= 1 doubled: 16-wide filter. Output i
   * (-6..5) sums v[7 + clamp(i + j, -7, 6)] over j = -6..6; the window
   * sums slide from one output to the next. */
  int x[14];
#pragma unroll
  for (int k = 0; k < 14; k++) x[k] = v[k];
  int wide = 0, mid = 0;
#pragma unroll
  for (int j = -6; j <= 6; j++) wide += x[7 + min(max(-6 + j, -7), 6)];
#pragma unroll
  for (int j = -1; j <= 1; j++) mid += x[7 + min(max(-6 + j, -7), 6)];
#pragma unroll
  for (int i = -6; i < 6; i++) {
    v[7 + i] = (wide + mid + 8) >> 4;
    wide += x[7 + min(i + 7, 6)] - x[7 + max(i - 6, -7)];
    mid += x[7 + min(i + 2, 6)] - x[7 + max(i - 1, -7)];
  }
  return true;
}

template <typename T>
struct Io;
template <>
struct Io<uint16_t> {
  typedef uint2 W; /* 4 samples */
  static __device__ __forceinline__ W ldw(const uint16_t *p) { return *reinterpret_cast<const uint2 *>(p); }
  static __device__ __forceinline__ void unpack(W u, int *o) {
    o[0] = u.x & 0xffff;
    o[1] = u.x >> 16;
    o[2] = u.y & 0xffff;
    o[3] = u.y >> 16;
  }
  /* 4 consecutive samples at p (8-byte aligned) */
  static __device__ __forceinline__ void ld4(const uint16_t *p, int *o) {
    const uint2 u = *reinterpret_cast<const uint2 *>(p);
    o[0] = u.x & 0xffff;
    o[1] = u.x >> 16;
    o[2] = u.y & 0xffff;
    o[3] = u.y >> 16;
  }
  static __device__ __forceinline__ void st4(uint16_t *p, const int *o) {
    uint2 u;
    u.x = (uint32_t)o[0] | ((uint32_t)o[1] << 16);
    u.y = (uint32_t)o[2] | ((uint32_t)o[3] << 16);
    *reinterpret_cast<uint2 *>(p) = u;
  }
  /* 2 consecutive samples at p (4-byte aligned) */
  static __device__ __forceinline__ void st2(uint16_t *p, int a, int b) {
    *reinterpret_cast<uint32_t *>(p) = (uint32_t)a | ((uint32_t)b << 16);
  }
};
template <>
struct Io<uint8_t> {
  typedef uint32_t W; /* 4 samples */
  static __device__ __forceinline__ W ldw(const uint8_t *p) { return *reinterpret_cast<const uint32_t *>(p); }
  static __device__ __forceinline__ void unpack(W u, int *o) {
    o[0] = u & 0xff;
    o[1] = (u >> 8) & 0xff;
    o[2] = (u >> 16) & 0xff;
    o[3] = u >> 24;
  }
  static __device__ __forceinline__ void ld4(const uint8_t *p, int *o) {
    const uint32_t u = *reinterpret_cast<const uint32_t *>(p);
    o[0] = u & 0xff;
    o[1] = (u >> 8) & 0xff;
    o[2] = (u >> 16) & 0xff;
    o[3] = u >> 24;
  }
  static __device__ __forceinline__ void st4(uint8_t *p, const int *o) {
    *reinterpret_cast<uint32_t *>(p) =
        (uint32_t)o[0] | ((uint32_t)o[1] << 8) | ((uint32_t)o[2] << 16) | ((uint32_t)o[3] << 24);
  }
  static __device__ __forceinline__ void st2(uint8_t *p, int a, int b) {
    *reinterpret_cast<uint16_t *>(p) = (uint16_t)(a | (b << 8));
  }
};

struct LfLaunch {
  Geom g;
  LfHdr h;
  LfTables t;
  LfSide s;
  PlanePtr planes[3];
  int plane_id[3];   /* planes filtered, in grid order */
  int blk_start[4];  /* first block of each grid plane (padded to 8), and the total */
  int blk_x[3];      /* blocks across each grid plane */
  int blk_n[3];      /* blocks of each grid plane */
  int nseg;
};

/* Mode info and transform sizes an edge decision reads, loaded together
 * with the samples: the block of the edge (m), the block before it (pm,
 * used when m's level is 0), the transform sizes of both sides. */
struct EdgeSide {
  rvpf_mi m, pm;
  int tx, ptx;
};

__device__ __forceinline__ EdgeSide load_side(const LfLaunch &a, int plane, int pass, int ex, int ey) {
  const Geom &g = a.g;
  const int ssx = plane ? g.ssx : 0, ssy = plane ? g.ssy : 0;
  const int rowo = (ey << ssy) | ssy, colo = (ex << ssx) | ssx;
  const int ms = a.s.mi_stride;
  const uint2 *mi = gptr(reinterpret_cast<const uint2 *>(a.s.mi));
  const int prow = pass == 0 ? rowo : rowo - (1 << ssy);
  const int pcol = pass == 0 ? colo - (1 << ssx) : colo;
  const uint8_t *tx = gptr(a.s.tx[plane]);
  EdgeSide e;
  e.m = __builtin_bit_cast(rvpf_mi, mi[(int64_t)rowo * ms + colo]);
  e.pm = __builtin_bit_cast(rvpf_mi, mi[(int64_t)prow * ms + pcol]);
  e.tx = tx[(int64_t)ey * ms + ex];
  e.ptx = pass == 0 ? tx[(int64_t)ey * ms + ex - 1] : tx[(int64_t)(ey - 1) * ms + ex];
  return e;
}

/* Edge decision of edge_loop_filter (postfilter.rs:115-172). Returns the
 * filter size (0: no filter) and the strength. t: the LDS tables. */
__device__ __forceinline__ int edge_decision(const LfLaunch &a, const LfTables &t, const EdgeSide &e, int plane,
                                             int pass, int ex, int ey, int *limit_bd, int *blimit_bd,
                                             int *thresh_bd) {
  const Geom &g = a.g;
  const int ssx = plane ? g.ssx : 0, ssy = plane ? g.ssy : 0;
  const int x = (ex << ssx) * 4, y = (ey << ssy) * 4;
  if (x >= g.width || y >= g.height) return 0;
  const int pos = pass == 0 ? x >> ssx : y >> ssy;
  const rvpf_mi &m = e.m;
  const int tx_sz = min(e.tx, 18), prev_tx_sz = min(e.ptx, 18);
  /* BLOCK_INVALID (log2 0xff): a block edge everywhere */
  const int bl = t.blk_log2[plane != 0][pass][min((int)m.mi_size, 21)];
  const bool is_block_edge = bl == 0xff || (pos & ((1 << bl) - 1)) == 0;
  const int tl = t.tx_log2[pass][tx_sz], ptl = t.tx_log2[pass][prev_tx_sz];
  const bool is_tx_edge = (pos & ((1 << tl) - 1)) == 0;
  const bool skip = m.seg_skip & 1;
  const bool is_intra = m.ref_frame0 <= 0;
  if (!(is_tx_edge && (is_block_edge || !skip || is_intra))) return 0;
  const int base_size = 1 << min(tl, ptl);
  int filter_size = min(plane == 0 ? 16 : 8, base_size);
  /* A 16-wide edge of legal side info is at least 16 samples from the
   * plane's edge; keep the reads in bounds for side info that is not. */
  if (filter_size == 16 && pos < 8) filter_size = 8;
  const int li = plane == 0 ? pass : plane + 1;
  int lvl = lf_level(a.h, t, m, li);
  if (lvl == 0) lvl = lf_level(a.h, t, e.pm, li);
  if (lvl == 0) return 0;
  const int sharp = a.h.sharpness;
  const int shift = sharp > 4 ? 2 : (sharp > 0 ? 1 : 0);
  const int limit = sharp > 0 ? clip3(1, 9 - sharp, lvl >> shift) : max(lvl >> shift, 1);
  /* unscaled (the filters shift them by bd - 8) */
  *limit_bd = limit;
  *blimit_bd = 2 * (lvl + 2) + limit;
  *thresh_bd = lvl >> 4;
  return filter_size;
}

constexpr int LF_TWORDS = sizeof(LfTables) / 4;

/* The block's samples staged in LDS (16-bit at every sample size): in the
 * vertical pass rows [4 ey0, 4 ey0 + 4 EY) x columns [4 ex0 - 8,
 * 4 ex0 + 4 EX + 8), in the horizontal pass rows [4 ey0 - 8,
 * 4 ey0 + 4 EY + 8) x columns [4 ex0, 4 ex0 + 4 EX): every sample a
 * filter of the block reads. */
template <int PASS, int EX>
struct LfTile {
  static constexpr int EY = 256 / EX;
  static constexpr int W = PASS == 0 ? 4 * EX + 16 : 4 * EX; /* samples per row */
  static constexpr int R = PASS == 0 ? 4 * EY : 4 * EY + 16;
  static constexpr int CW = W / 4; /* 4-sample chunks per row */
  static constexpr int CHUNKS = R * CW;
  static constexpr int PER = (CHUNKS + 255) / 256;
};

/* LDS of a deblocking workgroup: the staged samples, the tables, and the
 * filtered edges of the block sorted by filter class (local edge index |
 * limit << 8 | blimit << 16 | thresh << 24, unscaled). */
template <int PASS, int EX>
struct LfShared {
  __align__(16) uint16_t tile[LfTile<PASS, EX>::R * LfTile<PASS, EX>::W];
  LfTables t;
  int count[4];
  uint32_t list[4][256];
};

/* 4 staged samples (8-byte aligned) */
__device__ __forceinline__ void lds4(const uint16_t *p, int *o) {
  const uint2 u = *reinterpret_cast<const uint2 *>(p);
  o[0] = u.x & 0xffff;
  o[1] = u.x >> 16;
  o[2] = u.y & 0xffff;
  o[3] = u.y >> 16;
}

/* One filter class of the block: one line of an edge per item (a row in
 * the vertical pass, a column in the horizontal one), read from the staged
 * samples; only the filter's footprint is stored to the plane. */
template <typename T, int PASS, int EX, int C>
__device__ __forceinline__ void lf_class(const LfShared<PASS, EX> &sm, T *base, int64_t stride, int ex0, int ey0,
                                         int bd) {
  typedef LfTile<PASS, EX> TL;
  constexpr int N = LfClass<C>::N, H = N / 2;
  const int sh = bd - 8;
  const int n = sm.count[C];
  const int tx0 = PASS == 0 ? 4 * ex0 - 8 : 4 * ex0, ty0 = PASS == 0 ? 4 * ey0 : 4 * ey0 - 8;
  if (PASS == 0) {
    for (int it = threadIdx.x; it < 4 * n; it += 256) {
      const uint32_t e = sm.list[C][it >> 2];
      const int li = e & 255;
      const int xp = (ex0 + (li % EX)) * 4, yp = (ey0 + (li / EX)) * 4 + (it & 3);
      const uint16_t *trow = sm.tile + (yp - ty0) * TL::W + (xp - tx0);
      int s[16];
      if (C == 3) {
        lds4(trow - 8, s);
        lds4(trow + 4, s + 12);
      }
      lds4(trow - 4, s + 4);
      lds4(trow, s + 8);
      int v[N];
#pragma unroll
      for (int k = 0; k < N; k++) v[k] = s[8 - H + k];
      if (!filter_class<C>(v, ((e >> 8) & 255) << sh, ((e >> 16) & 255) << sh, (e >> 24) << sh, bd)) continue;
#pragma unroll
      for (int k = 0; k < N; k++) s[8 - H + k] = v[k];
      T *row = base + (int64_t)yp * stride + xp;
      if (C <= 1) {
        Io<T>::st2(row - 2, s[6], s[7]);
        Io<T>::st2(row, s[8], s[9]);
      } else {
        if (C == 3) {
          Io<T>::st4(row - 8, s);
          Io<T>::st4(row + 4, s + 12);
        }
        Io<T>::st4(row - 4, s + 4);
        Io<T>::st4(row, s + 8);
      }
    }
  } else {
    for (int it = threadIdx.x; it < 4 * n; it += 256) {
      const uint32_t e = sm.list[C][it >> 2];
      const int li = e & 255;
      const int xp = (ex0 + (li % EX)) * 4 + (it & 3), yp = (ey0 + (li / EX)) * 4;
      const uint16_t *tcol = sm.tile + (yp - ty0) * TL::W + (xp - tx0);
      int v[N];
#pragma unroll
      for (int k = 0; k < N; k++) v[k] = tcol[(k - H) * TL::W];
      if (!filter_class<C>(v, ((e >> 8) & 255) << sh, ((e >> 16) & 255) << sh, (e >> 24) << sh, bd)) continue;
      /* rows the filters modify: p1..q1, p2..q2 (8-tap), p5..q5 (16) */
      constexpr int R = C <= 1 ? 2 : (C == 2 ? 3 : 6);
      T *col = base + (int64_t)yp * stride + xp;
#pragma unroll
      for (int k = H - R; k < H + R; k++) col[(int64_t)(k - H) * stride] = (T)v[k];
    }
  }
}

/* Block lb of grid plane gp: EX x (256 / EX) edges. The block's samples
 * are loaded (every load in flight together with the side info) and staged
 * in LDS; each thread decides one edge (7.14.2-7.14.5); the filtered edges
 * are gathered into one list per filter class, and each list is then
 * filtered with every lane on the same code.
 *
 * Staging bounds: a filtered edge of the vertical pass reads columns
 * x - 8 .. x + 7 of rows inside the mode-info area, one of the horizontal
 * pass rows y - 8 .. y + 7 of columns inside it; x + 7 and y + 7 stay below
 * the area plus the 8-sample margin every deblocked plane has. Chunks
 * outside those bounds are loaded from a clamped position and never read. */
template <typename T, int PASS, int EX>
__device__ __forceinline__ void lf_block(const LfLaunch &a, int gp, int lb, LfShared<PASS, EX> *sm) {
  typedef LfTile<PASS, EX> TL;
  const int plane = a.plane_id[gp];
  const int ex0 = (lb % a.blk_x[gp]) * EX, ey0 = (lb / a.blk_x[gp]) * (256 / EX);
  const int ex = ex0 + (threadIdx.x % EX), ey = ey0 + (threadIdx.x / EX);
  const Geom &g = a.g;
  const int ssx = plane ? g.ssx : 0, ssy = plane ? g.ssy : 0;
  const int xp = ex * 4, yp = ey * 4;
  /* inside the mode-info area; the edge is not filtered at the frame's
   * left / top border */
  const bool valid = (ey << ssy) < g.mi_rows && (ex << ssx) < g.mi_cols && (PASS == 0 ? xp != 0 : yp != 0);
  const uint32_t tword = threadIdx.x < LF_TWORDS ? reinterpret_cast<const uint32_t *>(&a.t)[threadIdx.x] : 0u;
  EdgeSide side;
  if (valid) side = load_side(a, plane, PASS, ex, ey);
  T *base = gptr(reinterpret_cast<T *>(a.planes[plane].data));
  const int64_t stride = a.planes[plane].stride;
  const int area_w = (g.mi_cols * 4) >> ssx, area_h = (g.mi_rows * 4) >> ssy;
  const int tx0 = PASS == 0 ? 4 * ex0 - 8 : 4 * ex0, ty0 = PASS == 0 ? 4 * ey0 : 4 * ey0 - 8;
  const int cmax = PASS == 0 ? area_w + 4 : area_w - 4, rmax = PASS == 0 ? area_h - 1 : area_h + 7;
  typename Io<T>::W v[TL::PER];
#pragma unroll
  for (int k = 0; k < TL::PER; k++) {
    const int ch = min((int)threadIdx.x + 256 * k, TL::CHUNKS - 1);
    const int r = ch / TL::CW, c = (ch - r * TL::CW) * 4;
    const int y = min(max(ty0 + r, 0), rmax), x = min(max(tx0 + c, 0), cmax);
    v[k] = Io<T>::ldw(base + (int64_t)y * stride + x);
  }
  if (threadIdx.x < LF_TWORDS) reinterpret_cast<uint32_t *>(&sm->t)[threadIdx.x] = tword;
  if (threadIdx.x < 4) sm->count[threadIdx.x] = 0;
#pragma unroll
  for (int k = 0; k < TL::PER; k++) {
    const int ch = threadIdx.x + 256 * k;
    if (ch < TL::CHUNKS) {
      int o[4];
      Io<T>::unpack(v[k], o);
      *reinterpret_cast<uint2 *>(&sm->tile[ch * 4]) =
          make_uint2((uint32_t)o[0] | ((uint32_t)o[1] << 16), (uint32_t)o[2] | ((uint32_t)o[3] << 16));
    }
  }
  __syncthreads();
  int cls = -1; /* filter class of this lane's edge, -1: not filtered */
  uint32_t word = 0;
  if (valid) {
    int limit, blimit, thresh;
    const int fs = edge_decision(a, sm->t, side, plane, PASS, ex, ey, &limit, &blimit, &thresh);
    if (fs) {
      cls = fs == 4 ? 0 : (fs == 16 ? 3 : (plane ? 1 : 2));
      word = threadIdx.x | (limit << 8) | (blimit << 16) | (thresh << 24);
    }
  }
  /* wave-aggregated append: a ballot per class gives the wave's edges of
   * each class; lanes 0..3 reserve the wave's run of class 0..3 with one
   * LDS atomic instruction, a shuffle hands each lane its class's base and
   * its slot adds the number of the class's lanes below it (mbcnt). The
   * order within a list does not matter: the edges of a class write
   * disjoint footprints and read the staged samples. */
  {
    const uint64_t m0 = __ballot(cls == 0), m1 = __ballot(cls == 1), m2 = __ballot(cls == 2), m3 = __ballot(cls == 3);
    const int lane = threadIdx.x & 63;
    const uint64_t mk = lane == 0 ? m0 : (lane == 1 ? m1 : (lane == 2 ? m2 : m3));
    int base = 0;
    if (lane < 4 && mk != 0) base = atomicAdd(&sm->count[lane], __popcll(mk));
    const uint64_t mc = cls == 0 ? m0 : (cls == 1 ? m1 : (cls == 2 ? m2 : m3));
    const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(mc >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mc, 0u));
    base = __shfl(base, cls < 0 ? 0 : cls, 64);
    if (cls >= 0) sm->list[cls][base + below] = word;
  }
  __syncthreads();
  lf_class<T, PASS, EX, 0>(*sm, base, stride, ex0, ey0, g.bd);
  if (plane) {
    lf_class<T, PASS, EX, 1>(*sm, base, stride, ex0, ey0, g.bd);
  } else {
    lf_class<T, PASS, EX, 2>(*sm, base, stride, ex0, ey0, g.bd);
    lf_class<T, PASS, EX, 3>(*sm, base, stride, ex0, ey0, g.bd);
  }
}

template <typename T, int PASS, int EX>
__global__ __launch_bounds__(256) void lf_pass_kernel(LfLaunch a) {
  __shared__ LfShared<PASS, EX> tsm;
  int lb;
  const int gp = xcd_segment(blockIdx.x, a.blk_start, a.nseg, &lb);
  if (gp >= a.nseg || lb >= a.blk_n[gp]) return; /* padding */
  lf_block<T, PASS, EX>(a, gp, lb, &tsm);
}

/* ------------------------------------------------------------------ */
/* Both passes in one launch, out of place (the whole chain's deblock).
 *
 * A workgroup produces one 64 x 64 block of one plane's deblocked frame at
 * (X0, Y0). Its samples after both passes depend on the vertical edges at
 * X0, X0 + 4, .., X0 + 64 of rows Y0 - 8 .. Y0 + 71 (the horizontal edges
 * read 7 rows either side of theirs) and on the horizontal edges at Y0,
 * Y0 + 4, .., Y0 + 64 of its columns: the edges at X0 + 64 and Y0 + 64 are
 * the next blocks' first ones, whose p side reaches into this block, and no
 * edge left of X0 or above Y0 reaches into it (transform blocks do not
 * cross a 64 boundary, so their q-side footprint ends before it). Edges on
 * the block boundaries are filtered by both blocks alike.
 *
 * The block stages its 80 x 80 input samples (rows and columns 8 either
 * side) in LDS, decides every edge (both passes at once: the decision
 * reads side info only), filters the vertical class lists in LDS in place,
 * then the horizontal ones, and writes its 64 x 64 samples, clipped to the
 * mode-info area, to the output planes. Reading the input and writing
 * another buffer keeps blocks from seeing each other's results; in LDS a
 * line never reads another line's footprint (legal side info, as above). */
constexpr int FT = 80;                 /* staged samples per row and rows */
constexpr int FV_COLS = 17, FV_ROWS = 20; /* vertical edges: columns x rows of 4 lines */
constexpr int FH_ROWS = 17, FH_COLS = 16; /* horizontal edges */
constexpr int FV_N = FV_COLS * FV_ROWS, FH_N = FH_ROWS * FH_COLS;
constexpr int F_ITEMS = (FV_N + FH_N + 255) / 256; /* decisions per thread */
/* side-info positions (4x4 units of the plane) the decisions read: columns
 * X0 / 4 - 1 .. X0 / 4 + 16, rows Y0 / 4 - 2 .. Y0 / 4 + 17 */
constexpr int FM_COLS = 18, FM_ROWS = 20, FM_N = FM_COLS * FM_ROWS;
constexpr int FM_ITEMS = (FM_N + 255) / 256;

struct LfFusedShared {
  __align__(16) uint16_t tile[FT * FT];
  LfTables t;
  int count[8];             /* [pass * 4 + class] */
  /* Edge indices by class, two lists per pass. A plane produces at most
   * three classes (luma 4/8/16-wide: 0, 2, 3; chroma 4/6-wide: 0, 1) and
   * each edge is in one: class 0 fills list 0 from the front, class 1 or 2
   * (never both in one plane) from the back, class 3 fills list 1. Sized
   * this way the block's LDS stays below 20 KB, 8 workgroups per CU. */
  uint16_t lv[2][FV_N];     /* vertical edges */
  uint16_t lh[2][FH_N];     /* horizontal edges */
  uint8_t elv[FV_N], elh[FH_N]; /* the edge's level */
  rvpf_mi mi[FM_N];         /* mode info at each side-info position */
  uint8_t tx[FM_N];         /* transform size */
  uint8_t lvl[2][FM_N];     /* level of the position's block in the vertical / horizontal pass */
};

/* limit, blimit and thresh of a level (unscaled; filter_strength,
 * postfilter.rs:242-252) */
__device__ __forceinline__ void level_limits(int lvl, int sharp, int *limit, int *blimit, int *thresh) {
  const int shift = sharp > 4 ? 2 : (sharp > 0 ? 1 : 0);
  const int l = sharp > 0 ? clip3(1, 9 - sharp, lvl >> shift) : max(lvl >> shift, 1);
  *limit = l;
  *blimit = 2 * (lvl + 2) + l;
  *thresh = lvl >> 4;
}

/* Edge decision of edge_loop_filter (postfilter.rs:115-172) from the
 * staged side info: m / tx of the edge's block, ptx of the block before
 * it, and the level of each (lvl, plvl). Returns the filter size (0: no
 * filter) and, in *level, the level it filters with. */
__device__ __forceinline__ int edge_decision_staged(const Geom &g, const LfTables &t, const rvpf_mi &m, int tx,
                                                    int ptx, int lvl, int plvl, int plane, int pass, int ex, int ey,
                                                    int *level) {
  const int ssx = plane ? g.ssx : 0, ssy = plane ? g.ssy : 0;
  const int x = (ex << ssx) * 4, y = (ey << ssy) * 4;
  if (x >= g.width || y >= g.height) return 0;
  const int pos = pass == 0 ? x >> ssx : y >> ssy;
  const int tx_sz = min(tx, 18), prev_tx_sz = min(ptx, 18);
  const int bl = t.blk_log2[plane != 0][pass][min((int)m.mi_size, 21)];
  const bool is_block_edge = bl == 0xff || (pos & ((1 << bl) - 1)) == 0;
  const int tl = t.tx_log2[pass][tx_sz], ptl = t.tx_log2[pass][prev_tx_sz];
  const bool is_tx_edge = (pos & ((1 << tl) - 1)) == 0;
  const bool skip = m.seg_skip & 1;
  const bool is_intra = m.ref_frame0 <= 0;
  if (!(is_tx_edge && (is_block_edge || !skip || is_intra))) return 0;
  const int base_size = 1 << min(tl, ptl);
  int filter_size = min(plane == 0 ? 16 : 8, base_size);
  /* A 16-wide edge of legal side info is at least 16 samples from the
   * plane's edge; keep the reads in bounds for side info that is not. */
  if (filter_size == 16 && pos < 8) filter_size = 8;
  const int l = lvl ? lvl : plvl;
  *level = l;
  return l ? filter_size : 0;
}

__device__ __forceinline__ void lds_st4(uint16_t *p, const int *o) {
  *reinterpret_cast<uint2 *>(p) = make_uint2((uint32_t)o[0] | ((uint32_t)o[1] << 16), (uint32_t)o[2] | ((uint32_t)o[3] << 16));
}
__device__ __forceinline__ void lds_st2(uint16_t *p, int a, int b) {
  *reinterpret_cast<uint32_t *>(p) = (uint32_t)a | ((uint32_t)b << 16);
}

/* Entry i of class c's list in the two lists of a pass (N entries each). */
template <int N>
__device__ __forceinline__ uint16_t &class_slot(uint16_t (*l)[N], int c, int i) {
  return c == 3 ? l[1][i] : (c == 0 ? l[0][i] : l[0][N - 1 - i]);
}
template <int C, int N>
__device__ __forceinline__ int class_item(const uint16_t (*l)[N], int i) {
  return C == 3 ? l[1][i] : (C == 0 ? l[0][i] : l[0][N - 1 - i]);
}

/* One vertical-edge class: a row of an edge per item, in the staged tile. */
template <int C>
__device__ __forceinline__ void fused_v(LfFusedShared &sm, int bd, int sharp) {
  constexpr int N = LfClass<C>::N, H = N / 2;
  const int sh = bd - 8, n = sm.count[C];
  for (int it = threadIdx.x; it < 4 * n; it += 256) {
    const int li = class_item<C, FV_N>(sm.lv, it >> 2);
    int limit, blimit, thresh;
    level_limits(sm.elv[li], sharp, &limit, &blimit, &thresh);
    const int col = 8 + 4 * (li % FV_COLS), row = 4 * (li / FV_COLS) + (it & 3);
    uint16_t *trow = sm.tile + row * FT + col;
    int s[16];
    if (C == 3) {
      lds4(trow - 8, s);
      lds4(trow + 4, s + 12);
    }
    lds4(trow - 4, s + 4);
    lds4(trow, s + 8);
    int v[N];
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = s[8 - H + k];
    if (!filter_class<C>(v, limit << sh, blimit << sh, thresh << sh, bd)) continue;
#pragma unroll
    for (int k = 0; k < N; k++) s[8 - H + k] = v[k];
    if (C <= 1) {
      lds_st2(trow - 2, s[6], s[7]);
      lds_st2(trow, s[8], s[9]);
    } else {
      if (C == 3) {
        lds_st4(trow - 8, s);
        lds_st4(trow + 4, s + 12);
      }
      lds_st4(trow - 4, s + 4);
      lds_st4(trow, s + 8);
    }
  }
}

/* One horizontal-edge class: a column of an edge per item. */
template <int C>
__device__ __forceinline__ void fused_h(LfFusedShared &sm, int bd, int sharp) {
  constexpr int N = LfClass<C>::N, H = N / 2;
  constexpr int R = C <= 1 ? 2 : (C == 2 ? 3 : 6); /* rows the filter modifies either side */
  const int sh = bd - 8, n = sm.count[4 + C];
  for (int it = threadIdx.x; it < 4 * n; it += 256) {
    const int li = class_item<C, FH_N>(sm.lh, it >> 2);
    int limit, blimit, thresh;
    level_limits(sm.elh[li], sharp, &limit, &blimit, &thresh);
    const int row = 8 + 4 * (li / FH_COLS), col = 8 + 4 * (li % FH_COLS) + (it & 3);
    uint16_t *tcol = sm.tile + row * FT + col;
    int v[N];
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = tcol[(k - H) * FT];
    if (!filter_class<C>(v, limit << sh, blimit << sh, thresh << sh, bd)) continue;
#pragma unroll
    for (int k = H - R; k < H + R; k++) tcol[(k - H) * FT] = (uint16_t)v[k];
  }
}

template <typename T>
__device__ __forceinline__ void lf_fused_block(const LfLaunch &a, int gp, int lb, const PlanePtr &o0,
                                               const PlanePtr &o1, const PlanePtr &o2, LfFusedShared *sm) {
  const int plane = a.plane_id[gp];
  const int bx = lb % a.blk_x[gp], by = lb / a.blk_x[gp];
  const int X0 = 64 * bx, Y0 = 64 * by;
  const Geom &g = a.g;
  const int ssx = plane ? g.ssx : 0, ssy = plane ? g.ssy : 0;
  const uint32_t tword = threadIdx.x < LF_TWORDS ? reinterpret_cast<const uint32_t *>(&a.t)[threadIdx.x] : 0u;
  /* the side info at this thread's positions (mode info of the position's
   * 4x4 block, its transform size), loaded with the samples */
  const int ex0 = X0 / 4 - 1, ey0 = Y0 / 4 - 2;
  const int ms = a.s.mi_stride;
  uint2 pmi[FM_ITEMS];
  int ptx[FM_ITEMS];
#pragma unroll
  for (int k = 0; k < FM_ITEMS; k++) {
    const int i = threadIdx.x + 256 * k;
    const int ex = ex0 + i % FM_COLS, ey = ey0 + i / FM_COLS;
    pmi[k] = make_uint2(0, 0);
    ptx[k] = 0;
    /* positions outside the mode-info area are read by no filtered edge */
    if (i < FM_N && ex >= 0 && ey >= 0 && (ex << ssx) < g.mi_cols && (ey << ssy) < g.mi_rows) {
      const int rowo = (ey << ssy) | ssy, colo = (ex << ssx) | ssx;
      pmi[k] = gptr(reinterpret_cast<const uint2 *>(a.s.mi))[rowo * ms + colo];
      ptx[k] = gptr(a.s.tx[plane])[ey * ms + ex];
    }
  }
  const T *in = gptr(reinterpret_cast<const T *>(a.planes[plane].data));
  const int64_t stride = a.planes[plane].stride;
  const int area_w = (g.mi_cols * 4) >> ssx, area_h = (g.mi_rows * 4) >> ssy;
  /* staged rows Y0 - 8 .. Y0 + 71, columns X0 - 8 .. X0 + 71; a filtered
   * edge reads rows below area_h + 8 and columns below area_w + 8 (the
   * margin every deblocked input plane has); chunks past them are loaded
   * from a clamped position and never read */
  constexpr int CW = FT / 4, CHUNKS = FT * CW, PER = (CHUNKS + 255) / 256;
  typename Io<T>::W v[PER];
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int ch = min((int)threadIdx.x + 256 * k, CHUNKS - 1);
    const int r = ch / CW, c = (ch - r * CW) * 4;
    const int y = min(max(Y0 - 8 + r, 0), area_h + 7), x = min(max(X0 - 8 + c, 0), area_w + 4);
    v[k] = Io<T>::ldw(in + (int64_t)y * stride + x);
  }
  if (threadIdx.x < LF_TWORDS) reinterpret_cast<uint32_t *>(&sm->t)[threadIdx.x] = tword;
  if (threadIdx.x < 8) sm->count[threadIdx.x] = 0;
#pragma unroll
  for (int k = 0; k < PER; k++) {
    const int ch = threadIdx.x + 256 * k;
    if (ch < CHUNKS) {
      int o[4];
      Io<T>::unpack(v[k], o);
      lds_st4(&sm->tile[ch * 4], o);
    }
  }
  __syncthreads();
  /* the level of each position's block, per pass (luma: the vertical and
   * horizontal levels; chroma: the plane's) */
#pragma unroll
  for (int k = 0; k < FM_ITEMS; k++) {
    const int i = threadIdx.x + 256 * k;
    if (i < FM_N) {
      const rvpf_mi m = __builtin_bit_cast(rvpf_mi, pmi[k]);
      const int l0 = lf_level(a.h, sm->t, m, plane == 0 ? 0 : plane + 1);
      sm->mi[i] = m;
      sm->tx[i] = (uint8_t)ptx[k];
      sm->lvl[0][i] = (uint8_t)l0;
      sm->lvl[1][i] = (uint8_t)(plane == 0 ? lf_level(a.h, sm->t, m, 1) : l0);
    }
  }
  __syncthreads();
  /* decisions (item < FV_N: vertical edge (column i, row group j); else
   * horizontal), appended wave by wave to the 8 lists (pass, class) */
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int k = 0; k < F_ITEMS; k++) {
    const int item = threadIdx.x + 256 * k;
    int cls = -1, idx = 0, lvl = 0;
    if (item < FV_N + FH_N) {
      const int pass = item < FV_N ? 0 : 1;
      idx = pass ? item - FV_N : item;
      int ex, ey, pos, ppos;
      if (pass == 0) {
        const int i = idx % FV_COLS, j = idx / FV_COLS;
        ex = X0 / 4 + i;
        ey = Y0 / 4 - 2 + j;
        pos = j * FM_COLS + i + 1;
        ppos = pos - 1;
      } else {
        const int i = idx % FH_COLS, j = idx / FH_COLS;
        ex = X0 / 4 + i;
        ey = Y0 / 4 + j;
        pos = (j + 2) * FM_COLS + i + 1;
        ppos = pos - FM_COLS;
      }
      const bool valid = ey >= 0 && (ey << ssy) < g.mi_rows && (ex << ssx) < g.mi_cols && (pass == 0 ? ex != 0 : ey != 0);
      if (valid) {
        const int fs = edge_decision_staged(g, sm->t, sm->mi[pos], sm->tx[pos], sm->tx[ppos], sm->lvl[pass][pos],
                                            sm->lvl[pass][ppos], plane, pass, ex, ey, &lvl);
        if (fs) cls = 4 * pass + (fs == 4 ? 0 : (fs == 16 ? 3 : (plane ? 1 : 2)));
      }
    }
    uint64_t m[8];
#pragma unroll
    for (int c = 0; c < 8; c++) m[c] = __ballot(cls == c);
    uint64_t mk = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) mk = lane == c ? m[c] : mk;
    int base = 0;
    if (lane < 8 && mk != 0) base = atomicAdd(&sm->count[lane], __popcll(mk));
    uint64_t mc = 0;
#pragma unroll
    for (int c = 0; c < 8; c++) mc = cls == c ? m[c] : mc;
    const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(mc >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mc, 0u));
    base = __shfl(base, cls < 0 ? 0 : cls, 64);
    if (cls >= 0) {
      if (cls < 4) {
        class_slot<FV_N>(sm->lv, cls, base + below) = (uint16_t)idx;
        sm->elv[idx] = (uint8_t)lvl;
      } else {
        class_slot<FH_N>(sm->lh, cls - 4, base + below) = (uint16_t)idx;
        sm->elh[idx] = (uint8_t)lvl;
      }
    }
  }
  __syncthreads();
  const int sharp = a.h.sharpness;
  fused_v<0>(*sm, g.bd, sharp);
  if (plane) {
    fused_v<1>(*sm, g.bd, sharp);
  } else {
    fused_v<2>(*sm, g.bd, sharp);
    fused_v<3>(*sm, g.bd, sharp);
  }
  __syncthreads();
  fused_h<0>(*sm, g.bd, sharp);
  if (plane) {
    fused_h<1>(*sm, g.bd, sharp);
  } else {
    fused_h<2>(*sm, g.bd, sharp);
    fused_h<3>(*sm, g.bd, sharp);
  }
  __syncthreads();
  /* the block's 64 x 64 samples, clipped to the mode-info area (a chunk
   * that starts inside it is written whole: output rows are padded) */
  const PlanePtr &op = plane == 0 ? o0 : (plane == 1 ? o1 : o2);
  T *o = gptr(reinterpret_cast<T *>(op.data));
  const int64_t ostride = op.stride;
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int ch = threadIdx.x + 256 * k;
    const int r = ch >> 4, c = (ch & 15) * 4;
    if (Y0 + r < area_h && X0 + c < area_w) {
      int s4[4];
      lds4(&sm->tile[(8 + r) * FT + 8 + c], s4);
      Io<T>::st4(o + (int64_t)(Y0 + r) * ostride + X0 + c, s4);
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void lf_fused_kernel(LfLaunch a, PlanePtr o0, PlanePtr o1, PlanePtr o2) {
  __shared__ LfFusedShared sm;
  int lb;
  const int gp = xcd_segment(blockIdx.x, a.blk_start, a.nseg, &lb);
  if (gp >= a.nseg || lb >= a.blk_n[gp]) return; /* padding */
  lf_fused_block<T>(a, gp, lb, o0, o1, o2, &sm);
}

/* Batched: item j of a frame is block j of its (unpadded) grid. */
template <typename T, int PASS, int EX, bool DIRECT>
__global__ __launch_bounds__(256) void lf_pass_batch(BatchWork w) {
  __shared__ LfShared<PASS, EX> tsm;
  batch_loop<DIRECT>(w, [&](int f, int j) {
    const LfLaunch &a = batch_table<LfLaunch>(w.rec, f);
    int gp = 0;
    while (gp + 1 < a.nseg && j >= a.blk_start[gp + 1]) gp++;
    const int lb = j - a.blk_start[gp];
    if (lb < a.blk_n[gp]) lf_block<T, PASS, EX>(a, gp, lb, &tsm);
  });
}

int log2i(int v) {
  int l = 0;
  while ((1 << l) < v) l++;
  return l;
}

void fill_tables(const Geom &g, const LfHdr &h, LfTables &t) {
  memset(&t, 0, sizeof t);
  for (int tx = 0; tx < 19; tx++) {
    t.tx_log2[0][tx] = (uint8_t)log2i(AV1_TX_WIDTH[tx]);
    t.tx_log2[1][tx] = (uint8_t)log2i(AV1_TX_HEIGHT[tx]);
  }
  for (int c = 0; c < 2; c++)
    for (int b = 0; b < 22; b++) {
      const int ps = AV1_SUBSAMPLED_SIZE[b][c ? g.ssx : 0][c ? g.ssy : 0];
      t.blk_log2[c][0][b] = ps < 22 ? (uint8_t)log2i(4 * AV1_NUM_4X4_BLOCKS_WIDE[ps]) : 0xff;
      t.blk_log2[c][1][b] = ps < 22 ? (uint8_t)log2i(4 * AV1_NUM_4X4_BLOCKS_HIGH[ps]) : 0xff;
    }
  for (int i = 0; i < 8; i++) {
    t.ref_deltas[i] = (int8_t)h.ref_deltas[i];
    t.seg_lf_enabled[i] = (uint8_t)h.seg_lf_enabled[i];
    for (int j = 0; j < 4; j++) t.seg_lf_data[i][j] = (int8_t)h.seg_lf_data[i][j];
  }
  for (int i = 0; i < 2; i++) t.mode_deltas[i] = (int8_t)h.mode_deltas[i];
  for (int i = 0; i < 4; i++) t.level[i] = (uint8_t)h.level[i];
}

/* Fills the launch record; blocks of EX x (256 / EX) edges, each plane's
 * first block padded to a multiple of 8 when pad8 (per-frame launch with
 * the XCD segment mapping). Returns the number of blocks. */
int plan(const Geom &g, const LfHdr &h, const LfSide &s, const PlanePtr planes[3], int ex, bool pad8,
         LfLaunch &a) {
  a.g = g;
  a.h = h;
  a.s = s;
  fill_tables(g, h, a.t);
  int n = 0, total = 0;
  const int ey = 256 / ex;
  for (int p = 0; p < 3; p++) a.planes[p] = planes[p];
  for (int p = 0; p < g.nplanes; p++) {
    if (!(p == 0 || h.level[1 + p] != 0)) continue;
    const int ssx = p ? g.ssx : 0, ssy = p ? g.ssy : 0;
    const int cols4 = (g.mi_cols + (1 << ssx) - 1) >> ssx;
    const int rows4 = (g.mi_rows + (1 << ssy) - 1) >> ssy;
    a.plane_id[n] = p;
    a.blk_x[n] = (cols4 + ex - 1) / ex;
    a.blk_n[n] = a.blk_x[n] * ((rows4 + ey - 1) / ey);
    a.blk_start[n] = total;
    total += pad8 ? (a.blk_n[n] + 7) & ~7 : a.blk_n[n];
    n++;
  }
  a.nseg = n;
  for (int i = n; i < 3; i++) {
    a.plane_id[i] = 0;
    a.blk_x[i] = 1;
    a.blk_n[i] = 0;
    a.blk_start[i] = total;
  }
  a.blk_start[3] = total;
  return total;
}

constexpr int SHAPE_EX[LF_SHAPES] = {64, 32, 16};

/* Per-frame blocks: 16 x 16 edges (64 x 64 samples), whose 8-sample
 * borders add a quarter to the staged samples in either pass (64 x 4
 * edges add a sixteenth across and twice the rows down). */
constexpr int LF_EX = 16;

template <typename T, int EX>
void launch_batch_t(const BatchWork &wv, const BatchWork &wh, const StagePlan &sp, hipStream_t st) {
  if (wv.total > 0) {
    if (batch_direct(wv, sp.grid))
      lf_pass_batch<T, 0, EX, true><<<sp.grid, 256, 0, st>>>(wv);
    else
      lf_pass_batch<T, 0, EX, false><<<sp.grid, 256, 0, st>>>(wv);
  }
  if (wh.total > 0) {
    if (batch_direct(wh, sp.grid))
      lf_pass_batch<T, 1, EX, true><<<sp.grid, 256, 0, st>>>(wh);
    else
      lf_pass_batch<T, 1, EX, false><<<sp.grid, 256, 0, st>>>(wh);
  }
}

template <typename T>
void launch_batch_shape(const BatchWork &wv, const BatchWork &wh, const StagePlan &sp, hipStream_t st) {
  switch (sp.shape) {
    case 1: launch_batch_t<T, 32>(wv, wh, sp, st); break;
    case 2: launch_batch_t<T, 16>(wv, wh, sp, st); break;
    default: launch_batch_t<T, 64>(wv, wh, sp, st); break;
  }
}

}  // namespace

size_t lf_record_size() { return sizeof(LfLaunch); }

int lf_plan(const Geom &g, const LfHdr &h, const LfSide &s, const PlanePtr planes[3], int shape, void *rec) {
  return plan(g, h, s, planes, SHAPE_EX[shape < 0 || shape >= LF_SHAPES ? 0 : shape], false,
              *static_cast<LfLaunch *>(rec));
}

int launch_deblock_batch(const BatchWork &wv, const BatchWork &wh, const StagePlan &sp, int bps, hipStream_t st) {
  if (bps == 1)
    launch_batch_shape<uint8_t>(wv, wh, sp, st);
  else
    launch_batch_shape<uint16_t>(wv, wh, sp, st);
  return (int)hipGetLastError();
}

int launch_deblock_fused(const Geom &g, const LfHdr &h, const LfSide &s, const PlanePtr in[3], const PlanePtr out[3],
                         int bps, hipStream_t st) {
  LfLaunch a;
  const int total = plan(g, h, s, in, LF_EX, true, a);
  if (total == 0) return 0;
  if (bps == 1)
    lf_fused_kernel<uint8_t><<<total, 256, 0, st>>>(a, out[0], out[1], out[2]);
  else
    lf_fused_kernel<uint16_t><<<total, 256, 0, st>>>(a, out[0], out[1], out[2]);
  return (int)hipGetLastError();
}

int launch_deblock(const Geom &g, const LfHdr &h, const LfSide &s, PlanePtr planes[3], int bps,
                   hipStream_t st) {
  LfLaunch a;
  const int total = plan(g, h, s, planes, LF_EX, true, a);
  if (total == 0) return 0;
  if (bps == 1) {
    lf_pass_kernel<uint8_t, 0, LF_EX><<<total, 256, 0, st>>>(a);
    lf_pass_kernel<uint8_t, 1, LF_EX><<<total, 256, 0, st>>>(a);
  } else {
    lf_pass_kernel<uint16_t, 0, LF_EX><<<total, 256, 0, st>>>(a);
    lf_pass_kernel<uint16_t, 1, LF_EX><<<total, 256, 0, st>>>(a);
  }
  return (int)hipGetLastError();
}

}  // namespace rvpf
