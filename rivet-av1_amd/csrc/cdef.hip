/*
 * cdef.hip — CDEF (AV1 7.15) for gfx950.
 *
 * Follows rivet-av1's cdef / cdef_block / cdef_direction /
 * cdef_filter_block (src/decoder/postfilter.rs:273-513) and
 * dsp::cdef::filter_scalar (src/dsp/cdef.rs:48-90).
 *
 * One workgroup (256 lanes, 4 waves) per 64x64 filter block (one
 * cdef_idx):
 *   1. the luma tile plus a 2-sample border is staged in LDS; samples at or
 *      beyond the mode-info edge are the sentinel NA_S (unavailable); one
 *      wave lists the 8x8 blocks CDEF filters, then the skipped ones, and
 *      the four lanes of slot s take block s of that list, so the waves
 *      past the filtered blocks only copy theirs;
 *   2. the four lanes of an 8x8 block run its direction search, each
 *      through one symmetry of the block (no divergence), and one shuffle
 *      gathers the eight costs;
 *   3. the same four lanes filter that block (two luma columns each), so
 *      taps and strengths are set up once per lane; a wave's 16 blocks
 *      span two block rows, and its stores cover whole 64-sample rows.
 *      Each chroma plane is then staged and filtered the same way.
 * Unfiltered blocks (cdef_idx -1, all four 4x4 skipped, CDEF off) are
 * copied, so `out` receives the whole mode-info area.
 *
 * Unavailable taps: NA_S = 0xA000, i.e. -24576 as a signed and 40960 as an
 * unsigned 16-bit value. Its difference to any sample (0..4095) lies in
 * [-28671, -24576] without wrapping, so constrain() of it is 0 for every
 * legal strength/damping (the magnitude test |d| >> adj > strength holds
 * for adj <= 10); a signed max() and an unsigned min() never pick it. The
 * reference skips such taps, which therefore gives the same result.
 */
#include "av1_tables.h"
#include "rvpf_internal.h"

namespace rvpf {

namespace {

constexpr int NA_S = 0xA000;
constexpr int TS = 68; /* tile stride in samples: 64 + 2 * 2 */
/* word offset of the shifted tile S behind a tile P of `rows` rows plus
 * the 2-row borders */
constexpr int soff_rows(int rows) { return (rows + 4) * (TS / 2); }
constexpr int SOFF = soff_rows(64);

struct CdefLaunch {
  Geom g;
  CdefHdr h;
  const rvpf_mi *mi;
  int mi_stride;
  const int8_t *cdef_idx;
  int cdef_stride;
  PlanePtr in[3], out[3];
  int enabled;
  int fb_cols;
  int vec; /* every input row is 16-byte (u16) or 8-byte (u8) aligned: chunked staging */
};

template <typename T>
__device__ __forceinline__ T *pp(const PlanePtr &p) {
  return gptr(reinterpret_cast<T *>(p.data));
}


/* Stages rows [y0-2, y0+h+2) x cols [x0-2, x0+w+2) of a plane into the
 * u16 tile P (stride TS); outside [0, xlim) x [0, ylim) -> NA_S. The same
 * samples shifted left by one go to S (S[r][c] = P[r][c + 1]), so that any
 * horizontal pair (c + dx, c + dx + 1) with c even is one aligned 32-bit
 * word of P (dx even) or of S (dx odd). w <= 64. */
template <typename T, int NW>
__device__ void stage_tile(uint16_t *t, uint16_t *sh, const T *src, int64_t stride, int x0, int y0, int w,
                           int h, int xlim, int ylim) {
  const int tw = w + 4, th = h + 4;
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int xa = x0 + tx - 2, xb = xa + 64;
  const bool xa_ok = xa >= 0 && xa < xlim, xb_ok = tx + 64 < tw && xb < xlim;
  const int xac = min(max(xa, 0), xlim - 1), xbc = min(xb, xlim - 1);
  /* ROWS rows per batch: every load (clamped into the plane) is issued
   * before the LDS stores */
  constexpr int ROWS = 6;
  for (int r0 = ty; r0 < th; r0 += NW * ROWS) {
    uint16_t va[ROWS], vb[ROWS];
#pragma unroll
    for (int k = 0; k < ROWS; k++) {
      const int y = y0 + r0 + NW * k - 2;
      const bool yok = y >= 0 && y < ylim;
      const T *row = src + (int64_t)min(max(y, 0), ylim - 1) * stride;
      const int a = row[xac], b = row[xbc];
      va[k] = (uint16_t)(yok && xa_ok ? a : NA_S);
      vb[k] = (uint16_t)(yok && xb_ok ? b : NA_S);
    }
#pragma unroll
    for (int k = 0; k < ROWS; k++) {
      const int r = r0 + NW * k;
      if (r < th) {
        if (tx < tw) {
          t[r * TS + tx] = va[k];
          if (tx > 0) sh[r * TS + tx - 1] = va[k];
        }
        if (tx + 64 < tw) {
          t[r * TS + tx + 64] = vb[k];
          sh[r * TS + tx + 63] = vb[k];
        }
        if (tx == 0) sh[r * TS + tw - 1] = (uint16_t)NA_S;
      }
    }
  }
}

/* Chunked staging, for planes whose rows are 16-byte (u16) or 8-byte (u8)
 * aligned. Item (r, k) of a tile with NCH chunks per row covers the 8
 * samples x0 - 8 + 8k .. x0 - 1 + 8k of tile row r: P words 4k - 3 .. 4k
 * (tile columns 2m, 2m + 1 for word m), loaded as one access, and the
 * next chunk's first word, from which the S words (S[m] = samples 2m + 1,
 * 2m + 2) are aligned shifts. Loads read clamped addresses inside
 * [0, xlim) (xlim >= 8) and unavailable words become NA_S pairs. x0 is a
 * multiple of 8 and xlim of 4 (8 for luma): a chunk lies wholly inside,
 * wholly outside, or (chroma) its first half inside, read from the access
 * at xlim - 8. */
constexpr uint32_t NA_W = (uint32_t)NA_S * 0x10001u;

template <typename T>
__device__ __forceinline__ uint4 load_chunk(const T *p);
template <>
__device__ __forceinline__ uint4 load_chunk<uint16_t>(const uint16_t *p) {
  return *reinterpret_cast<const uint4 *>(p);
}
template <>
__device__ __forceinline__ uint4 load_chunk<uint8_t>(const uint8_t *p) {
  const uint2 v = *reinterpret_cast<const uint2 *>(p);
  return make_uint4(__builtin_amdgcn_perm(0u, v.x, 0x0c010c00u), __builtin_amdgcn_perm(0u, v.x, 0x0c030c02u),
                    __builtin_amdgcn_perm(0u, v.y, 0x0c010c00u), __builtin_amdgcn_perm(0u, v.y, 0x0c030c02u));
}
/* the first two samples at p as one word */
template <typename T>
__device__ __forceinline__ uint32_t load_word(const T *p);
template <>
__device__ __forceinline__ uint32_t load_word<uint16_t>(const uint16_t *p) {
  return *reinterpret_cast<const uint32_t *>(p);
}
template <>
__device__ __forceinline__ uint32_t load_word<uint8_t>(const uint8_t *p) {
  return __builtin_amdgcn_perm(0u, (uint32_t)*reinterpret_cast<const uint16_t *>(p), 0x0c010c00u);
}

/* A chunk as loaded (every address clamped into the plane, no branch, so
 * that nothing waits for it before its store); chunk_store applies the
 * availability. */
struct Chunk {
  uint4 c;    /* P words 4k - 3 .. 4k */
  uint32_t n; /* word 4k + 1 */
};

template <typename T, int NCH>
__device__ __forceinline__ Chunk chunk_load(const T *src, int64_t stride, int x0, int y0, int it, int xlim,
                                            int ylim) {
  const int r = it / NCH, k = it - r * NCH;
  const int y = y0 - 2 + r, x = x0 - 8 + 8 * k;
  const T *row = src + (int64_t)min(max(y, 0), ylim - 1) * stride;
  Chunk ch;
  ch.c = load_chunk<T>(row + min(max(x, 0), xlim - 8));
  ch.n = load_word<T>(row + min(x + 8, xlim - 2));
  return ch;
}

template <int NCH, int RW = TS / 2>
__device__ __forceinline__ void chunk_store(const Chunk &ch, uint32_t *pw, uint32_t *sw, int it, int x0, int y0,
                                            int xlim, int ylim) {
  const int r = it / NCH, k = it - r * NCH;
  const int y = y0 - 2 + r, x = x0 - 8 + 8 * k;
  const bool yok = y >= 0 && y < ylim;
  const bool ok = yok && x >= 0 && x + 8 <= xlim, half = yok && x + 4 == xlim, ok1 = yok && x + 8 < xlim;
  const uint32_t w[5] = {ok ? ch.c.x : (half ? ch.c.z : NA_W), ok ? ch.c.y : (half ? ch.c.w : NA_W),
                         ok ? ch.c.z : NA_W, ok ? ch.c.w : NA_W, ok1 ? ch.n : NA_W};
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const int m = 4 * k - 3 + j;
    if (m >= 0 && m < RW) {
      pw[r * RW + m] = w[j];
      sw[r * RW + m] = __builtin_amdgcn_alignbit(w[j + 1], w[j], 16);
    }
  }
}

/* A 4:2:0 chroma tile held beside the luma tile (chunked staging): 32 + 4
 * samples per row. */
constexpr int TSC = 36;
constexpr int csoff_rows(int rows) { return (rows + 4) * (TSC / 2); }

/* stage_tile (w = 64, or 32 with NCH = 6) from chunks: every load of a
 * thread's items before its LDS stores */
template <typename T, int NW, int NCH>
__device__ __forceinline__ void stage_chunks(uint32_t *pw, uint32_t *sw, const T *src, int64_t stride, int x0,
                                             int y0, int th, int xlim, int ylim) {
  constexpr int NT = 64 * NW, MAXI = ((16 * NW + 4) * NCH + NT - 1) / NT;
  Chunk c[MAXI];
#pragma unroll
  for (int i = 0; i < MAXI; i++) c[i] = chunk_load<T, NCH>(src, stride, x0, y0, threadIdx.x + i * NT, xlim, ylim);
#pragma unroll
  for (int i = 0; i < MAXI; i++) {
    const int it = threadIdx.x + i * NT;
    if (it < th * NCH) chunk_store<NCH>(c[i], pw, sw, it, x0, y0, xlim, ylim);
  }
}

/* A chroma tile of a 4:2:0 frame ((32 + 4) x (8 NW + 4) samples, the
 * chroma of an NW-wave tile) loaded into registers, CR = 8 + 4 / NW rows
 * per thread (rows ty, ty + NW, ...; column tx < 36), with the unavailable
 * samples already NA_S, so that its loads can be in flight while the luma
 * tile filters; chroma_store writes it to P and S. */
struct ChromaRegs {
  uint32_t v[5]; /* row k in half k & 1 of word k >> 1 */
};

template <int NW>
constexpr int chroma_rows() { return 8 + 4 / NW; }

template <typename T, int NW>
__device__ __forceinline__ void chroma_load(ChromaRegs &cr, const T *src, int64_t stride, int x0, int y0, int xlim,
                                            int ylim) {
  constexpr int CR = chroma_rows<NW>();
  static_assert(CR <= 10, "ChromaRegs holds 10 rows");
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  const int xa = x0 + tx - 2;
  const bool xa_ok = tx < 36 && xa >= 0 && xa < xlim;
  const int xac = min(max(xa, 0), xlim - 1);
  uint32_t v[CR];
#pragma unroll
  for (int k = 0; k < CR; k++) {
    const int y = y0 + ty + NW * k - 2;
    const bool yok = y >= 0 && y < ylim;
    const int a = src[(int64_t)min(max(y, 0), ylim - 1) * stride + xac];
    v[k] = yok && xa_ok ? (uint32_t)a : (uint32_t)NA_S;
  }
#pragma unroll
  for (int k = 0; k < (CR + 1) / 2; k++) cr.v[k] = v[2 * k] | (2 * k + 1 < CR ? v[2 * k + 1] << 16 : 0u);
}

template <int NW>
__device__ __forceinline__ void chroma_store(const ChromaRegs &cr, uint16_t *t, uint16_t *sh) {
  constexpr int CR = chroma_rows<NW>();
  const int tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
  if (tx >= 36) return;
#pragma unroll
  for (int k = 0; k < CR; k++) {
    const int r = ty + NW * k;
    const uint16_t v = (uint16_t)(cr.v[k >> 1] >> (16 * (k & 1)));
    t[r * TS + tx] = v;
    if (tx > 0) sh[r * TS + tx - 1] = v;
    if (tx == 0) sh[r * TS + 35] = (uint16_t)NA_S;
  }
}

/* Product of two values known to fit in 16 bits (one 24-bit multiply). */
__device__ __forceinline__ int mul16(int a, int b) { return (int)(int16_t)a * (int)(int16_t)b; }

/* Two horizontally adjacent samples filtered at once in packed 16-bit
 * arithmetic (dsp::cdef::filter_scalar, src/dsp/cdef.rs:61-88). Every
 * quantity fits in int16: samples <= 4095, the sentinel 0xA000, their
 * differences, constrain() results <= 240 and |sum| <= 24 * 240. */
typedef short s16x2 __attribute__((ext_vector_type(2)));
typedef unsigned short u16x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ s16x2 as_s2(uint32_t w) { return __builtin_bit_cast(s16x2, w); }
__device__ __forceinline__ s16x2 splat(int v) { return (s16x2){(short)v, (short)v}; }

/* constrain() of 7.15.3, sign(d) * min(|d|, max(0, t - (|d| >> adj))),
 * of the tap p against the centre x, all unsigned: the positive and
 * negative parts of d = p - x are saturating differences (one of them is
 * 0, so their or is |d|), m = max(0, t - (|d| >> adj)) is a saturating
 * difference too, and the result is min(pos, m) - min(neg, m). An
 * unavailable tap (NA_S = 40960) has |d| >= 36865, which keeps m at 0. */
__device__ __forceinline__ s16x2 constrain2(u16x2 p, u16x2 x, u16x2 t, u16x2 adj, u16x2 &neg) {
  const u16x2 pos = __builtin_elementwise_sub_sat(p, x);
  neg = __builtin_elementwise_sub_sat(x, p);
  const u16x2 m = __builtin_elementwise_sub_sat(t, (pos | neg) >> adj);
  return __builtin_bit_cast(s16x2, __builtin_elementwise_min(pos, m)) -
         __builtin_bit_cast(s16x2, __builtin_elementwise_min(neg, m));
}

/* Packed maximum of pairs of samples (0 .. 4095) or NA_S, read as half
 * floats (v_pk_maximum3_f16): a sample's bit pattern is a non-negative
 * half (normal or subnormal; the kernels keep subnormal halves) whose
 * order is the integer order, and NA_S is a negative half, below every
 * sample. */
typedef _Float16 h16x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ uint32_t max3_h(uint32_t a, uint32_t b, uint32_t c) {
  const h16x2 m = __builtin_elementwise_maximum(__builtin_bit_cast(h16x2, a), __builtin_bit_cast(h16x2, b));
  return __builtin_bit_cast(uint32_t, __builtin_elementwise_maximum(m, __builtin_bit_cast(h16x2, c)));
}

/* Per-lane filter set-up: strengths, damping shifts, primary tap weights
 * and the 32-bit word offsets of the 12 taps relative to the pair's word. */
struct Taps {
  u16x2 pri, sec;
  u16x2 pri_adj, sec_adj;
  s16x2 pt0, pt1;
  int po[4];  /* +d0, -d0, +d1, -d1 */
  int so[8];  /* (lo, hi) x (k0, k1) x (+, -) */
  bool has_pri, has_sec, clamp;
};

__device__ __forceinline__ int floor_log2(int x) { return 31 - __clz(x); }

/* word offset of tap (dy, dx) for a pair at an even column; soff = the
 * word offset of S behind P */
__device__ __forceinline__ int word_off(int dy, int dx, int soff, int rw) {
  return (dx & 1) ? soff + dy * rw + ((dx - 1) >> 1) : dy * rw + (dx >> 1);
}

__device__ __forceinline__ Taps make_taps(int pri, int sec, int damping, int dir, int coeff_shift, int soff,
                                          int rw = TS / 2) {
  Taps k;
  k.has_pri = pri != 0;
  k.has_sec = sec != 0;
  k.clamp = k.has_pri && k.has_sec;
  k.pri = (u16x2){(unsigned short)pri, (unsigned short)pri};
  k.sec = (u16x2){(unsigned short)sec, (unsigned short)sec};
  const int pa = pri ? max(damping - floor_log2(pri), 0) : 0;
  const int sa = sec ? max(damping - floor_log2(sec), 0) : 0;
  k.pri_adj = (u16x2){(unsigned short)pa, (unsigned short)pa};
  k.sec_adj = (u16x2){(unsigned short)sa, (unsigned short)sa};
  const int sel = (pri >> coeff_shift) & 1;
  k.pt0 = splat(AV1_CDEF_PRI_TAPS[sel][0]);
  k.pt1 = splat(AV1_CDEF_PRI_TAPS[sel][1]);
  const int dlo = (dir + 6) & 7, dhi = (dir + 2) & 7;
#pragma unroll
  for (int t = 0; t < 2; t++) {
    const int dy = AV1_CDEF_DIRECTIONS[dir][t][0], dx = AV1_CDEF_DIRECTIONS[dir][t][1];
    k.po[2 * t] = word_off(dy, dx, soff, rw);
    k.po[2 * t + 1] = word_off(-dy, -dx, soff, rw);
    const int ly = AV1_CDEF_DIRECTIONS[dlo][t][0], lx = AV1_CDEF_DIRECTIONS[dlo][t][1];
    const int hy = AV1_CDEF_DIRECTIONS[dhi][t][0], hx = AV1_CDEF_DIRECTIONS[dhi][t][1];
    k.so[4 * t + 0] = word_off(ly, lx, soff, rw);
    k.so[4 * t + 1] = word_off(-ly, -lx, soff, rw);
    k.so[4 * t + 2] = word_off(hy, hx, soff, rw);
    k.so[4 * t + 3] = word_off(-hy, -hx, soff, rw);
  }
  return k;
}

/* The filtered pair whose own word is w[base]. PRI / SEC: the tile has
 * primary / secondary taps at all (a lane's strength of 0 makes its taps
 * contribute nothing, since constrain() with t = 0 is 0). The clamp to
 * [min, max] of the available taps is computed only with both kinds: with
 * one kind the taps' weights total 12 and |constrain(d)| <= |d|, so
 * (8 + sum - (sum < 0)) >> 4 lies between -max(x - p) and max(p - x) and
 * the update stays inside by itself; with no taps it is the identity.
 * Every tap is loaded before the arithmetic, and the taps are
 * independent, so the packed operations interleave. */
/* The filtered pair xw from its loaded taps pw (primary) and sw
 * (secondary). The clamp's bounds: max = the largest of x and the
 * available taps (three at a time, max3_h), min = x - the largest
 * saturating difference x - tap (0 for NA_S), which constrain() computes
 * anyway. */
template <bool PRI, bool SEC>
__device__ __forceinline__ uint32_t filter_pair_taps(uint32_t xw, const uint32_t *pw, const uint32_t *sw,
                                                     const Taps &k) {
  const s16x2 x = as_s2(xw);
  const u16x2 xu = __builtin_bit_cast(u16x2, xw);
  constexpr bool CLAMP = PRI && SEC;
  s16x2 sum = splat(0);
  uint32_t mx = 0, mneg = 0, n3 = 0; /* CLAMP: running maxima, folded in as the taps go */
  const auto w32 = [](u16x2 v) { return __builtin_bit_cast(uint32_t, v); };
  if (PRI) {
    s16x2 c[4];
    u16x2 n[4];
#pragma unroll
    for (int t = 0; t < 4; t++) c[t] = constrain2(__builtin_bit_cast(u16x2, pw[t]), xu, k.pri, k.pri_adj, n[t]);
    if (CLAMP) {
      mx = max3_h(max3_h(xw, pw[0], pw[1]), pw[2], pw[3]);
      mneg = max3_h(w32(n[0]), w32(n[1]), w32(n[2]));
      n3 = w32(n[3]);
    }
    sum = k.pt0 * (c[0] + c[1]) + k.pt1 * (c[2] + c[3]);
  }
  if (SEC) {
    /* Cdef_Sec_Taps is {2, 1} in both rows */
    s16x2 c[8];
    u16x2 n[8];
#pragma unroll
    for (int t = 0; t < 8; t++) {
      c[t] = constrain2(__builtin_bit_cast(u16x2, sw[t]), xu, k.sec, k.sec_adj, n[t]);
      if (CLAMP && (t & 1)) mx = max3_h(mx, sw[t - 1], sw[t]);
    }
    if (CLAMP) {
      mneg = max3_h(mneg, n3, w32(n[0]));
      mneg = max3_h(mneg, w32(n[1]), w32(n[2]));
      mneg = max3_h(mneg, w32(n[3]), w32(n[4]));
      mneg = max3_h(mneg, w32(n[5]), w32(n[6]));
      mneg = max3_h(mneg, w32(n[7]), w32(n[7]));
    }
    const s16x2 s0 = (c[0] + c[1]) + (c[2] + c[3]), s1 = (c[4] + c[5]) + (c[6] + c[7]);
    sum += s0 + s0 + s1;
  }
  /* x + ((8 + sum - (sum < 0)) >> 4), clamped to [min, max] of the taps */
  s16x2 v = x + ((sum + splat(8) + (sum >> splat(15))) >> splat(4));
  if (CLAMP) {
    const s16x2 mn = __builtin_bit_cast(s16x2, xu - __builtin_bit_cast(u16x2, mneg));
    v = __builtin_elementwise_min(__builtin_elementwise_max(v, mn), as_s2(mx));
  }
  return __builtin_bit_cast(uint32_t, v);
}

template <bool PRI, bool SEC>
__device__ __forceinline__ uint32_t filter_pair(const uint32_t *w, int base, const Taps &k) {
  uint32_t pw[4], sw[8];
  if (PRI) {
#pragma unroll
    for (int t = 0; t < 4; t++) pw[t] = w[base + k.po[t]];
  }
  if (SEC) {
#pragma unroll
    for (int t = 0; t < 8; t++) sw[t] = w[base + k.so[t]];
  }
  return filter_pair_taps<PRI, SEC>(w[base], pw, sw, k);
}

/* Three of the direction costs of 7.15.2 (postfilter.rs:345-390) for an
 * 8x8 block read through a symmetry of the square: pixel (i, j) of the
 * view is at tile offset b0 + i * rs + j * cs. On the view the lane sums
 * lines i + j (direction 0's lines), i + j / 2 (direction 1's) and i
 * (direction 2's). Under a horizontal flip these are directions 4, 3, 2;
 * under a transpose 0, 7, 6; under transpose + flip 4, 5, 6 with the line
 * order reversed, which leaves the costs unchanged (both cost formulas are
 * symmetric in the line index). */
__device__ void dir_costs3(const uint16_t *t, int b0, int rs, int cs, int sh, int32_t *c0, int32_t *c1,
                           int32_t *c2) {
  int32_t p0[15], p1[11], p2[8];
#pragma unroll
  for (int i = 0; i < 15; i++) p0[i] = 0;
#pragma unroll
  for (int i = 0; i < 11; i++) p1[i] = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    int row = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int x = (t[b0 + i * rs + j * cs] >> sh) - 128;
      p0[i + j] += x;
      p1[i + j / 2] += x;
      row += x;
    }
    p2[i] = row;
  }
  /* |partial| <= 1024: squares and the 840/k weights stay within 24-bit
   * multiplies, their products within 32 bits */
  int32_t a = __mul24(mul16(p0[7], p0[7]), AV1_DIV_TABLE[8]);
#pragma unroll
  for (int i = 0; i < 7; i++)
    a += __mul24(mul16(p0[i], p0[i]) + mul16(p0[14 - i], p0[14 - i]), AV1_DIV_TABLE[i + 1]);
  int32_t b = 0;
#pragma unroll
  for (int j = 0; j < 5; j++) b += mul16(p1[3 + j], p1[3 + j]);
  b = __mul24(b, AV1_DIV_TABLE[8]);
#pragma unroll
  for (int j = 0; j < 3; j++)
    b += __mul24(mul16(p1[j], p1[j]) + mul16(p1[10 - j], p1[10 - j]), AV1_DIV_TABLE[2 * j + 2]);
  int32_t c = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) c += mul16(p2[i], p2[i]);
  *c0 = a;
  *c1 = b;
  *c2 = __mul24(c, AV1_DIV_TABLE[8]);
}

/* Direction and variance (7.15.2) of the 8x8 block of lane tid (block
 * tid / 4 of the tile at block row by, column bx): its four lanes each run
 * dir_costs3 through one symmetry of the block, and one shuffle gathers
 * the eight costs; the first largest wins. Lanes whose block is not
 * filtered get direction 0 and variance 0 (all four lanes of a block
 * agree on filt). */
__device__ __forceinline__ void block_direction(const uint16_t *tile, int tid, bool filt, int by, int bx, int cs,
                                                int *ydir, int32_t *var) {
  const int q = tid & 3;
  int32_t c0 = 0, c1 = 0, c2 = 0;
  if (filt) {
    const int base = (by * 8 + 2) * TS + bx * 8 + 2;
    /* q: 0 identity, 1 horizontal flip, 2 transpose, 3 transpose + flip */
    const int rs = q == 0 || q == 1 ? TS : (q == 2 ? 1 : -1);
    const int csx = q == 0 ? 1 : (q == 1 ? -1 : TS);
    const int b0 = base + (q == 1 || q == 3 ? 7 : 0);
    dir_costs3(tile, b0, rs, csx, cs, &c0, &c1, &c2);
  }
  /* lane q of the block: q0 -> costs 0,1,2; q1 -> 4,3,2; q2 -> 0,7,6; q3 -> 4,5,6 */
  const int lb = (tid & 63) & ~3;
  const int32_t k0 = __shfl(c0, lb + 0, 64), k1 = __shfl(c1, lb + 0, 64), k2 = __shfl(c2, lb + 0, 64);
  const int32_t k4 = __shfl(c0, lb + 1, 64), k3 = __shfl(c1, lb + 1, 64);
  const int32_t k7 = __shfl(c1, lb + 2, 64), k6 = __shfl(c2, lb + 2, 64);
  const int32_t k5 = __shfl(c1, lb + 3, 64);
  *ydir = 0;
  *var = 0;
  if (filt) {
    const int32_t cost[8] = {k0, k1, k2, k3, k4, k5, k6, k7};
    int32_t best = 0;
    int d = 0;
#pragma unroll
    for (int k = 0; k < 8; k++)
      if (cost[k] > best) {
        best = cost[k];
        d = k;
      }
    *ydir = d;
    *var = (best - cost[(d + 4) & 7]) >> 10;
  }
}

/* The luma primary strength after the variance adjustment of cdef_block
 * (postfilter.rs:318-326); pri is already shifted by coeff_shift. */
__device__ __forceinline__ int adjust_pri(int pri, int32_t var) {
  const int var_str = (var >> 6) != 0 ? min(floor_log2(var >> 6), 12) : 0;
  return var != 0 ? (pri * (4 + var_str) + 8) >> 4 : 0;
}

/* Filters this lane's pairs of one block: rows r0 .. r0 + NROWS - 1, the
 * column pair starting at block column c (even); t0 = the block's top-left
 * tile sample (even), o = the block's top-left output sample. Lanes whose
 * block is not filtered (on == false) store the input pair. The words of
 * the 12 taps on the first row are addressed once; the rows follow at
 * constant offsets (immediates of the LDS reads). */
template <typename T, bool PRI, bool SEC, int NROWS, int TSX>
__device__ __forceinline__ void filter_block_rows(const uint32_t *w, int t0, T *o, int64_t os, int c, int r0,
                                                  const Taps &k, bool on) {
  constexpr int RW = TSX / 2; /* words per tile row */
  o += (int64_t)r0 * os + c;
  const uint32_t *b = w + ((t0 + r0 * TSX + c) >> 1);
  const uint32_t *pt[4], *st[8];
#pragma unroll
  for (int t = 0; t < 4; t++) pt[t] = b + k.po[t];
#pragma unroll
  for (int t = 0; t < 8; t++) st[t] = b + k.so[t];
#pragma clang loop unroll(full)
  for (int r = 0; r < NROWS; r++, o += os) {
    uint32_t pw[4], sw[8];
    if (PRI) {
#pragma unroll
      for (int t = 0; t < 4; t++) pw[t] = pt[t][r * RW];
    }
    if (SEC) {
#pragma unroll
      for (int t = 0; t < 8; t++) sw[t] = st[t][r * RW];
    }
    const uint32_t x = b[r * RW];
    uint32_t v = x;
    if (PRI || SEC) {
      const uint32_t f = filter_pair_taps<PRI, SEC>(x, pw, sw, k); /* branch-free: every lane filters */
      v = on ? f : x;
    }
    if (sizeof(T) == 2) {
      *reinterpret_cast<uint32_t *>(o) = v;
    } else {
      *reinterpret_cast<uint16_t *>(o) = (uint16_t)((v & 0xff) | ((v >> 8) & 0xff00));
    }
    /* the chroma tiles beside the luma tile: one row at a time (the
     * scheduler would interleave the two rows of a chroma lane and take
     * 105 registers) */
    if (TSX == TSC) __builtin_amdgcn_sched_barrier(0);
  }
}

template <typename T, bool PRI, bool SEC, int TSX>
__device__ __forceinline__ void filter_block_pairs(const uint32_t *w, int t0, T *o, int64_t os, int c, int r0,
                                                   int nrows, const Taps &k, bool on) {
  switch (nrows) {
    case 8: filter_block_rows<T, PRI, SEC, 8, TSX>(w, t0, o, os, c, r0, k, on); break;
    case 4: filter_block_rows<T, PRI, SEC, 4, TSX>(w, t0, o, os, c, r0, k, on); break;
    default: filter_block_rows<T, PRI, SEC, 2, TSX>(w, t0, o, os, c, r0, k, on); break;
  }
}

/* filter_block_pairs with the tile's tap kinds as template arguments. */
template <typename T, int TSX = TS>
__device__ __forceinline__ void filter_block_any(bool pri, bool sec, const uint32_t *w, int t0, T *o, int64_t os,
                                                 int c, int r0, int nrows, const Taps &k, bool on) {
  if (pri && sec)
    filter_block_pairs<T, true, true, TSX>(w, t0, o, os, c, r0, nrows, k, on);
  else if (pri)
    filter_block_pairs<T, true, false, TSX>(w, t0, o, os, c, r0, nrows, k, on);
  else if (sec)
    filter_block_pairs<T, false, true, TSX>(w, t0, o, os, c, r0, nrows, k, on);
  else
    filter_block_pairs<T, false, false, TSX>(w, t0, o, os, c, r0, nrows, k, false);
}

/* The 8x8 blocks of a filter block in the order lanes take them: the
 * blocks CDEF filters first (nfilt of them), then the ones it copies
 * (ncopy: all four 4x4 skipped); blocks past the mode-info edge are in
 * neither. */
struct CdefBlocks {
  uint8_t list[64];
  int nfilt, ncopy;
};

/* One tile of NW waves: the 64 x 16 NW luma samples (and their chroma)
 * of part tile_id % (4 / NW) of the 64x64 filter block tile_id / (4 / NW)
 * (raster order); words holds the tile P, then the shifted tile S. The
 * four lanes of slot s = tid / 4 take block list[s]: only the waves whose
 * slots hold filtered blocks run the direction search and the filters,
 * the rest copy their blocks from the staged tile. */
template <typename T, int NW, bool VEC>
__device__ __forceinline__ void cdef_tile(const CdefLaunch &a, int tile_id, uint32_t *words, CdefBlocks &bl) {
  constexpr int TH = 16 * NW;      /* luma rows of the tile */
  constexpr int NPART = 64 / TH;   /* tiles per filter block */
  constexpr int SO = soff_rows(TH);
  uint16_t *tile = reinterpret_cast<uint16_t *>(words);
  uint16_t *shifted = reinterpret_cast<uint16_t *>(words + SO);
  const Geom &g = a.g;
  const int tid = threadIdx.x;
  const int tx = tid & 63, ty = tid >> 6;
  const int fb = tile_id / NPART, part = tile_id % NPART;
  const int fbx = fb % a.fb_cols, fby = fb / a.fb_cols;
  const int cs = g.bd - 8;
  const int xlim = g.mi_cols * 4, ylim = g.mi_rows * 4;
  const int x0 = fbx * 64, y0 = fby * 64 + part * TH;
  if (y0 >= ylim) return; /* the lower part of a filter block below the area */
  int idx = -1;
  if (a.enabled) idx = gptr(a.cdef_idx)[(int64_t)fby * a.cdef_stride + fbx];
  if (idx >= 0) idx &= 7; /* cdef_bits <= 3 */
  const int bw = min(64, xlim - x0), bh = min(TH, ylim - y0); /* multiples of 8 */

  if (idx < 0) {
    /* not filtered: copy the block area of every plane */
    for (int p = 0; p < g.nplanes; p++) {
      const int ssx = p ? g.ssx : 0, ssy = p ? g.ssy : 0;
      const int w = bw >> ssx, h = bh >> ssy, px = x0 >> ssx, py = y0 >> ssy;
      if (tx < w) {
        const T *src = pp<T>(a.in[p]) + (int64_t)py * a.in[p].stride + px + tx;
        T *dst = pp<T>(a.out[p]) + (int64_t)py * a.out[p].stride + px + tx;
        /* all loads of a batch (rows clamped to the block) before the stores */
        constexpr int ROWS = 8;
        for (int r0 = ty; r0 < h; r0 += NW * ROWS) {
          T v[ROWS];
#pragma unroll
          for (int k = 0; k < ROWS; k++) v[k] = src[(int64_t)min(r0 + NW * k, h - 1) * a.in[p].stride];
#pragma unroll
          for (int k = 0; k < ROWS; k++)
            if (r0 + NW * k < h) dst[(int64_t)(r0 + NW * k) * a.out[p].stride] = v[k];
        }
      }
    }
    return;
  }

  /* cdef_block's selection (postfilter.rs:304-310): wave 0, a lane per
   * 8x8 block, lists the filtered blocks, then the skipped ones */
  if (tid < 64) {
    const int by = tid >> 3, bx = tid & 7;
    const bool inside = by * 8 < bh && bx * 8 < bw;
    bool skip = true;
    if (inside) {
      const int r = (y0 >> 2) + by * 2, c = (x0 >> 2) + bx * 2;
      const int ms = a.mi_stride;
      const rvpf_mi *mi = gptr(a.mi);
      skip = mi[(int64_t)r * ms + c].seg_skip & mi[(int64_t)(r + 1) * ms + c].seg_skip &
             mi[(int64_t)r * ms + c + 1].seg_skip & mi[(int64_t)(r + 1) * ms + c + 1].seg_skip & 1;
    }
    const bool filt = inside && !skip;
    const uint64_t mf = __ballot(filt), mc = __ballot(inside && skip);
    const int nf = __popcll(mf);
    const uint32_t flo = (uint32_t)mf, fhi = (uint32_t)(mf >> 32), clo = (uint32_t)mc, chi = (uint32_t)(mc >> 32);
    if (filt) bl.list[__builtin_amdgcn_mbcnt_hi(fhi, __builtin_amdgcn_mbcnt_lo(flo, 0u))] = (uint8_t)tid;
    if (inside && skip) bl.list[nf + __builtin_amdgcn_mbcnt_hi(chi, __builtin_amdgcn_mbcnt_lo(clo, 0u))] = (uint8_t)tid;
    if (tid == 0) {
      bl.nfilt = nf;
      bl.ncopy = __popcll(mc);
    }
  }
  /* 4:2:0 chroma: with chunked staging both chroma tiles are staged with
   * the luma tile, into their own LDS beside it (no barrier or restaging
   * between planes); else their loads go out now into registers and land
   * while luma filters */
  const bool c420 = g.nplanes > 1 && g.ssx && g.ssy;
  constexpr int CSO = csoff_rows(8 * NW);
  uint32_t *cwords = words + 2 * SO; /* VEC: Cb P, Cb S, Cr P, Cr S */
  ChromaRegs cpre[VEC ? 1 : 2];
  if constexpr (VEC) {
    /* luma (bh + 4) x 10 chunks, chroma ((bh >> 1) + 4) x 6 per plane (at
     * most 64 NW: one per thread); every load before the first store */
    constexpr int NT = 64 * NW, MAXI = ((16 * NW + 4) * 10 + NT - 1) / NT;
    const T *ysrc = pp<T>(a.in[0]);
    Chunk cy[MAXI], cc[2];
#pragma unroll
    for (int i = 0; i < MAXI; i++) cy[i] = chunk_load<T, 10>(ysrc, a.in[0].stride, x0, y0, tid + i * NT, xlim, ylim);
    if (c420)
#pragma unroll
      for (int p = 1; p < 3; p++)
        cc[p - 1] = chunk_load<T, 6>(pp<T>(a.in[p]), a.in[p].stride, x0 >> 1, y0 >> 1, tid, xlim >> 1, ylim >> 1);
#pragma unroll
    for (int i = 0; i < MAXI; i++)
      if (tid + i * NT < (bh + 4) * 10) chunk_store<10>(cy[i], words, words + SO, tid + i * NT, x0, y0, xlim, ylim);
    if (c420 && tid < ((bh >> 1) + 4) * 6)
#pragma unroll
      for (int p = 1; p < 3; p++)
        chunk_store<6, TSC / 2>(cc[p - 1], cwords + (p - 1) * 2 * CSO, cwords + (p - 1) * 2 * CSO + CSO, tid,
                                x0 >> 1, y0 >> 1, xlim >> 1, ylim >> 1);
  } else {
    stage_tile<T, NW>(tile, shifted, pp<T>(a.in[0]), a.in[0].stride, x0, y0, 64, TH, xlim, ylim);
    if (c420)
      for (int p = 1; p < 3; p++)
        chroma_load<T, NW>(cpre[p - 1], pp<T>(a.in[p]), a.in[p].stride, x0 >> 1, y0 >> 1, xlim >> 1, ylim >> 1);
  }
  __syncthreads();

  const int nf = bl.nfilt, nact = nf + bl.ncopy;
  const int slot = tid >> 2, q = tid & 3;
  const bool active = slot < nact, filt = slot < nf;
  /* this wave's slots hold filtered blocks (wave-uniform) */
  const bool wave_filters = 16 * (tid >> 6) < nf;
  const int b = active ? bl.list[slot] : 0;
  const int by = b >> 3, bx = b & 7;
  int ydir = 0;
  int32_t var = 0;
  if (wave_filters) block_direction(tile, tid, filt, by, bx, cs, &ydir, &var);
  const int pri = filt ? adjust_pri(a.h.y_pri[idx] << cs, var) : 0;

  /* luma: lane q filters (or copies) the column pair 2q, 2q + 1 of its block */
  if (active) {
    const int sec = a.h.y_sec[idx] << cs;
    const Taps k = make_taps(pri, sec, a.h.damping + cs, a.h.y_pri[idx] ? ydir : 0, cs, SO);
    T *o = pp<T>(a.out[0]) + (int64_t)(y0 + by * 8) * a.out[0].stride + x0 + bx * 8;
    const int t0 = (by * 8 + 2) * TS + bx * 8 + 2;
    if (wave_filters)
      filter_block_any<T>(a.h.y_pri[idx] != 0, sec != 0, words, t0, o, a.out[0].stride, 2 * q, 0, 8, k,
                          filt && (pri | sec));
    else
      filter_block_any<T>(false, false, words, t0, o, a.out[0].stride, 2 * q, 0, 8, k, false);
  }
  if (g.nplanes == 1) return;

  /* chroma planes: block of (8 >> ssx) x (8 >> ssy) = ncp column pairs;
   * lane q takes pair q % ncp and its share of the rows */
  const int ssx = g.ssx, ssy = g.ssy;
  const int cw = 64 >> ssx, ch = TH >> ssy;
  const int upri = a.h.uv_pri[idx] << cs, usec = a.h.uv_sec[idx] << cs;
  const int bws = 8 >> ssx, bhs = 8 >> ssy;
  const int ncp = bws >> 1;               /* 4 (4:4:4) or 2 */
  const int nrows = (bhs * ncp) >> 2;     /* rows per lane */
  const int cpair = q % ncp, r0 = (q / ncp) * nrows;
  const int udir = upri ? AV1_CDEF_UV_DIR[ssx][ssy][ydir] : 0;
  if (VEC && c420) {
    const Taps k = make_taps(upri, usec, a.h.damping + cs - 1, udir, cs, CSO, TSC / 2);
    if (active) {
      const int t0 = (by * 4 + 2) * TSC + bx * 4 + 2;
#pragma unroll
      for (int p = 1; p < 3; p++) {
        T *o = pp<T>(a.out[p]) + (int64_t)((y0 >> 1) + by * 4) * a.out[p].stride + (x0 >> 1) + bx * 4;
        const uint32_t *pw = cwords + (p - 1) * 2 * CSO;
        if (wave_filters)
          filter_block_any<T, TSC>(upri != 0, usec != 0, pw, t0, o, a.out[p].stride, 2 * cpair, r0, nrows, k,
                                   filt && (upri | usec));
        else
          filter_block_any<T, TSC>(false, false, pw, t0, o, a.out[p].stride, 2 * cpair, r0, nrows, k, false);
      }
    }
    return;
  }
  const Taps k = make_taps(upri, usec, a.h.damping + cs - 1, udir, cs, SO);
#pragma unroll
  for (int p = 1; p < 3; p++) {
    __syncthreads();
    /* (with chunked staging only 4:2:2 and 4:4:4 get here) */
    if (!VEC && c420) {
      chroma_store<NW>(cpre[p - 1], tile, shifted);
    } else {
      stage_tile<T, NW>(tile, shifted, pp<T>(a.in[p]), a.in[p].stride, x0 >> ssx, y0 >> ssy, cw, ch, xlim >> ssx,
                        ylim >> ssy);
    }
    __syncthreads();
    if (active) {
      T *o = pp<T>(a.out[p]) + (int64_t)((y0 >> ssy) + by * bhs) * a.out[p].stride + (x0 >> ssx) + bx * bws;
      const int t0 = (by * bhs + 2) * TS + bx * bws + 2;
      if (wave_filters)
        filter_block_any<T>(upri != 0, usec != 0, words, t0, o, a.out[p].stride, 2 * cpair, r0, nrows, k,
                            filt && (upri | usec));
      else
        filter_block_any<T>(false, false, words, t0, o, a.out[p].stride, 2 * cpair, r0, nrows, k, false);
    }
  }
}

/* Tiles of the per-frame kernel: CDEF_WAVES waves, 64 x 16 CDEF_WAVES
 * luma samples each. */
constexpr int CDEF_WAVES = 4;

/* 5 waves per SIMD: 96 VGPRs, no spills (at 6 waves, 80 VGPRs, the
 * prefetched 4:2:0 chroma tiles spill and the kernel is 6 % slower per 4K
 * frame). */
template <typename T, bool VEC>
__global__ __launch_bounds__(64 * CDEF_WAVES, 5) void cdef_kernel(CdefLaunch a) {
  /* the tile P and S; chunked staging adds both 4:2:0 chroma tiles' */
  __shared__ uint32_t words[2 * soff_rows(16 * CDEF_WAVES) + (VEC ? 4 * csoff_rows(8 * CDEF_WAVES) : 0)];
  __shared__ CdefBlocks bl;
  cdef_tile<T, CDEF_WAVES, VEC>(a, xcd_remap(blockIdx.x, gridDim.x), words, bl);
}

/* Batched. The persistent form is held at 5 waves per SIMD (inside the
 * grab loop the compiler otherwise takes 192 VGPRs, 2 waves; the few
 * spills the bound costs are cheaper than the lost latency hiding); the
 * direct form at 6, as the per-frame kernel. */
template <typename T, bool DIRECT, bool VEC>
__global__ __launch_bounds__(256, DIRECT ? 6 : 5) void cdef_batch(BatchWork w) {
  __shared__ uint32_t words[2 * SOFF + (VEC ? 4 * csoff_rows(32) : 0)];
  __shared__ CdefBlocks bl;
  batch_loop<DIRECT>(w, [&](int f, int j) { cdef_tile<T, 4, VEC>(batch_table<CdefLaunch>(w.rec, f), j, words, bl); });
}

/* ---- the encoder's strength search (src/encoder/cdef.rs:52-197) ---- */

constexpr int MAX_YC = 44, MAX_UVC = 24;

struct CdefSearchLaunch {
  Geom g;
  const rvpf_mi *mi;
  int mi_stride;
  const int8_t *cdef_idx;
  int cdef_stride;
  PlanePtr in[3], src[3];
  int fb_cols;
  int damping; /* cdef_damping (damping_minus_3 + 3) */
  int ny, nuv;
  int8_t ypri[MAX_YC], ysec[MAX_YC], uvpri[MAX_UVC], uvsec[MAX_UVC]; /* strengths (secondary 3 -> 4) */
  uint64_t *ysse, *uvsse; /* [64x64 block][candidate] */
  int8_t *coded;          /* [64x64 block]: 1 when it codes an index */
};

/* Two horizontally adjacent samples (even column) as a 16-bit pair. */
template <typename T>
__device__ __forceinline__ uint32_t load_pair(const T *p) {
  if constexpr (sizeof(T) == 2) {
    return *reinterpret_cast<const uint32_t *>(p);
  } else {
    const uint32_t w = *reinterpret_cast<const uint16_t *>(p);
    return (w & 0xffu) | ((w & 0xff00u) << 8);
  }
}

/* The squared error of this lane's filtered pairs (rows r0 .. r0 + n - 1
 * of the column pair c of the block at tile sample t0) against the source
 * pairs sp[], components outside the frame masked off. Differences fit 16
 * bits and 16 squares fit 31, so a 16-bit dot product accumulates them. */
template <bool PRI, bool SEC>
__device__ __forceinline__ int pairs_sse(const uint32_t *w, int t0, int c, int r0, int n, const Taps &k,
                                         const uint32_t *sp, const uint32_t *mask) {
  int e = 0;
#pragma unroll 2
  for (int r = 0; r < n; r++) {
    const uint32_t f = filter_pair<PRI, SEC>(w, (t0 + (r0 + r) * TS + c) >> 1, k);
    const uint32_t d = __builtin_bit_cast(uint32_t, as_s2(f) - as_s2(sp[r])) & mask[r];
    e = __builtin_amdgcn_sdot2(as_s2(d), as_s2(d), e, false);
  }
  return e;
}

__device__ __forceinline__ int pairs_sse_any(bool pri, bool sec, const uint32_t *w, int t0, int c, int r0, int n,
                                             const Taps &k, const uint32_t *sp, const uint32_t *mask) {
  if (pri && sec) return pairs_sse<true, true>(w, t0, c, r0, n, k, sp, mask);
  if (pri) return pairs_sse<true, false>(w, t0, c, r0, n, k, sp, mask);
  if (sec) return pairs_sse<false, true>(w, t0, c, r0, n, k, sp, mask);
  return pairs_sse<false, false>(w, t0, c, r0, n, k, sp, mask);
}

/* Adds the wave's sum of v to *acc (LDS). */
__device__ __forceinline__ void wave_add(uint64_t *acc, uint32_t v) {
  uint64_t x = v;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(reinterpret_cast<unsigned long long *>(acc), (unsigned long long)x);
}

/* One 64x64 block: every luma candidate on its 8x8 blocks, then every
 * chroma candidate on both chroma planes, each candidate's squared error
 * summed over the block. The direction search, the skip rule and the
 * filter are the decoder's (cdef_tile); the source pairs of a lane are
 * loaded once per plane and reused by every candidate. */
template <typename T>
__device__ void cdef_search_tile(const CdefSearchLaunch &a, int tile_id, uint32_t *words, uint64_t *acc) {
  uint16_t *tile = reinterpret_cast<uint16_t *>(words);
  uint16_t *shifted = reinterpret_cast<uint16_t *>(words + SOFF);
  const Geom &g = a.g;
  const int tid = threadIdx.x;
  const int fbx = tile_id % a.fb_cols, fby = tile_id / a.fb_cols;
  const int cs = g.bd - 8;
  if (gptr(a.cdef_idx)[(int64_t)fby * a.cdef_stride + fbx] < 0) {
    if (tid == 0) gptr(a.coded)[tile_id] = 0;
    return;
  }
  const int nacc = a.ny + a.nuv;
  if (tid < nacc) acc[tid] = 0;
  const int xlim = g.mi_cols * 4, ylim = g.mi_rows * 4;
  const int x0 = fbx * 64, y0 = fby * 64;
  const int bw = min(64, xlim - x0), bh = min(64, ylim - y0);
  stage_tile<T, 4>(tile, shifted, pp<T>(a.in[0]), a.in[0].stride, x0, y0, 64, 64, xlim, ylim);
  __syncthreads();

  const int b = tid >> 2, q = tid & 3;
  const int by = b >> 3, bx = b & 7;
  const bool inside = by * 8 < bh && bx * 8 < bw;
  bool skip = true;
  if (inside) {
    const int r = (y0 >> 2) + by * 2, c = (x0 >> 2) + bx * 2;
    const int ms = a.mi_stride;
    const rvpf_mi *mi = gptr(a.mi);
    skip = mi[(int64_t)r * ms + c].seg_skip & mi[(int64_t)(r + 1) * ms + c].seg_skip &
           mi[(int64_t)r * ms + c + 1].seg_skip & mi[(int64_t)(r + 1) * ms + c + 1].seg_skip & 1;
  }
  const bool filt = inside && !skip;
  int ydir;
  int32_t var;
  block_direction(tile, tid, filt, by, bx, cs, &ydir, &var);

  uint32_t sp[8], mask[8];
  {
    /* this lane's luma pairs: columns 2q, 2q + 1, rows 0..7 of its block */
    const int fw = g.width, fh = g.height;
    const int px = x0 + bx * 8 + 2 * q, py = y0 + by * 8;
    const T *src = pp<T>(a.src[0]);
    const int64_t ss = a.src[0].stride;
#pragma unroll
    for (int r = 0; r < 8; r++) {
      /* clamped into the mode-info area, which the source covers */
      sp[r] = load_pair<T>(src + (int64_t)min(py + r, ylim - 1) * ss + min(px, xlim - 2));
      mask[r] = !filt || py + r >= fh || px >= fw ? 0u : (px + 1 < fw ? 0xffffffffu : 0xffffu);
    }
  }
  const int t0 = (by * 8 + 2) * TS + bx * 8 + 2;
  for (int k = 0; k < a.ny; k++) {
    const int pri0 = a.ypri[k] << cs, sec = a.ysec[k] << cs;
    const Taps tk = make_taps(filt ? adjust_pri(pri0, var) : 0, sec, a.damping + cs, pri0 ? ydir : 0, cs, SOFF);
    wave_add(&acc[k], (uint32_t)pairs_sse_any(pri0 != 0, sec != 0, words, t0, 2 * q, 0, 8, tk, sp, mask));
  }

  if (g.nplanes > 1) {
    const int ssx = g.ssx, ssy = g.ssy;
    const int cw = 64 >> ssx, ch = 64 >> ssy;
    const int bws = 8 >> ssx, bhs = 8 >> ssy;
    const int ncp = bws >> 1;
    const int nrows = (bhs * ncp) >> 2;
    const int cpair = q % ncp, r0 = (q / ncp) * nrows;
    const int cdir = AV1_CDEF_UV_DIR[ssx][ssy][ydir];
    const int fw = (g.width + ssx) >> ssx, fh = (g.height + ssy) >> ssy;
    const int cxl = xlim >> ssx, cyl = ylim >> ssy;
    const int ct0 = (by * bhs + 2) * TS + bx * bws + 2;
    for (int p = 1; p < 3; p++) {
      __syncthreads();
      stage_tile<T, 4>(tile, shifted, pp<T>(a.in[p]), a.in[p].stride, x0 >> ssx, y0 >> ssy, cw, ch, cxl, cyl);
      __syncthreads();
      const int px = (x0 >> ssx) + bx * bws + 2 * cpair, py = (y0 >> ssy) + by * bhs + r0;
      const T *src = pp<T>(a.src[p]);
      const int64_t ss = a.src[p].stride;
#pragma unroll
      for (int r = 0; r < 8; r++) {
        sp[r] = load_pair<T>(src + (int64_t)min(py + r, cyl - 1) * ss + min(px, cxl - 2));
        mask[r] = !filt || r >= nrows || py + r >= fh || px >= fw ? 0u : (px + 1 < fw ? 0xffffffffu : 0xffffu);
      }
      for (int k = 0; k < a.nuv; k++) {
        const int pri = a.uvpri[k] << cs, sec = a.uvsec[k] << cs;
        const Taps tk = make_taps(pri, sec, a.damping + cs - 1, pri ? cdir : 0, cs, SOFF);
        wave_add(&acc[a.ny + k], (uint32_t)pairs_sse_any(pri != 0, sec != 0, words, ct0, 2 * cpair, r0, nrows, tk,
                                                         sp, mask));
      }
    }
  }
  __syncthreads();
  if (tid < a.ny)
    gptr(a.ysse)[(int64_t)tile_id * a.ny + tid] = acc[tid];
  else if (tid < nacc)
    gptr(a.uvsse)[(int64_t)tile_id * a.nuv + tid - a.ny] = acc[tid];
  if (tid == 0) gptr(a.coded)[tile_id] = 1;
}

template <typename T>
__global__ __launch_bounds__(256) void cdef_search_kernel(CdefSearchLaunch a) {
  __shared__ uint32_t words[2 * SOFF];
  __shared__ uint64_t acc[MAX_YC + MAX_UVC];
  cdef_search_tile<T>(a, xcd_remap(blockIdx.x, gridDim.x), words, acc);
}

int plan(const Geom &g, const CdefHdr &h, const rvpf_mi *mi, int mi_stride, const int8_t *cdef_idx,
         int cdef_stride, int enabled, const PlanePtr in[3], const PlanePtr out[3], CdefLaunch &a) {
  a.g = g;
  a.h = h;
  a.mi = mi;
  a.mi_stride = mi_stride;
  a.cdef_idx = cdef_idx;
  a.cdef_stride = cdef_stride;
  for (int p = 0; p < 3; p++) {
    a.in[p] = in[p];
    a.out[p] = out[p];
  }
  a.enabled = enabled;
  a.fb_cols = (g.mi_cols + 15) / 16;
  /* chunked staging: 8 samples per access, 16 bytes (u16) or 8 (u8);
   * 16-byte aligned planes with strides of 8k samples suit both */
  a.vec = (g.mi_cols * 4) >> g.ssx >= 8; /* a chunk's access stays inside the mode-info width */
  for (int p = 0; p < g.nplanes; p++)
    if ((uintptr_t)in[p].data % 16 != 0 || in[p].stride % 8 != 0) a.vec = 0;
  return a.fb_cols * ((g.mi_rows + 15) / 16);
}

}  // namespace

size_t cdef_record_size() { return sizeof(CdefLaunch); }

int cdef_plan(const Geom &g, const CdefHdr &h, const rvpf_mi *mi, int mi_stride, const int8_t *cdef_idx,
              int cdef_stride, int enabled, const PlanePtr in[3], const PlanePtr out[3], void *rec) {
  return plan(g, h, mi, mi_stride, cdef_idx, cdef_stride, enabled, in, out, *static_cast<CdefLaunch *>(rec));
}

int cdef_record_vec(const void *rec) { return static_cast<const CdefLaunch *>(rec)->vec; }

template <typename T, bool VEC>
void cdef_batch_launch(const BatchWork &w, const StagePlan &sp, hipStream_t st) {
  if (batch_direct(w, sp.grid))
    cdef_batch<T, true, VEC><<<sp.grid, 256, 0, st>>>(w);
  else
    cdef_batch<T, false, VEC><<<sp.grid, 256, 0, st>>>(w);
}

int launch_cdef_batch(const BatchWork &w, const StagePlan &sp, int bps, int vec, hipStream_t st) {
  if (w.total == 0) return 0;
  if (bps == 1 && vec)
    cdef_batch_launch<uint8_t, true>(w, sp, st);
  else if (bps == 1)
    cdef_batch_launch<uint8_t, false>(w, sp, st);
  else if (vec)
    cdef_batch_launch<uint16_t, true>(w, sp, st);
  else
    cdef_batch_launch<uint16_t, false>(w, sp, st);
  return (int)hipGetLastError();
}

int launch_cdef(const Geom &g, const CdefHdr &h, const rvpf_mi *mi, int mi_stride,
                const int8_t *cdef_idx, int cdef_stride, int enabled, const PlanePtr in[3],
                PlanePtr out[3], int bps, hipStream_t st) {
  CdefLaunch a;
  const int n = plan(g, h, mi, mi_stride, cdef_idx, cdef_stride, enabled, in, out, a) * (4 / CDEF_WAVES);
  if (bps == 1 && a.vec)
    cdef_kernel<uint8_t, true><<<n, 64 * CDEF_WAVES, 0, st>>>(a);
  else if (bps == 1)
    cdef_kernel<uint8_t, false><<<n, 64 * CDEF_WAVES, 0, st>>>(a);
  else if (a.vec)
    cdef_kernel<uint16_t, true><<<n, 64 * CDEF_WAVES, 0, st>>>(a);
  else
    cdef_kernel<uint16_t, false><<<n, 64 * CDEF_WAVES, 0, st>>>(a);
  return (int)hipGetLastError();
}

int launch_cdef_search(const Geom &g, const rvpf_mi *mi, int mi_stride, const int8_t *cdef_idx, int cdef_stride,
                       const PlanePtr in[3], const PlanePtr src[3], int damping, const int8_t *ypri,
                       const int8_t *ysec, int ny, const int8_t *uvpri, const int8_t *uvsec, int nuv, uint64_t *ysse,
                       uint64_t *uvsse, int8_t *coded, int bps, hipStream_t st) {
  if (ny < 1 || ny > MAX_YC || nuv < 1 || nuv > MAX_UVC) return (int)hipErrorInvalidValue;
  CdefSearchLaunch a;
  a.g = g;
  a.mi = mi;
  a.mi_stride = mi_stride;
  a.cdef_idx = cdef_idx;
  a.cdef_stride = cdef_stride;
  for (int p = 0; p < 3; p++) {
    a.in[p] = in[p];
    a.src[p] = src[p];
  }
  a.fb_cols = (g.mi_cols + 15) / 16;
  a.damping = damping;
  a.ny = ny;
  a.nuv = nuv;
  for (int k = 0; k < MAX_YC; k++) {
    a.ypri[k] = k < ny ? ypri[k] : 0;
    a.ysec[k] = k < ny ? ysec[k] : 0;
  }
  for (int k = 0; k < MAX_UVC; k++) {
    a.uvpri[k] = k < nuv ? uvpri[k] : 0;
    a.uvsec[k] = k < nuv ? uvsec[k] : 0;
  }
  a.ysse = ysse;
  a.uvsse = uvsse;
  a.coded = coded;
  const int n = a.fb_cols * ((g.mi_rows + 15) / 16);
  if (bps == 1)
    cdef_search_kernel<uint8_t><<<n, 256, 0, st>>>(a);
  else
    cdef_search_kernel<uint16_t><<<n, 256, 0, st>>>(a);
  return (int)hipGetLastError();
}

}  // namespace rvpf
