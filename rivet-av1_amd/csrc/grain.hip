/*
 * grain.hip — film grain synthesis (AV1 7.18.3) for gfx950.
 *
 * Follows rivet-av1's grain::apply (src/decoder/grain.rs:8-354).
 *
 * grain_prepare_kernel: one workgroup of two waves per frame.
 *   - Gaussian templates (7.18.3.3): luma 73x82, cb/cr 38x44 (4:2:0) or
 *     larger. The 16-bit LFSR is linear over GF(2), so lane t jumps to the
 *     state after t*C steps with precomputed powers A^(2^k) of the step
 *     matrix and then generates its C values.
 *   - The auto-regressive filter is a recursion: sample (y, x) reads rows
 *     y-lag..y and columns x-lag..x+lag. Processing it along the diagonals
 *     T = (lag+1)*y + x visits every dependency at an earlier T, so all
 *     samples of one T run in parallel (one lane per template row); the
 *     chroma recursion runs on the second wave, interleaved with luma.
 *   - Scaling LUTs (7.18.3.4) and the per-stripe random block offsets.
 * grain_apply_kernel: one thread per chroma sample (and the 1, 2 or 4 luma
 *   samples it covers). The noise of a sample is rebuilt directly from the
 *   template: its 32x32 block's offset, the horizontal blend with the block
 *   to the left, then the vertical blend with the stripe above (whose value
 *   is itself horizontally blended), exactly the order in which the
 *   reference fills its noise stripes and noise image. Chroma reads the
 *   un-noised luma of the input, as the reference blends chroma first.
 */
#include "av1_tables.h"
#include "rvpf_internal.h"

namespace rvpf {

namespace {

/* --- LFSR jump tables ------------------------------------------------ */
constexpr uint16_t lfsr_step(uint16_t r) {
  return (uint16_t)((r >> 1) | ((((r) ^ (r >> 1) ^ (r >> 3) ^ (r >> 12)) & 1) << 15));
}

struct JumpTable {
  uint16_t col[16][16]; /* col[k][j] = A^(2^k) e_j */
};

constexpr uint16_t apply_cols(const uint16_t *cols, uint16_t s) {
  uint16_t r = 0;
  for (int j = 0; j < 16; j++)
    if ((s >> j) & 1) r ^= cols[j];
  return r;
}

constexpr JumpTable make_jump_table() {
  JumpTable t{};
  for (int j = 0; j < 16; j++) t.col[0][j] = lfsr_step((uint16_t)(1u << j));
  for (int k = 1; k < 16; k++)
    for (int j = 0; j < 16; j++) t.col[k][j] = apply_cols(t.col[k - 1], t.col[k - 1][j]);
  return t;
}

constexpr JumpTable JUMP = make_jump_table();

__device__ __forceinline__ uint32_t jump(uint32_t s, uint32_t n) {
  for (int k = 0; k < 16 && n; k++, n >>= 1)
    if (n & 1) {
      uint32_t r = 0;
#pragma unroll
      for (int j = 0; j < 16; j++) r ^= ((s >> j) & 1) ? JUMP.col[k][j] : 0u;
      s = r;
    }
  return s;
}

__device__ __forceinline__ uint32_t step(uint32_t r) {
  const uint32_t bit = (r ^ (r >> 1) ^ (r >> 3) ^ (r >> 12)) & 1;
  return (r >> 1) | (bit << 15);
}

__device__ __forceinline__ int round2(int x, int n) { return (x + ((1 << n) >> 1)) >> n; }
__device__ __forceinline__ int clip3(int lo, int hi, int x) { return min(max(x, lo), hi); }

constexpr int TW = 82, TH = 73;

/* Gaussian values of the LFSR sequence from `seed` into t (rows x cols, stride TW). */
__device__ void gaussian_fill(int16_t *t, const int16_t *gauss, uint32_t seed, int rows, int cols, int shift,
                              bool on, int lane) {
  const int n = rows * cols;
  const int per = (n + 63) / 64;
  const int i0 = lane * per, i1 = min(i0 + per, n);
  int r = i0 / cols, c = i0 - r * cols;
  uint32_t s = on ? jump(seed, (uint32_t)i0) : 0;
  for (int i = i0; i < i1; i++) {
    int v = 0;
    if (on) {
      s = step(s);
      v = round2(gauss[(s >> 5) & 2047], shift);
    }
    t[r * TW + c] = (int16_t)v;
    if (++c == cols) {
      c = 0;
      r++;
    }
  }
}

/* One diagonal step of the luma AR of 7.18.3.3 (grain.rs:57-73): the
 * samples (y, x) with (LAG+1)*y + x = T, one lane each. Four partial sums
 * keep the multiply-adds out of one dependency chain. */
template <int LAG>
__device__ __forceinline__ void luma_step(int16_t *tl, const int *c, int T, int lane, int ars, int gmin, int gmax) {
  constexpr int K = LAG + 1;
  const int ymin = max(3, (T - 78 + K - 1) / K);
  const int y = ymin + lane;
  const int x = T - K * y;
  if (y > 72 || x < 3) return;
  int s[4] = {0, 0, 0, 0}, pos = 0;
#pragma unroll
  for (int dr = -LAG; dr <= 0; dr++)
#pragma unroll
    for (int dc = -LAG; dc <= LAG; dc++) {
      if (dr == 0 && dc >= 0) continue;
      s[pos & 3] += __mul24((int)tl[(y + dr) * TW + x + dc], c[pos]);
      pos++;
    }
  tl[y * TW + x] = (int16_t)clip3(gmin, gmax, tl[y * TW + x] + round2((s[0] + s[1]) + (s[2] + s[3]), ars));
}

/* One chroma AR sample (grain.rs:101-134): cb and cr together, plus the
 * tap on the co-located (averaged) luma. */
template <int LAG>
__device__ __forceinline__ void chroma_one(int16_t *tcb, int16_t *tcr, const int16_t *tl, const int *c0,
                                           const int *c1, const GrainHdr &h, int y, int x, int gmin, int gmax) {
  constexpr int NC = 2 * LAG * (LAG + 1) + 1;
  int s0[2] = {0, 0}, s1[2] = {0, 0}, pos = 0;
#pragma unroll
  for (int dr = -LAG; dr <= 0; dr++)
#pragma unroll
    for (int dc = -LAG; dc <= LAG; dc++) {
      if (dr == 0 && dc >= 0) continue;
      s0[pos & 1] += __mul24((int)tcb[(y + dr) * TW + x + dc], c0[pos]);
      s1[pos & 1] += __mul24((int)tcr[(y + dr) * TW + x + dc], c1[pos]);
      pos++;
    }
  int a0 = s0[0] + s0[1], a1 = s1[0] + s1[1];
  if (h.num_y > 0) {
    const int lx = ((x - 3) << h.ssx) + 3, ly = ((y - 3) << h.ssy) + 3;
    int l = 0;
    for (int i = 0; i <= h.ssy; i++)
      for (int j = 0; j <= h.ssx; j++) l += tl[(ly + i) * TW + lx + j];
    l = round2(l, h.ssx + h.ssy);
    a0 += __mul24(l, c0[NC - 1]);
    a1 += __mul24(l, c1[NC - 1]);
  }
  tcb[y * TW + x] = (int16_t)clip3(gmin, gmax, tcb[y * TW + x] + round2(a0, h.ar_shift));
  tcr[y * TW + x] = (int16_t)clip3(gmin, gmax, tcr[y * TW + x] + round2(a1, h.ar_shift));
}

/* One diagonal step of the chroma AR: the samples with
 * (LAG+1)*y + x = T, 3 <= x <= cw-4, 3 <= y < ch. */
template <int LAG>
__device__ __forceinline__ void chroma_step(int16_t *tcb, int16_t *tcr, const int16_t *tl, const int *c0,
                                            const int *c1, const GrainHdr &h, int cw, int ch, int T, int lane,
                                            int gmin, int gmax) {
  constexpr int K = LAG + 1;
  const int xmax = cw - 4;
  const int ymin = max(3, (T - xmax + K - 1) / K);
  for (int y = ymin + lane; y < ch; y += 64) {
    const int x = T - K * y;
    if (x < 3) break;
    chroma_one<LAG>(tcb, tcr, tl, c0, c1, h, y, x, gmin, gmax);
  }
}

/* Both AR recursions in lock step: wave 0 walks the luma diagonals
 * T = K*y + x (K = LAG+1) at global step G = T; wave 1 walks the chroma
 * ones at G = 2T - 2K - 1 (subsampled chroma) or G = T + 1 (4:4:4). Every
 * luma sample a chroma sample (y, x) reads was finished at a smaller G:
 * for 4:2:0 the luma samples (2y-3+i, 2x-3+j), i, j in {0, 1}, finish by
 * G = 2T - 2K - 2; for 4:2:2, (y, 2x-3+j) finish by T + x - 2 <= 2T - 3K - 2;
 * for 4:4:4, (y, x) finishes at G = T. */
template <int LAG>
__device__ void ar_both(int16_t *tl, int16_t *tcb, int16_t *tcr, const GrainHdr &h, bool chroma, int cw, int ch,
                        int gmin, int gmax) {
  constexpr int K = LAG + 1, NL = 2 * LAG * (LAG + 1), NC = NL + 1;
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int cl[NL > 0 ? NL : 1], c0[NC], c1[NC];
#pragma unroll
  for (int i = 0; i < NL; i++) cl[i] = h.ar_y[i];
#pragma unroll
  for (int i = 0; i < NC; i++) {
    c0[i] = h.ar_cb[i];
    c1[i] = h.ar_cr[i];
  }
  if (LAG == 0) {
    /* no neighbour taps: luma samples are only clipped, and each chroma
     * sample depends on the (final) luma alone */
    for (int i = threadIdx.x; i < 70 * 76; i += 128) {
      const int y = 3 + i / 76, x = 3 + i % 76;
      tl[y * TW + x] = (int16_t)clip3(gmin, gmax, tl[y * TW + x]);
    }
    __syncthreads();
    if (chroma) {
      const int rw = cw - 6, nr = ch - 3;
      for (int i = threadIdx.x; i < rw * nr; i += 128) {
        const int y = 3 + i / rw, x = 3 + i % rw;
        chroma_one<0>(tcb, tcr, tl, c0, c1, h, y, x, gmin, gmax);
      }
    }
    __syncthreads();
    return;
  }
  const int sub = h.ssx | h.ssy;
  const int tl_lo = 3 * K + 3, tl_hi = 72 * K + 78;
  const int tc_lo = 3 * K + 3, tc_hi = (ch - 1) * K + (cw - 4);
  const int g_hi = max(tl_hi, chroma ? (sub ? 2 * tc_hi - 2 * K - 1 : tc_hi + 1) : 0);
  for (int G = tl_lo; G <= g_hi; G++) {
    if (wave == 0) {
      if (G <= tl_hi) luma_step<LAG>(tl, cl, G, lane, h.ar_shift, gmin, gmax);
    } else if (chroma) {
      const int T = sub ? ((G & 1) ? (G + 2 * K + 1) >> 1 : -1) : G - 1;
      if (T >= tc_lo && T <= tc_hi) chroma_step<LAG>(tcb, tcr, tl, c0, c1, h, cw, ch, T, lane, gmin, gmax);
    }
    __syncthreads();
  }
}

struct PrepShared {
  int16_t tl[TH * TW], tcb[TH * TW], tcr[TH * TW];
  int16_t gauss[2048];
};

/* Templates, LUTs and block offsets of one frame (128 threads). */
__device__ __forceinline__ void grain_prepare_body(const GrainHdr &h, const GrainScratch &sc, PrepShared &sm) {
  int16_t *tl = sm.tl, *tcb = sm.tcb, *tcr = sm.tcr, *gauss = sm.gauss;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < 2048; i += 128) gauss[i] = AV1_GAUSSIAN_SEQUENCE[i];
  __syncthreads();
  const int bd = h.bd;
  const int gmin = -(128 << (bd - 8)), gmax = (256 << (bd - 8)) - 1 - (128 << (bd - 8));
  const int shift = 12 - bd + h.grain_scale_shift;
  const bool chroma = h.nplanes > 1;
  const int cw = h.ssx ? 44 : 82, ch = h.ssy ? 38 : 73;
  if (wave == 0) {
    gaussian_fill(tl, gauss, h.seed, TH, TW, shift, h.num_y > 0, lane);
  } else if (chroma) {
    gaussian_fill(tcb, gauss, h.seed ^ 0xb524, ch, cw, shift, h.num_cb > 0 || h.cfl, lane);
    gaussian_fill(tcr, gauss, h.seed ^ 0x49d8, ch, cw, shift, h.num_cr > 0 || h.cfl, lane);
  }
  __syncthreads();
  switch (h.ar_lag) {
    case 0: ar_both<0>(tl, tcb, tcr, h, chroma, cw, ch, gmin, gmax); break;
    case 1: ar_both<1>(tl, tcb, tcr, h, chroma, cw, ch, gmin, gmax); break;
    case 2: ar_both<2>(tl, tcb, tcr, h, chroma, cw, ch, gmin, gmax); break;
    default: ar_both<3>(tl, tcb, tcr, h, chroma, cw, ch, gmin, gmax); break;
  }
  /* templates out */
  for (int i = threadIdx.x; i < TH * TW; i += 128) {
    sc.tmpl[i] = tl[i];
    sc.tmpl[TH * TW + i] = chroma ? tcb[i] : 0;
    sc.tmpl[2 * TH * TW + i] = chroma ? tcr[i] : 0;
  }
  /* scaling LUTs (grain.rs:137-165) */
  for (int p = 0; p < h.nplanes; p++) {
    const int np = h.npoints[p];
    const int *xs = h.points[p][0], *ys = h.points[p][1];
    for (int x = threadIdx.x; x < 256; x += 128) {
      int v = 0;
      if (np > 0) {
        if (x < xs[0]) {
          v = ys[0];
        } else if (x >= xs[np - 1]) {
          v = ys[np - 1];
        } else {
          int i = 0;
          while (i + 1 < np && xs[i + 1] <= x) i++;
          const int dy = ys[i + 1] - ys[i], dx = xs[i + 1] - xs[i];
          const int delta = dy * ((65536 + (dx >> 1)) / dx);
          v = ys[i] + (((x - xs[i]) * delta + 32768) >> 16);
        }
      }
      sc.lut[p * 256 + x] = (int16_t)v;
    }
  }
  /* block offsets per noise stripe (grain.rs:186-199) */
  const int nstripes = ((h.h + 1) / 2 + 15) / 16;
  const int nblocks = ((h.w + 1) / 2 + 15) / 16;
  for (int ln = threadIdx.x; ln < nstripes; ln += 128) {
    uint32_t r = h.seed;
    r ^= (uint32_t)(((ln * 37 + 178) & 255) << 8);
    r ^= (uint32_t)((ln * 173 + 105) & 255);
    for (int b = 0; b < nblocks; b++) {
      r = step(r);
      sc.offsets[ln * sc.max_blocks + b] = (uint8_t)((r >> 8) & 255);
    }
  }
}

/* Noise of plane sample (x, y): blocks of 32 (16 subsampled) with a 2 (1)
 * sample overlap, horizontally blended, in noise-stripe row i of stripe ln. */
__device__ __forceinline__ int stripe_noise(const int16_t *t, const uint8_t *offs, int mb, int ln,
                                            int i, int x, int psx, int psy, bool overlap, int gmin,
                                            int gmax) {
  const int bs = 32 >> psx;
  const int bx = x / bs, j = x - bx * bs;
  const int o = offs[ln * mb + bx];
  const int ox = o >> 4, oy = o & 15;
  const int pox = psx ? 6 + ox : 9 + ox * 2, poy = psy ? 6 + oy : 9 + oy * 2;
  int gv = t[(poy + i) * TW + pox + j];
  if (overlap && bx > 0 && j < (psx ? 1 : 2)) {
    const int o2 = offs[ln * mb + bx - 1];
    const int ox2 = o2 >> 4, oy2 = o2 & 15;
    const int pox2 = psx ? 6 + ox2 : 9 + ox2 * 2, poy2 = psy ? 6 + oy2 : 9 + oy2 * 2;
    const int old = t[(poy2 + i) * TW + pox2 + bs + j];
    if (psx)
      gv = old * 23 + gv * 22;
    else
      gv = j == 0 ? old * 27 + gv * 17 : old * 17 + gv * 27;
    gv = clip3(gmin, gmax, round2(gv, 5));
  }
  return gv;
}

__device__ __forceinline__ int noise_at(const int16_t *t, const uint8_t *offs, int mb, int x, int y,
                                        int psx, int psy, bool overlap, int gmin, int gmax) {
  const int sh = 5 - psy;
  const int ln = y >> sh, i = y - (ln << sh);
  int gv = stripe_noise(t, offs, mb, ln, i, x, psx, psy, overlap, gmin, gmax);
  if (overlap && ln > 0 && i < (psy ? 1 : 2)) {
    const int old = stripe_noise(t, offs, mb, ln - 1, i + (32 >> psy), x, psx, psy, overlap, gmin, gmax);
    if (psy)
      gv = old * 23 + gv * 22;
    else
      gv = i == 0 ? old * 27 + gv * 17 : old * 17 + gv * 27;
    gv = clip3(gmin, gmax, round2(gv, 5));
  }
  return gv;
}

__device__ __forceinline__ int scale_lut(const int16_t *lut, int bd, int index) {
  const int sh = bd - 8;
  const int x = index >> sh;
  const int rem = index - (x << sh);
  if (bd == 8 || x == 255) return lut[x];
  const int start = lut[x], end = lut[x + 1];
  return start + round2((end - start) * rem, sh);
}

/* Tile tile_id of EX x (256 / EX) chroma positions (raster order). */
template <typename TI, typename TO, int EX>
__device__ __forceinline__ void grain_apply_tile(const GrainHdr &h, const GrainScratch &sc, const PlanePtr &i0,
                                                 const PlanePtr &i1, const PlanePtr &i2, const PlanePtr &o0,
                                                 const PlanePtr &o1, const PlanePtr &o2, int tile_id, int16_t *lut) {
  for (int i = threadIdx.x; i < h.nplanes * 256; i += 256) lut[i] = sc.lut[i];
  __syncthreads();
  const int ssx = h.nplanes > 1 ? h.ssx : 0, ssy = h.nplanes > 1 ? h.ssy : 0;
  const int cw = (h.w + ssx) >> ssx, ch = (h.h + ssy) >> ssy;
  const int nbx = (cw + EX - 1) / EX;
  const int cx = (tile_id % nbx) * EX + (threadIdx.x % EX);
  const int cy = (tile_id / nbx) * (256 / EX) + (threadIdx.x / EX);
  if (cx >= cw || cy >= ch) return;
  const int bd = h.bd;
  const int gmin = -(128 << (bd - 8)), gmax = (256 << (bd - 8)) - 1 - (128 << (bd - 8));
  const bool ov = h.overlap;
  const int mb = sc.max_blocks;
  const TI *lin = gptr(reinterpret_cast<const TI *>(i0.data));
  TO *lout = gptr(reinterpret_cast<TO *>(o0.data));
  if (h.nplanes > 1) {
    const int lx = cx << ssx, ly = cy << ssy;
    const int lnx = min(lx + 1, h.w - 1);
    const int64_t lrow = (int64_t)ly * i0.stride;
    const int avg = ssx ? round2((int)lin[lrow + lx] + (int)lin[lrow + lnx], 1) : (int)lin[lrow + lx];
    const int maxv = (1 << bd) - 1;
    for (int p = 1; p < 3; p++) {
      const PlanePtr &ip = p == 1 ? i1 : i2;
      const PlanePtr &op = p == 1 ? o1 : o2;
      const int orig = gptr(reinterpret_cast<const TI *>(ip.data))[(int64_t)cy * ip.stride + cx];
      int v = orig;
      const int npts = p == 1 ? h.num_cb : h.num_cr;
      if (npts > 0 || h.cfl) {
        int merged;
        if (h.cfl) {
          merged = avg;
        } else {
          const int lm = p == 1 ? h.cb_luma_mult : h.cr_luma_mult;
          const int mu = p == 1 ? h.cb_mult : h.cr_mult;
          const int of = p == 1 ? h.cb_offset : h.cr_offset;
          const int combined = avg * (lm - 128) + orig * (mu - 128);
          merged = clip3(0, maxv, (combined >> 6) + (of - 256) * (1 << (bd - 8)));
        }
        const int noise = noise_at(sc.tmpl + p * TH * TW, sc.offsets, mb, cx, cy, ssx, ssy, ov, gmin, gmax);
        const int sn = round2(scale_lut(lut + p * 256, bd, merged) * noise, h.scaling_shift);
        v = clip3(h.min_value, h.max_chroma, orig + sn);
      }
      gptr(reinterpret_cast<TO *>(op.data))[(int64_t)cy * op.stride + cx] = (TO)v;
    }
  }
  /* luma samples of this chroma position */
  for (int dy = 0; dy <= ssy; dy++) {
    const int ly = (cy << ssy) + dy;
    if (ly >= h.h) break;
    for (int dx = 0; dx <= ssx; dx++) {
      const int lx = (cx << ssx) + dx;
      if (lx >= h.w) break;
      const int orig = lin[(int64_t)ly * i0.stride + lx];
      int v = orig;
      if (h.num_y > 0) {
        const int noise = noise_at(sc.tmpl, sc.offsets, mb, lx, ly, 0, 0, ov, gmin, gmax);
        const int sn = round2(scale_lut(lut, bd, orig) * noise, h.scaling_shift);
        v = clip3(h.min_value, h.max_luma, orig + sn);
      }
      lout[(int64_t)ly * o0.stride + lx] = (TO)v;
    }
  }
}

__global__ __launch_bounds__(128) void grain_prepare_kernel(GrainHdr h, GrainScratch sc) {
  __shared__ PrepShared sm;
  grain_prepare_body(h, sc, sm);
}

template <typename TI, typename TO>
__global__ __launch_bounds__(256) void grain_apply_kernel(GrainHdr h, GrainScratch sc, PlanePtr i0,
                                                          PlanePtr i1, PlanePtr i2, PlanePtr o0,
                                                          PlanePtr o1, PlanePtr o2) {
  __shared__ int16_t lut[3 * 256];
  grain_apply_tile<TI, TO, 64>(h, sc, i0, i1, i2, o0, o1, o2, xcd_remap(blockIdx.x, gridDim.x), lut);
}

struct PrepRecord {
  GrainHdr h;
  GrainScratch s;
};

struct ApplyRecord {
  GrainHdr h;
  GrainScratch s;
  PlanePtr in[3], out[3];
};

__global__ __launch_bounds__(128) void grain_prepare_batch(BatchWork w) {
  __shared__ PrepShared sm;
  batch_loop(w, [&](int f, int) {
    const PrepRecord &r = batch_table<PrepRecord>(w.rec, f);
    grain_prepare_body(r.h, r.s, sm);
  });
}

template <typename T, int EX>
__global__ __launch_bounds__(256) void grain_apply_batch(BatchWork w) {
  __shared__ int16_t lut[3 * 256];
  batch_loop(w, [&](int f, int j) {
    const ApplyRecord &r = batch_table<ApplyRecord>(w.rec, f);
    grain_apply_tile<T, T, EX>(r.h, r.s, r.in[0], r.in[1], r.in[2], r.out[0], r.out[1], r.out[2], j, lut);
  });
}

constexpr int SHAPE_EX[GRAIN_SHAPES] = {64, 32, 16};

int apply_tiles(const GrainHdr &h, int ex) {
  const int ssx = h.nplanes > 1 ? h.ssx : 0, ssy = h.nplanes > 1 ? h.ssy : 0;
  const int cw = (h.w + ssx) >> ssx, ch = (h.h + ssy) >> ssy;
  return ((cw + ex - 1) / ex) * ((ch + 256 / ex - 1) / (256 / ex));
}

template <typename T>
void launch_apply_shape(const BatchWork &w, const StagePlan &sp, hipStream_t st) {
  switch (sp.shape) {
    case 1: grain_apply_batch<T, 32><<<sp.grid, 256, 0, st>>>(w); break;
    case 2: grain_apply_batch<T, 16><<<sp.grid, 256, 0, st>>>(w); break;
    default: grain_apply_batch<T, 64><<<sp.grid, 256, 0, st>>>(w); break;
  }
}

}  // namespace

int launch_grain_prepare(const GrainHdr &h, const GrainScratch &s, hipStream_t st) {
  grain_prepare_kernel<<<1, 128, 0, st>>>(h, s);
  return (int)hipGetLastError();
}

size_t grain_prep_record_size() { return sizeof(PrepRecord); }

int grain_prep_plan(const GrainHdr &h, const GrainScratch &s, void *rec) {
  PrepRecord &r = *static_cast<PrepRecord *>(rec);
  r.h = h;
  r.s = s;
  return 1;
}

int launch_grain_prepare_batch(const BatchWork &w, const StagePlan &sp, hipStream_t st) {
  if (w.total == 0) return 0;
  grain_prepare_batch<<<sp.grid, 128, 0, st>>>(w);
  return (int)hipGetLastError();
}

size_t grain_apply_record_size() { return sizeof(ApplyRecord); }

int grain_apply_plan(const GrainHdr &h, const GrainScratch &s, const PlanePtr in[3], const PlanePtr out[3],
                     int shape, void *rec) {
  ApplyRecord &r = *static_cast<ApplyRecord *>(rec);
  r.h = h;
  r.s = s;
  for (int p = 0; p < 3; p++) {
    r.in[p] = in[p];
    r.out[p] = out[p];
  }
  return apply_tiles(h, SHAPE_EX[shape < 0 || shape >= GRAIN_SHAPES ? 0 : shape]);
}

int launch_grain_apply_batch(const BatchWork &w, const StagePlan &sp, int bps, hipStream_t st) {
  if (w.total == 0) return 0;
  if (bps == 1)
    launch_apply_shape<uint8_t>(w, sp, st);
  else
    launch_apply_shape<uint16_t>(w, sp, st);
  return (int)hipGetLastError();
}

int launch_grain_apply(const GrainHdr &h, const GrainScratch &s, const PlanePtr in[3],
                       PlanePtr out[3], int bps, hipStream_t st) {
  const int grid = apply_tiles(h, 64);
  if (bps == 1)
    grain_apply_kernel<uint8_t, uint8_t><<<grid, 256, 0, st>>>(h, s, in[0], in[1], in[2], out[0], out[1], out[2]);
  else
    grain_apply_kernel<uint16_t, uint16_t><<<grid, 256, 0, st>>>(h, s, in[0], in[1], in[2], out[0], out[1], out[2]);
  return (int)hipGetLastError();
}

}  // namespace rvpf
