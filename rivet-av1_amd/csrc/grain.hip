/*
 * grain.hip — film grain synthesis (AV1 7.18.3) for gfx950.
 *
 * Follows rivet-av1's grain::apply (src/decoder/grain.rs:8-354).
 *
 * grain_prepare_kernel: one workgroup of three waves per frame.
 *   - Gaussian templates (7.18.3.3): luma 73x82, cb/cr 38x44 (4:2:0) or
 *     larger. The 16-bit LFSR is linear over GF(2), so lane t jumps to the
 *     state after t*C steps with precomputed powers A^(2^k) of the step
 *     matrix and then generates its C values.
 *   - The auto-regressive filter is a recursion: sample (y, x) reads rows
 *     y-lag..y and columns x-lag..x+lag. Processing it along the diagonals
 *     T = (lag+1)*y + x visits every dependency at an earlier T, so all
 *     samples of one T run in parallel (one lane per template row). The
 *     rows above a lane's row belong to the lanes below it, so the taps
 *     are passed between lanes by DPP rotations and stay in registers.
 *     Luma runs on wave 0, cb and cr on waves 1 and 2, a fixed number of
 *     diagonals behind the luma samples they read.
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
constexpr int PREP_THREADS = 192; /* three waves: luma, cb, cr */

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

/* One chroma AR sample (grain.rs:101-134) with no neighbour taps (lag 0):
 * cb and cr from the co-located (averaged) luma alone. */
__device__ __forceinline__ void chroma_lag0(int16_t *tcb, int16_t *tcr, const int16_t *tl, const GrainHdr &h, int y,
                                            int x, int gmin, int gmax) {
  int a0 = 0, a1 = 0;
  if (h.num_y > 0) {
    const int lx = ((x - 3) << h.ssx) + 3, ly = ((y - 3) << h.ssy) + 3;
    int l = 0;
    for (int i = 0; i <= h.ssy; i++)
      for (int j = 0; j <= h.ssx; j++) l += tl[(ly + i) * TW + lx + j];
    l = round2(l, h.ssx + h.ssy);
    a0 = l * h.ar_cb[0];
    a1 = l * h.ar_cr[0];
  }
  tcb[y * TW + x] = (int16_t)clip3(gmin, gmax, tcb[y * TW + x] + round2(a0, h.ar_shift));
  tcr[y * TW + x] = (int16_t)clip3(gmin, gmax, tcr[y * TW + x] + round2(a1, h.ar_shift));
}

/* Lane l receives the value of lane (l - 1) mod 64 (DPP wave_ror:1). */
__device__ __forceinline__ int from_lane_below(int v) {
  return __builtin_amdgcn_update_dpp(0, v, 0x13C, 0xF, 0xF, false);
}

/* Diagonals per round of the AR recursion (one workgroup barrier each). */
constexpr int AR_ROUND = 16;

/* Schedule of the three AR waves: rounds of AR_ROUND diagonals, a barrier
 * after each. Wave 0 (luma) runs its round r in barrier round r. A chroma
 * round j reads luma samples finished by diagonal 2*(16j+15) - 2K - 2
 * (subsampled) or 16j+15 (4:4:4), so it runs in barrier round 2j+2 or j+1,
 * when every luma round it depends on is behind a barrier. */
struct ArPlan {
  int first, every, rounds;
};

__device__ __forceinline__ int ar_rounds(int K, int rows, int cols) {
  return (K * (rows - 1) + cols - 1) / AR_ROUND + 1;
}

/* The auto-regressive filter of 7.18.3.3 (grain.rs:57-73 luma, 101-134
 * chroma) over one template, on one wave. Samples are visited along the
 * diagonals T = K*y + x (K = LAG + 1): every sample (y - d, x + dc) that
 * (y, x) reads lies on an earlier diagonal, T - (d*K - dc).
 *
 * Lane l owns the template rows y = (l + 3) mod 64 and y + 64 and handles,
 * on diagonal T, the sample of its row at x = T - K*y (the two rows of a
 * lane are never both inside the template on one diagonal). Each lane
 * emits one value per diagonal: the filtered sample inside the filtered
 * area, the unfiltered template value on the 3-sample border. It keeps the
 * last 16 values emitted by itself (O) and by the lanes 1, 2 and 3 below
 * it (N1, N2, N3, each filled by one lane rotation of the previous array
 * per diagonal), which are exactly the rows above its row: sample
 * (y - d, x + dc) is N_d[T - (d*K - dc)] (O for d = 0). The taps therefore
 * come from registers, in registers indexed by T mod 16 (the rounds of 16
 * diagonals are unrolled, so every index is a constant).
 *
 * CHROMA adds the tap on the co-located luma average (read from the finished
 * luma template in LDS). coef holds the NL neighbour coefficients (then the
 * luma one). */
template <int LAG, bool CHROMA>
__device__ void ar_wave(int16_t *t, int rows, int cols, const int *coef, int luma_coef, const int16_t *tl, int lsx,
                        int lsy, bool luma_tap, int ars, int gmin, int gmax, ArPlan plan) {
  constexpr int K = LAG + 1, NL = 2 * LAG * K;
  const int lane = threadIdx.x & 63;
  const int ya = (lane + 3) & 63;
  int c[NL > 0 ? NL : 1];
#pragma unroll
  for (int i = 0; i < NL; i++) c[i] = coef[i];
  int O[AR_ROUND], N1[AR_ROUND], N2[AR_ROUND], N3[AR_ROUND];
#pragma unroll
  for (int i = 0; i < AR_ROUND; i++) O[i] = N1[i] = N2[i] = N3[i] = 0;
  const int nchunks = ar_rounds(K, rows, cols);
  int chunk = 0;
  for (int r = 0; r < plan.rounds; r++) {
    if (r >= plan.first && (r - plan.first) % plan.every == 0 && chunk < nchunks) {
      const int T0 = chunk * AR_ROUND;
      int tv[AR_ROUND], lv[AR_ROUND], addr[AR_ROUND];
      bool act[AR_ROUND];
      /* every sample of this round is read before any is written: each
       * template position is written only on its own diagonal */
#pragma unroll
      for (int s = 0; s < AR_ROUND; s++) {
        int x = T0 + s - K * ya, y = ya;
        if (x >= cols) {
          x -= 64 * K;
          y += 64;
        }
        const bool valid = y < rows && x >= 0 && x < cols;
        act[s] = valid && y >= 3 && x >= 3 && x < cols - 3;
        addr[s] = valid ? y * TW + x : 0;
        tv[s] = t[addr[s]];
        lv[s] = 0;
        if (CHROMA && luma_tap) {
          const int ly = act[s] ? ((y - 3) << lsy) + 3 : 0, lx = act[s] ? ((x - 3) << lsx) + 3 : 0;
          const int16_t *q = tl + ly * TW + lx;
          int l = q[0];
          if (lsx) l += q[1];
          if (lsy) l += q[TW];
          if (lsx && lsy) l += q[TW + 1];
          lv[s] = round2(l, lsx + lsy);
        }
      }
#pragma unroll
      for (int s = 0; s < AR_ROUND; s++) {
        int acc[4] = {0, 0, 0, 0};
        int pos = 0;
#pragma unroll
        for (int d = LAG; d >= 0; d--)
#pragma unroll
          for (int dc = -LAG; dc <= (d ? LAG : -1); dc++) {
            const int idx = (s - (d * K - dc)) & (AR_ROUND - 1);
            const int v = d == 0 ? O[idx] : d == 1 ? N1[idx] : d == 2 ? N2[idx] : N3[idx];
            acc[pos & 3] += __mul24(v, c[pos]);
            pos++;
          }
        int sum = (acc[0] + acc[1]) + (acc[2] + acc[3]);
        if (CHROMA && luma_tap) sum += __mul24(lv[s], luma_coef);
        const int out = act[s] ? clip3(gmin, gmax, tv[s] + round2(sum, ars)) : tv[s];
        if (act[s]) t[addr[s]] = (int16_t)out;
        O[s] = out;
        N1[s] = from_lane_below(O[s]);
        N2[s] = from_lane_below(N1[s]);
        N3[s] = from_lane_below(N2[s]);
      }
      chunk++;
    }
    __syncthreads();
  }
}

/* The AR stage of the three templates: wave 0 luma, waves 1 and 2 cb and
 * cr (on the schedule of ArPlan). Lag 0 has no recursion: luma samples are
 * only clipped and each chroma sample depends on the finished luma alone. */
template <int LAG>
__device__ void ar_all(int16_t *tl, int16_t *tcb, int16_t *tcr, const GrainHdr &h, bool chroma, int cw, int ch,
                       int gmin, int gmax) {
  constexpr int K = LAG + 1;
  const int wave = threadIdx.x >> 6;
  if constexpr (LAG == 0) {
    for (int i = threadIdx.x; i < 70 * 76; i += blockDim.x) {
      const int y = 3 + i / 76, x = 3 + i % 76;
      tl[y * TW + x] = (int16_t)clip3(gmin, gmax, tl[y * TW + x]);
    }
    __syncthreads();
    if (chroma) {
      const int rw = cw - 6, nr = ch - 3;
      for (int i = threadIdx.x; i < rw * nr; i += blockDim.x) chroma_lag0(tcb, tcr, tl, h, 3 + i / rw, 3 + i % rw, gmin, gmax);
    }
    __syncthreads();
  } else {
  constexpr int NL = 2 * LAG * K;
  const int sub = h.ssx | h.ssy;
  const int lr = ar_rounds(K, TH, TW);
  const int cr = chroma ? ar_rounds(K, ch, cw) : 0;
  const ArPlan pl_luma = {0, 1, 0}, pl_chroma = {sub ? 2 : 1, sub ? 2 : 1, 0};
  const int rounds = max(lr, chroma ? pl_chroma.first + pl_chroma.every * (cr - 1) + 1 : 0);
  if (wave == 0) {
    ArPlan p = pl_luma;
    p.rounds = rounds;
    ar_wave<LAG, false>(tl, TH, TW, h.ar_y, 0, tl, 0, 0, false, h.ar_shift, gmin, gmax, p);
  } else if (wave == 1 || wave == 2) {
    ArPlan p = pl_chroma;
    p.rounds = chroma ? rounds : 0;
    const int *cf = wave == 1 ? h.ar_cb : h.ar_cr;
    int16_t *t = wave == 1 ? tcb : tcr;
    if (chroma)
      ar_wave<LAG, true>(t, ch, cw, cf, cf[NL], tl, h.ssx, h.ssy, h.num_y > 0, h.ar_shift, gmin, gmax, p);
    else
      for (int r = 0; r < rounds; r++) __syncthreads();
  }
  }
}

struct PrepShared {
  int16_t tl[TH * TW], tcb[TH * TW], tcr[TH * TW];
  int16_t gauss[2048];
};

/* Templates, LUTs and block offsets of one frame (PREP_THREADS threads). */
__device__ __forceinline__ void grain_prepare_body(const GrainHdr &h, const GrainScratch &sc, PrepShared &sm) {
  int16_t *tl = sm.tl, *tcb = sm.tcb, *tcr = sm.tcr, *gauss = sm.gauss;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < 2048; i += PREP_THREADS) gauss[i] = AV1_GAUSSIAN_SEQUENCE[i];
  __syncthreads();
  const int bd = h.bd;
  const int gmin = -(128 << (bd - 8)), gmax = (256 << (bd - 8)) - 1 - (128 << (bd - 8));
  const int shift = 12 - bd + h.grain_scale_shift;
  const bool chroma = h.nplanes > 1;
  const int cw = h.ssx ? 44 : 82, ch = h.ssy ? 38 : 73;
  if (wave == 0)
    gaussian_fill(tl, gauss, h.seed, TH, TW, shift, h.num_y > 0, lane);
  else if (wave == 1 && chroma)
    gaussian_fill(tcb, gauss, h.seed ^ 0xb524, ch, cw, shift, h.num_cb > 0 || h.cfl, lane);
  else if (wave == 2 && chroma)
    gaussian_fill(tcr, gauss, h.seed ^ 0x49d8, ch, cw, shift, h.num_cr > 0 || h.cfl, lane);
  __syncthreads();
  switch (h.ar_lag) {
    case 0: ar_all<0>(tl, tcb, tcr, h, chroma, cw, ch, gmin, gmax); break;
    case 1: ar_all<1>(tl, tcb, tcr, h, chroma, cw, ch, gmin, gmax); break;
    case 2: ar_all<2>(tl, tcb, tcr, h, chroma, cw, ch, gmin, gmax); break;
    default: ar_all<3>(tl, tcb, tcr, h, chroma, cw, ch, gmin, gmax); break;
  }
  /* templates out */
  for (int i = threadIdx.x; i < TH * TW; i += PREP_THREADS) {
    sc.tmpl[i] = tl[i];
    sc.tmpl[TH * TW + i] = chroma ? tcb[i] : 0;
    sc.tmpl[2 * TH * TW + i] = chroma ? tcr[i] : 0;
  }
  /* scaling LUTs (grain.rs:137-165) */
  for (int p = 0; p < h.nplanes; p++) {
    const int np = h.npoints[p];
    const int *xs = h.points[p][0], *ys = h.points[p][1];
    for (int x = threadIdx.x; x < 256; x += PREP_THREADS) {
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
  for (int ln = threadIdx.x; ln < nstripes; ln += PREP_THREADS) {
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

__global__ __launch_bounds__(PREP_THREADS) void grain_prepare_kernel(GrainHdr h, GrainScratch sc) {
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

__global__ __launch_bounds__(PREP_THREADS) void grain_prepare_batch(BatchWork w) {
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
  grain_prepare_kernel<<<1, PREP_THREADS, 0, st>>>(h, s);
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
  grain_prepare_batch<<<sp.grid, PREP_THREADS, 0, st>>>(w);
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
