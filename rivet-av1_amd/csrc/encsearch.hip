/*
 * encsearch.hip — the encoder's loop filter level search and CDEF strength
 * search on the GPU: rivet-av1's search_lf (src/encoder/mod.rs:1460-1490)
 * and cdef::search (src/encoder/cdef.rs:52-253).
 *
 * The filtering and the error sums run on the device with the decoder's
 * kernels (launch_deblock, and cdef.hip's search kernel, which applies
 * every candidate strength to every 8x8 block). What is left is a handful
 * of decisions taken on the host between launches:
 *   - search_lf: six scales of the start levels; per scale the
 *     reconstruction is copied, deblocked and its squared error per plane
 *     summed (sse_kernel, exact 64-bit sums); each plane keeps the first
 *     scale with the least error.
 *   - cdef::search: the greedy choice of up to eight (luma, chroma)
 *     strength pairs. Each step evaluates every pair not yet chosen at once
 *     (greedy_eval: one workgroup per pair, the total over the coded 64x64
 *     blocks of min(current best, pair error)), the host takes the first
 *     least total, and greedy_update folds it into the per-block bests. The
 *     reference sums block errors in double precision; the totals here are
 *     exact 64-bit integers, which equal those sums while they stay below
 *     2^53 (squared errors of frames up to ~3.6e8 samples).
 */
#include <float.h>
#include <math.h>
#include <string.h>

#include <vector>

#include "rvpf_host.h"

namespace rvpf {

namespace {

constexpr int NT = 256;

/* Sums of a u64 value over the workgroup into *out (thread 0). */
__device__ __forceinline__ uint64_t block_sum(uint64_t x, uint64_t *part) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = x;
  __syncthreads();
  uint64_t s = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < NT / 64; w++) s += part[w];
  return s;
}

/* ---- squared error of planes (plane_sse, src/encoder/mod.rs:1492-1505) ---- */

struct SseLaunch {
  PlanePtr a[3], b[3];
  int w[3], h[3];
  int blk_x[3];
  int blk_start[4];
  uint64_t *out; /* [3], added to */
};

constexpr int SSE_ROWS = 8;

/* Tile: 256 columns x SSE_ROWS rows of one plane, one column per thread. */
template <typename T>
__global__ __launch_bounds__(NT) void sse_kernel(SseLaunch a) {
  const int bid = blockIdx.x;
  int p = 0;
  if (bid >= a.blk_start[1]) p = 1;
  if (bid >= a.blk_start[2]) p = 2;
  const int lb = bid - a.blk_start[p];
  const int x = (lb % a.blk_x[p]) * NT + threadIdx.x, y0 = (lb / a.blk_x[p]) * SSE_ROWS;
  uint32_t e = 0; /* SSE_ROWS squares of at most 4095^2 */
  if (x < a.w[p]) {
    const T *pa = gptr(reinterpret_cast<const T *>(a.a[p].data));
    const T *pb = gptr(reinterpret_cast<const T *>(a.b[p].data));
    int va[SSE_ROWS], vb[SSE_ROWS];
#pragma unroll
    for (int r = 0; r < SSE_ROWS; r++) {
      const int y = min(y0 + r, a.h[p] - 1);
      va[r] = pa[(int64_t)y * a.a[p].stride + x];
      vb[r] = pb[(int64_t)y * a.b[p].stride + x];
    }
#pragma unroll
    for (int r = 0; r < SSE_ROWS; r++) {
      const int d = va[r] - vb[r];
      if (y0 + r < a.h[p]) e += (uint32_t)(d * d);
    }
  }
  __shared__ uint64_t part[NT / 64];
  const uint64_t s = block_sum(e, part);
  if (threadIdx.x == 0) atomicAdd(reinterpret_cast<unsigned long long *>(&gptr(a.out)[p]), (unsigned long long)s);
}

int launch_sse(const Geom &g, const PlanePtr a[3], const PlanePtr b[3], const int w[3], const int h[3], uint64_t *out,
               int bps, hipStream_t st) {
  SseLaunch l;
  int total = 0;
  for (int p = 0; p < 3; p++) {
    const bool on = p < g.nplanes && w[p] > 0 && h[p] > 0;
    l.a[p] = a[p];
    l.b[p] = b[p];
    l.w[p] = on ? w[p] : 0;
    l.h[p] = on ? h[p] : 0;
    l.blk_x[p] = on ? (w[p] + NT - 1) / NT : 1;
    l.blk_start[p] = total;
    if (on) total += l.blk_x[p] * ((h[p] + SSE_ROWS - 1) / SSE_ROWS);
  }
  l.blk_start[3] = total;
  l.out = out;
  if (total == 0) return 0;
  if (bps == 1)
    sse_kernel<uint8_t><<<total, NT, 0, st>>>(l);
  else
    sse_kernel<uint16_t><<<total, NT, 0, st>>>(l);
  return (int)hipGetLastError();
}

__global__ __launch_bounds__(NT) void fill_u64(uint64_t *p, int n, uint64_t v) {
  const int i = blockIdx.x * NT + threadIdx.x;
  if (i < n) gptr(p)[i] = v;
}

int launch_fill(uint64_t *p, int n, uint64_t v, hipStream_t st) {
  fill_u64<<<(n + NT - 1) / NT, NT, 0, st>>>(p, n, v);
  return (int)hipGetLastError();
}

/* ---- the greedy choice of cdef::search (cdef.rs:199-236) ---- */

/* Workgroup yk * nuv + uk: the frame total with the pair added to the set
 * whose per-block bests are cur (UINT64_MAX for the empty set). */
__global__ __launch_bounds__(NT) void greedy_eval(const uint64_t *ys, const uint64_t *uvs, const int8_t *coded,
                                                  const uint64_t *cur, int nfb, int ny, int nuv, uint64_t *tot) {
  const int yk = blockIdx.x / nuv, uk = blockIdx.x - yk * nuv;
  uint64_t acc = 0;
  for (int b = threadIdx.x; b < nfb; b += NT)
    if (gptr(coded)[b]) {
      const uint64_t v = gptr(ys)[(int64_t)b * ny + yk] + gptr(uvs)[(int64_t)b * nuv + uk];
      acc += min(gptr(cur)[b], v);
    }
  __shared__ uint64_t part[NT / 64];
  const uint64_t s = block_sum(acc, part);
  if (threadIdx.x == 0) gptr(tot)[blockIdx.x] = s;
}

__global__ __launch_bounds__(NT) void greedy_update(const uint64_t *ys, const uint64_t *uvs, const int8_t *coded,
                                                    uint64_t *cur, int nfb, int ny, int nuv, int yk, int uk) {
  const int b = blockIdx.x * NT + threadIdx.x;
  if (b >= nfb || !gptr(coded)[b]) return;
  const uint64_t v = gptr(ys)[(int64_t)b * ny + yk] + gptr(uvs)[(int64_t)b * nuv + uk];
  gptr(cur)[b] = min(gptr(cur)[b], v);
}

struct PairSet {
  int sy[8], su[8];
  int n;
};

/* best_for's pick (cdef.rs:201-216): each coded block's first pair of the
 * set with its least error. */
__global__ __launch_bounds__(NT) void greedy_pick(const uint64_t *ys, const uint64_t *uvs, const int8_t *coded,
                                                  int nfb, int ny, int nuv, PairSet set, int8_t *pick) {
  const int b = blockIdx.x * NT + threadIdx.x;
  if (b >= nfb) return;
  int8_t best_i = -1;
  if (gptr(coded)[b]) {
    uint64_t best = UINT64_MAX;
    for (int i = 0; i < set.n; i++) {
      const uint64_t v = gptr(ys)[(int64_t)b * ny + set.sy[i]] + gptr(uvs)[(int64_t)b * nuv + set.su[i]];
      if (v < best) {
        best = v;
        best_i = (int8_t)i;
      }
    }
  }
  gptr(pick)[b] = best_i;
}

/* The candidates of cdef.rs:60-84: (primary, secondary as coded). */
void cdef_candidates(bool thorough, std::vector<int> &yp, std::vector<int> &ys, std::vector<int> &up,
                     std::vector<int> &us) {
  static const int yp_n[] = {0, 1, 2, 4, 6, 9, 12}, yp_t[] = {0, 1, 2, 3, 4, 5, 6, 8, 10, 12, 15};
  static const int up_n[] = {0, 1, 2, 4, 7}, up_t[] = {0, 1, 2, 3, 4, 6, 8, 11};
  const int *ypl = thorough ? yp_t : yp_n, *upl = thorough ? up_t : up_n;
  const int nyp = thorough ? 11 : 7, nup = thorough ? 8 : 5;
  for (int i = 0; i < nyp; i++)
    for (int j = 0; j < 4; j++) {
      yp.push_back(ypl[i]);
      ys.push_back(j);
    }
  for (int i = 0; i < nup; i++)
    for (int j = 0; j < 3; j++) {
      up.push_back(upl[i]);
      us.push_back(j);
    }
}

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

}  // namespace
}  // namespace rvpf

using namespace rvpf;
using namespace rvpf::host;

namespace {
/* Releases the search scratch on every return path, so a later call on
 * another stream waits for whatever this call left in flight. */
struct SearchGateGuard {
  rvpf_ctx *ctx;
  hipStream_t st;
  ~SearchGateGuard() { (void)scratch_release(ctx, &ctx->search_gate, st); }
};
}  // namespace

extern "C" {

int rvpf_lf_search(rvpf_ctx *ctx, const rvpf_frame_params *p, const rvpf_side_info *s, const rvpf_planes *recon,
                   const rvpf_planes *source, int32_t levels[4], void *stream) {
  if (!ctx) return fail(nullptr, RVPF_E_INVALID, "ctx is NULL");
  int rc;
  if ((rc = validate_params(p, ctx->err)) || (rc = validate_side(p, s, ctx->err)) ||
      (rc = validate_planes(p, recon, "recon", 0, 0, ctx->err)) ||
      (rc = validate_planes(p, source, "source", 1, 0, ctx->err)))
    return rc;
  if (p->use_superres) return fail(ctx, RVPF_E_UNSUPPORTED, "the searches run on frames without super-resolution");
  if (!levels) return fail(ctx, RVPF_E_INVALID, "levels is NULL");
  if (recon->bytes_per_sample != source->bytes_per_sample)
    return fail(ctx, RVPF_E_INVALID, "recon and source sample sizes differ");
  const int bps = recon->bytes_per_sample;
  const Geom g = make_geom(p);
  hipStream_t st = (hipStream_t)stream;
  /* work planes: the mode-info area plus the deblocking margin */
  int aw[3] = {0, 0, 0}, ah[3] = {0, 0, 0}, ww[3] = {0, 0, 0}, wh[3] = {0, 0, 0};
  for (int i = 0; i < g.nplanes; i++) {
    area(p, i, &aw[i], &ah[i]);
    ww[i] = aw[i] + 8;
    wh[i] = ah[i] + 8;
  }
  const size_t plane_bytes = layout_planes(ww, wh, bps, nullptr, nullptr);
  if ((rc = ensure_buf(ctx, &ctx->search_buf, &ctx->search_bytes, plane_bytes + 256, "search scratch"))) return rc;
  if ((rc = scratch_acquire(ctx, &ctx->search_gate, st))) return rc;
  SearchGateGuard guard{ctx, st};
  PlanePtr work[3] = {}, rec[3], src[3];
  layout_planes(ww, wh, bps, ctx->search_buf, work);
  uint64_t *sse = reinterpret_cast<uint64_t *>((char *)ctx->search_buf + plane_bytes);
  to_ptrs(recon, rec);
  to_ptrs(source, src);
  LfSide side;
  side.mi = s->mi;
  side.mi_stride = s->mi_stride;
  for (int i = 0; i < 3; i++) side.tx[i] = s->tx_size[i];
  int cw[3] = {0, 0, 0}, ch[3] = {0, 0, 0};
  for (int i = 0; i < g.nplanes; i++) crop(p, i, &cw[i], &ch[i]);
  if ((rc = check_launch(ctx, launch_fill(sse, 18, 0, st), "fill"))) return rc;
  /* search_lf's scales; levels rounded half away from zero, at most 63 */
  static const double scales[6] = {0.0, 0.5, 0.75, 1.0, 1.25, 1.5};
  int lv[6][4];
  for (int k = 0; k < 6; k++) {
    rvpf_frame_params q = *p;
    for (int i = 0; i < 4; i++) {
      lv[k][i] = (int)round((double)p->loop_filter_level[i] * scales[k]);
      if (lv[k][i] > 63) lv[k][i] = 63;
      q.loop_filter_level[i] = lv[k][i];
    }
    if ((rc = check_launch(ctx, launch_copy_planes(g, rec, work, aw, ah, bps, st), "search copy"))) return rc;
    if (lv[k][0] != 0 || lv[k][1] != 0)
      if ((rc = check_launch(ctx, launch_deblock(g, make_lf(&q), side, work, bps, st), "search deblock"))) return rc;
    if ((rc = check_launch(ctx, launch_sse(g, work, src, cw, ch, sse + 3 * k, bps, st), "search sse"))) return rc;
  }
  uint64_t h_sse[18];
  hipError_t e = hipMemcpyAsync(h_sse, sse, sizeof h_sse, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(ctx, e, "search readback");
  uint64_t best_e[3] = {UINT64_MAX, UINT64_MAX, UINT64_MAX};
  int best_l[3] = {0, 0, 0};
  for (int k = 0; k < 6; k++)
    for (int pl = 0; pl < g.nplanes; pl++)
      if (h_sse[3 * k + pl] < best_e[pl]) {
        best_e[pl] = h_sse[3 * k + pl];
        best_l[pl] = pl == 0 ? lv[k][0] : lv[k][1 + pl];
      }
  if (best_l[0] == 0) {
    levels[0] = levels[1] = levels[2] = levels[3] = 0;
  } else {
    levels[0] = levels[1] = best_l[0];
    levels[2] = best_l[1];
    levels[3] = best_l[2];
  }
  return RVPF_OK;
}

int rvpf_cdef_search(rvpf_ctx *ctx, const rvpf_frame_params *p, const rvpf_side_info *s,
                     const rvpf_planes *deblocked, const rvpf_planes *source, int32_t base_q_idx, double lambda,
                     int32_t thorough, rvpf_cdef_search_result *out, int8_t *cdef_idx_out, void *stream) {
  if (!ctx) return fail(nullptr, RVPF_E_INVALID, "ctx is NULL");
  int rc;
  if ((rc = validate_params(p, ctx->err)) || (rc = validate_planes(p, deblocked, "deblocked", 0, 0, ctx->err)) ||
      (rc = validate_planes(p, source, "source", 0, 0, ctx->err, true)))
    return rc;
  if (p->use_superres) return fail(ctx, RVPF_E_UNSUPPORTED, "the searches run on frames without super-resolution");
  const int fb_rows = (p->mi_rows + 15) / 16, fb_cols = (p->mi_cols + 15) / 16;
  if (!s || !s->mi || s->mi_stride < p->mi_cols + 1 || !s->cdef_idx || s->cdef_stride < fb_cols)
    return fail(ctx, RVPF_E_INVALID, "side info: mi and cdef_idx are needed");
  if (!out || !cdef_idx_out) return fail(ctx, RVPF_E_INVALID, "out or cdef_idx_out is NULL");
  if (base_q_idx < 0 || base_q_idx > 255) return fail(ctx, RVPF_E_INVALID, "base_q_idx out of range");
  if (!(lambda >= 0.0)) return fail(ctx, RVPF_E_INVALID, "lambda must be non-negative");
  if (deblocked->bytes_per_sample != source->bytes_per_sample)
    return fail(ctx, RVPF_E_INVALID, "deblocked and source sample sizes differ");
  const int bps = deblocked->bytes_per_sample;
  const Geom g = make_geom(p);
  hipStream_t st = (hipStream_t)stream;
  std::vector<int> yp, ys, up, us;
  cdef_candidates(thorough != 0, yp, ys, up, us);
  const int ny = (int)yp.size(), nuv = (int)up.size(), npair = ny * nuv;
  const int nfb = fb_rows * fb_cols;
  /* scratch: per-block errors, the per-block bests, the pair totals,
   * coded flags and picks */
  const size_t o_ys = 0, o_uvs = align256(o_ys + (size_t)nfb * ny * 8), o_cur = align256(o_uvs + (size_t)nfb * nuv * 8),
               o_tot = align256(o_cur + (size_t)nfb * 8), o_coded = align256(o_tot + (size_t)npair * 8),
               o_pick = align256(o_coded + (size_t)nfb), bytes = align256(o_pick + (size_t)nfb);
  if ((rc = ensure_buf(ctx, &ctx->search_buf, &ctx->search_bytes, bytes, "search scratch"))) return rc;
  if ((rc = scratch_acquire(ctx, &ctx->search_gate, st))) return rc;
  SearchGateGuard guard{ctx, st};
  char *base = (char *)ctx->search_buf;
  uint64_t *d_ys = (uint64_t *)(base + o_ys), *d_uvs = (uint64_t *)(base + o_uvs), *d_cur = (uint64_t *)(base + o_cur),
           *d_tot = (uint64_t *)(base + o_tot);
  int8_t *d_coded = (int8_t *)(base + o_coded), *d_pick = (int8_t *)(base + o_pick);
  std::vector<int8_t> hy(2 * ny), hu(2 * nuv);
  for (int k = 0; k < ny; k++) {
    hy[k] = (int8_t)yp[k];
    hy[ny + k] = (int8_t)(ys[k] == 3 ? 4 : ys[k]);
  }
  for (int k = 0; k < nuv; k++) {
    hu[k] = (int8_t)up[k];
    hu[nuv + k] = (int8_t)(us[k] == 3 ? 4 : us[k]);
  }
  const int dm3 = base_q_idx >> 6 < 3 ? base_q_idx >> 6 : 3; /* damping_for, cdef.rs:256-258 */
  PlanePtr in[3], src[3];
  to_ptrs(deblocked, in);
  to_ptrs(source, src);
  if ((rc = check_launch(ctx,
                         launch_cdef_search(g, s->mi, s->mi_stride, s->cdef_idx, s->cdef_stride, in, src, dm3 + 3,
                                            hy.data(), hy.data() + ny, ny, hu.data(), hu.data() + nuv, nuv, d_ys, d_uvs,
                                            d_coded, bps, st),
                         "cdef search")))
    return rc;
  std::vector<int8_t> coded(nfb);
  hipError_t e = hipMemcpyAsync(coded.data(), d_coded, nfb, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess)
    e = hipMemcpyAsync(cdef_idx_out, s->cdef_idx, (size_t)fb_rows * s->cdef_stride, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(ctx, e, "search readback");
  int nb = 0;
  for (int b = 0; b < nfb; b++) nb += coded[b] != 0;
  memset(out, 0, sizeof *out);
  if (nb == 0) return RVPF_OK; /* CdefParams::off */
  if ((rc = check_launch(ctx, launch_fill(d_cur, nfb, UINT64_MAX, st), "fill"))) return rc;
  PairSet set;
  memset(&set, 0, sizeof set);
  std::vector<uint64_t> tot(npair);
  double best_cost = DBL_MAX;
  int best_n = 0;
  for (int bits = 0; bits <= 3; bits++) {
    uint64_t t = 0;
    while (set.n < (1 << bits)) {
      greedy_eval<<<npair, NT, 0, st>>>(d_ys, d_uvs, d_coded, d_cur, nfb, ny, nuv, d_tot);
      if ((rc = check_launch(ctx, (int)hipGetLastError(), "greedy eval"))) return rc;
      e = hipMemcpyAsync(tot.data(), d_tot, (size_t)npair * 8, hipMemcpyDeviceToHost, st);
      if (e == hipSuccess) e = hipStreamSynchronize(st);
      if (e != hipSuccess) return hip_fail(ctx, e, "search readback");
      int add = -1;
      for (int pr = 0; pr < npair; pr++) {
        bool in_set = false;
        for (int i = 0; i < set.n; i++) in_set |= set.sy[i] * nuv + set.su[i] == pr;
        if (!in_set && (add < 0 || tot[pr] < tot[add])) add = pr;
      }
      set.sy[set.n] = add / nuv;
      set.su[set.n] = add % nuv;
      set.n++;
      t = tot[add];
      greedy_update<<<(nfb + NT - 1) / NT, NT, 0, st>>>(d_ys, d_uvs, d_coded, d_cur, nfb, ny, nuv, add / nuv, add % nuv);
      if ((rc = check_launch(ctx, (int)hipGetLastError(), "greedy update"))) return rc;
    }
    const double cost = (double)t + lambda * (double)bits * (double)nb;
    if (cost < best_cost) {
      best_cost = cost;
      best_n = set.n;
    }
  }
  PairSet best = set;
  best.n = best_n;
  greedy_pick<<<(nfb + NT - 1) / NT, NT, 0, st>>>(d_ys, d_uvs, d_coded, nfb, ny, nuv, best, d_pick);
  if ((rc = check_launch(ctx, (int)hipGetLastError(), "greedy pick"))) return rc;
  std::vector<int8_t> pick(nfb);
  e = hipMemcpyAsync(pick.data(), d_pick, nfb, hipMemcpyDeviceToHost, st);
  if (e == hipSuccess) e = hipStreamSynchronize(st);
  if (e != hipSuccess) return hip_fail(ctx, e, "search readback");
  for (int b = 0; b < nfb; b++)
    if (coded[b]) cdef_idx_out[(b / fb_cols) * s->cdef_stride + b % fb_cols] = pick[b];
  out->damping_minus_3 = dm3;
  out->bits = best_n == 1 ? 0 : (best_n == 2 ? 1 : (best_n == 4 ? 2 : 3));
  for (int i = 0; i < best_n; i++) {
    out->y_pri[i] = yp[best.sy[i]];
    out->y_sec[i] = ys[best.sy[i]];
    out->uv_pri[i] = up[best.su[i]];
    out->uv_sec[i] = us[best.su[i]];
  }
  return RVPF_OK;
}

}  // extern "C"
