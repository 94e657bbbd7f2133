/*
 * lrdiff.hip — float-relaxed loop restoration and its adjoint, for the
 * encoder's restoration search (rivet-av1 src/encoder/lr.rs:64-430).
 *
 * The encoder fits, per restoration unit, Wiener taps and self-guided
 * projection weights that minimise the squared error against the source.
 * Here the filters are relaxed to real arithmetic (no rounding, no clip):
 *
 *   Wiener (7.17.4 without rounding): H = W + sum_k h_k (W<<d + W>>d - 2W),
 *     out = H + sum_k v_k (H^d + H_d - 2H), d = 3 - k, taps as fractions of
 *     128 (k = 0 outermost), horizontal on the window W, then vertical. This
 *     is the model fit_wiener solves (lr.rs:240-325).
 *   Self-guided (7.17.3 without rounding): box sums in real arithmetic,
 *     z = p s / 2^20, A = 256 z / (z + 1) (256 from z >= 255), B = (256 - A)
 *     mean, the 3x3 weighting; projection out = u + (w0 (f0 - 16u) + w2
 *     (f1 - 16u)) / 2048 (7.17.2 without rounding).
 *
 * Work items are rectangles (one stripe of one unit, at most 64 x 64): a
 * workgroup of 256 threads per item. Forward kernels write the residual
 * out - src and the item's squared error (double); backward kernels give
 * the item's gradient of its squared error with respect to the unit's taps
 * (double), which the caller sums per unit.
 *
 * Wiener adjoint: with e = 2 (out - src),
 *   dSSE/dv_k = sum e (H(i+3-d) + H(i+3+d) - 2 H(i+3))
 *   dSSE/dh_k = sum_r E(r) (W(c+3-d) + W(c+3+d) - 2 W(c+3)),
 *   E(r, j)   = sum_t V_t e(r - t, j), V = (v0, v1, v2, 1 - 2 sum v, v2, v1, v0)
 * (E is the vertical filter's transpose applied to e).
 */
#include "av1_tables.h"
#include "rvpf_host.h"

namespace rvpf {

namespace {

constexpr int NTH = 256;
constexpr int T = 64;      /* item tile */
constexpr int WS = T + 6;  /* window side */

__device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }
__device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }

/* An item as the kernels use it: sizes clamped to the tile, and an item
 * whose unit or plane is out of range treated as empty. */
__device__ __forceinline__ rvpf_lrd_item load_item(const rvpf_lrd_item *items, int b, int nunits) {
  rvpf_lrd_item it = items[b];
  it.w = imin(imax(it.w, 0), T);
  it.h = imin(imax(it.h, 0), T);
  if (it.unit < 0 || it.unit >= nunits || it.plane < 0 || it.plane > 2) {
    it.w = it.h = 0;
    it.unit = 0;
    it.plane = 0;
  }
  return it;
}

/* Sum over the workgroup of one double per thread (NTH threads). */
__device__ double block_sum(double v, double *red) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) red[wave] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < NTH / 64; i++) s += red[i];
  __syncthreads();
  return s;
}

struct GatherArgs {
  PlanePtr deb[3], cdef[3], src[3];
  int bps;
  int plane_end_x[3], plane_end_y[3];
  const rvpf_lrd_item *items;
  int nunits;
  float *win, *u, *s;
};

template <typename TS>
__global__ __launch_bounds__(NTH) void gather_kernel(GatherArgs a) {
  const rvpf_lrd_item it = load_item(a.items, blockIdx.x, a.nunits);
  const int p = it.plane;
  const TS *deb = reinterpret_cast<const TS *>(a.deb[p].data);
  const TS *cdf = reinterpret_cast<const TS *>(a.cdef[p].data);
  const TS *src = reinterpret_cast<const TS *>(a.src[p].data);
  const int64_t ds = a.deb[p].stride, cs = a.cdef[p].stride, ss = a.src[p].stride;
  const int pex = a.plane_end_x[p], pey = a.plane_end_y[p];
  float *win = a.win + (size_t)blockIdx.x * WS * WS;
  float *u = a.u + (size_t)blockIdx.x * T * T;
  float *s = a.s + (size_t)blockIdx.x * T * T;
  /* LrCtx::sample (postfilter.rs:584-596): clamp to the plane, rows
   * beyond the stripe from the deblocked frame, at most 2 rows out */
  for (int k = threadIdx.x; k < WS * WS; k += NTH) {
    const int r = k / WS, c = k - r * WS;
    float v = 0.f;
    if (r < it.h + 6 && c < it.w + 6) {
      const int x = imax(imin(it.x0 + c - 3, pex), 0);
      int y = imax(imin(it.y0 + r - 3, pey), 0);
      /* (the final clamp only matters for items outside their plane) */
      if (y < it.stripe_start) {
        y = imax(imin(imax(it.stripe_start - 2, y), pey), 0);
        v = (float)deb[(int64_t)y * ds + x];
      } else if (y > it.stripe_end) {
        y = imax(imin(imin(it.stripe_end + 2, y), pey), 0);
        v = (float)deb[(int64_t)y * ds + x];
      } else {
        v = (float)cdf[(int64_t)y * cs + x];
      }
    }
    win[k] = v;
  }
  for (int k = threadIdx.x; k < T * T; k += NTH) {
    const int i = k >> 6, j = k & 63;
    const bool in = i < it.h && j < it.w && it.y0 + i >= 0 && it.y0 + i <= pey && it.x0 + j >= 0 && it.x0 + j <= pex;
    u[k] = in ? (float)cdf[(int64_t)(it.y0 + i) * cs + it.x0 + j] : 0.f;
    s[k] = in ? (float)src[(int64_t)(it.y0 + i) * ss + it.x0 + j] : 0.f;
  }
}

/* LDS of the Wiener kernels. */
struct WienerShared {
  float w[WS * WS];
  float hbuf[WS * T]; /* H; the backward pass reuses it for E */
  float e[T * T];
  double red[NTH / 64];
};

/* H(r, c) = W(r, c+3) + sum_k h_k (W(r, c+3-d) + W(r, c+3+d) - 2 W(r, c+3)). */
__device__ __forceinline__ void wiener_h(const WienerShared &sm, float *hb, const float h[3], int rows, int w) {
  for (int k = threadIdx.x; k < rows * T; k += NTH) {
    const int r = k >> 6, c = k & 63;
    float v = 0.f;
    if (c < w) {
      const float *row = sm.w + r * WS + c + 3;
      const float ctr = row[0];
      v = ctr + h[0] * (row[-3] + row[3] - 2.f * ctr) + h[1] * (row[-2] + row[2] - 2.f * ctr) +
          h[2] * (row[-1] + row[1] - 2.f * ctr);
    }
    hb[k] = v;
  }
}

struct WienerArgs {
  const rvpf_lrd_item *items;
  const float *win, *src;
  const float *taps; /* [unit][2][3]: vertical, horizontal */
  int nunits;
  float *resid;
  double *out;       /* forward: sse [n]; backward: grad [n][6] (v0..v2, h0..h2) */
};

__device__ __forceinline__ void load_window(WienerShared &sm, const float *win) {
  for (int k = threadIdx.x; k < WS * WS; k += NTH) sm.w[k] = win[k];
}

/* The forward pass runs in double: its error is differenced against
 * itself by finite-difference checks at small tap steps, so its rounding
 * must stay far below the change a step of 1e-3 makes. */
struct ForwardShared {
  float w[WS * WS];
  double hbuf[WS * T];
  double red[NTH / 64];
};

__global__ __launch_bounds__(NTH) void wiener_forward_kernel(WienerArgs a) {
  __shared__ ForwardShared sm;
  const rvpf_lrd_item it = load_item(a.items, blockIdx.x, a.nunits);
  const float *tp = a.taps + (size_t)it.unit * 6;
  const double v[3] = {tp[0], tp[1], tp[2]}, h[3] = {tp[3], tp[4], tp[5]};
  const float *win = a.win + (size_t)blockIdx.x * WS * WS;
  for (int k = threadIdx.x; k < WS * WS; k += NTH) sm.w[k] = win[k];
  __syncthreads();
  for (int k = threadIdx.x; k < (it.h + 6) * T; k += NTH) {
    const int r = k >> 6, c = k & 63;
    double val = 0.0;
    if (c < it.w) {
      const float *row = sm.w + r * WS + c + 3;
      const double ctr = row[0];
      val = ctr + h[0] * ((double)row[-3] + row[3] - 2.0 * ctr) + h[1] * ((double)row[-2] + row[2] - 2.0 * ctr) +
            h[2] * ((double)row[-1] + row[1] - 2.0 * ctr);
    }
    sm.hbuf[k] = val;
  }
  __syncthreads();
  const float *src = a.src + (size_t)blockIdx.x * T * T;
  float *res = a.resid + (size_t)blockIdx.x * T * T;
  double acc = 0.0;
  for (int k = threadIdx.x; k < T * T; k += NTH) {
    const int i = k >> 6, j = k & 63;
    double e = 0.0;
    if (i < it.h && j < it.w) {
      const double *col = sm.hbuf + (i + 3) * T + j;
      const double ctr = col[0];
      const double out = ctr + v[0] * (col[-3 * T] + col[3 * T] - 2.0 * ctr) +
                         v[1] * (col[-2 * T] + col[2 * T] - 2.0 * ctr) + v[2] * (col[-T] + col[T] - 2.0 * ctr);
      e = out - (double)src[k];
      acc += e * e;
    }
    res[k] = (float)e;
  }
  const double s = block_sum(acc, sm.red);
  if (threadIdx.x == 0) a.out[blockIdx.x] = s;
}

__global__ __launch_bounds__(NTH) void wiener_backward_kernel(WienerArgs a) {
  __shared__ WienerShared sm;
  const rvpf_lrd_item it = load_item(a.items, blockIdx.x, a.nunits);
  const float *tp = a.taps + (size_t)it.unit * 6;
  const float v[3] = {tp[0], tp[1], tp[2]}, h[3] = {tp[3], tp[4], tp[5]};
  const float vc = 1.f - 2.f * (v[0] + v[1] + v[2]);
  load_window(sm, a.win + (size_t)blockIdx.x * WS * WS);
  const float *res = a.resid + (size_t)blockIdx.x * T * T;
  for (int k = threadIdx.x; k < T * T; k += NTH) sm.e[k] = 2.f * res[k]; /* zero outside the item */
  __syncthreads();
  wiener_h(sm, sm.hbuf, h, it.h + 6, it.w);
  __syncthreads();
  /* dSSE/dv_k */
  double gv[3] = {0.0, 0.0, 0.0};
  for (int k = threadIdx.x; k < T * T; k += NTH) {
    const int i = k >> 6, j = k & 63;
    if (i < it.h && j < it.w) {
      const float *col = sm.hbuf + (i + 3) * T + j;
      const float ctr = col[0], e = sm.e[k];
      gv[0] += (double)(e * (col[-3 * T] + col[3 * T] - 2.f * ctr));
      gv[1] += (double)(e * (col[-2 * T] + col[2 * T] - 2.f * ctr));
      gv[2] += (double)(e * (col[-T] + col[T] - 2.f * ctr));
    }
  }
  __syncthreads();
  /* E(r, j) = sum_t V_t e(r - t, j), into hbuf */
  const float V[7] = {v[0], v[1], v[2], vc, v[2], v[1], v[0]};
  for (int k = threadIdx.x; k < (it.h + 6) * T; k += NTH) {
    const int r = k >> 6, j = k & 63;
    float s = 0.f;
#pragma unroll
    for (int t = 0; t < 7; t++) {
      const int i = r - t;
      if (i >= 0 && i < it.h) s += V[t] * sm.e[i * T + j];
    }
    sm.hbuf[k] = s;
  }
  __syncthreads();
  double gh[3] = {0.0, 0.0, 0.0};
  for (int k = threadIdx.x; k < (it.h + 6) * T; k += NTH) {
    const int r = k >> 6, c = k & 63;
    if (c < it.w) {
      const float E = sm.hbuf[k];
      const float *row = sm.w + r * WS + c + 3;
      const float ctr = row[0];
      gh[0] += (double)(E * (row[-3] + row[3] - 2.f * ctr));
      gh[1] += (double)(E * (row[-2] + row[2] - 2.f * ctr));
      gh[2] += (double)(E * (row[-1] + row[1] - 2.f * ctr));
    }
  }
  double g[6];
  for (int q = 0; q < 3; q++) {
    g[q] = block_sum(gv[q], sm.red);
    g[3 + q] = block_sum(gh[q], sm.red);
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 6; q++) a.out[(size_t)blockIdx.x * 6 + q] = g[q];
}

/* --- self-guided ------------------------------------------------------ */

struct SgrFilterArgs {
  const rvpf_lrd_item *items;
  const float *win, *u;
  const int32_t *sets; /* [unit] */
  int nunits;
  int bd;
  float *flt;          /* [n][2][64][64] */
};

struct SgrShared {
  float w[WS * WS];
  float a[(T + 2) * (T + 2)], b[(T + 2) * (T + 2)];
};

__global__ __launch_bounds__(NTH) void sgr_filter_kernel(SgrFilterArgs g) {
  __shared__ SgrShared sm;
  const rvpf_lrd_item it = load_item(g.items, blockIdx.x, g.nunits);
  const int set = min(max(g.sets[it.unit], 0), 15);
  const float *u = g.u + (size_t)blockIdx.x * T * T;
  for (int k = threadIdx.x; k < WS * WS; k += NTH) sm.w[k] = g.win[(size_t)blockIdx.x * WS * WS + k];
  __syncthreads();
  const float s2 = 1.f / (float)(1 << (2 * (g.bd - 8))), s1 = 1.f / (float)(1 << (g.bd - 8));
  for (int pass = 0; pass < 2; pass++) {
    float *flt = g.flt + ((size_t)blockIdx.x * 2 + pass) * T * T;
    const int r = AV1_SGR_PARAMS[set][pass * 2];
    if (r == 0) {
      /* the pass is off: the projection uses u itself */
      for (int k = threadIdx.x; k < T * T; k += NTH) flt[k] = 16.f * u[k];
      continue;
    }
    const int eps = AV1_SGR_PARAMS[set][pass * 2 + 1];
    const int n = (2 * r + 1) * (2 * r + 1);
    const int n2e = n * n * eps;
    const float s = (float)(((1 << 20) + n2e / 2) / n2e);
    const int aw = it.w + 2;
    for (int k = threadIdx.x; k < (it.h + 2) * aw; k += NTH) {
      const int i = k / aw, j = k - i * aw;
      if (pass == 0 && ((it.y0 + i - 1) & 1) == 0) {
        /* weighted by 0 in the first filter; keep the values finite */
        sm.a[i * (T + 2) + j] = 0.f;
        sm.b[i * (T + 2) + j] = 0.f;
        continue;
      }
      const int wy = i + 2, wx = j + 2; /* window position of rectangle (i - 1, j - 1) */
      float sum = 0.f, sq = 0.f;
      for (int dy = -r; dy <= r; dy++)
        for (int dx = -r; dx <= r; dx++) {
          const float x = sm.w[(wy + dy) * WS + wx + dx];
          sum += x;
          sq += x * x;
        }
      const float av = sq * s2, dv = sum * s1;
      const float p = fmaxf(av * n - dv * dv, 0.f);
      const float z = p * s * (1.f / 1048576.f);
      const float a2 = z >= 255.f ? 256.f : 256.f * z / (z + 1.f);
      sm.a[i * (T + 2) + j] = a2;
      sm.b[i * (T + 2) + j] = (256.f - a2) * sum / (float)n;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < T * T; k += NTH) {
      const int i = k >> 6, j = k & 63;
      float v = 0.f;
      if (i < it.h && j < it.w) {
        const int yabs = it.y0 + i;
        const int shift = (pass == 0 && (yabs & 1)) ? 4 : 5;
        float A = 0.f, B = 0.f;
        for (int dy = -1; dy <= 1; dy++)
          for (int dx = -1; dx <= 1; dx++) {
            float wt;
            if (pass == 0)
              wt = ((yabs + dy) & 1) ? (dx == 0 ? 6.f : 5.f) : 0.f;
            else
              wt = (dx == 0 || dy == 0) ? 4.f : 3.f;
            const int q = (i + 1 + dy) * (T + 2) + j + 1 + dx;
            A += wt * sm.a[q];
            B += wt * sm.b[q];
          }
        v = (A * u[k] + B) / (float)(1 << (8 + shift - 4));
      }
      flt[k] = v;
    }
    __syncthreads();
  }
}

struct SgrArgs {
  const rvpf_lrd_item *items;
  const float *flt, *u, *src;
  const float *w;  /* [unit][2]: w0, w2 */
  int nunits;
  float *resid;
  double *out;     /* forward: sse [n]; backward: grad [n][2] */
};

__global__ __launch_bounds__(NTH) void sgr_forward_kernel(SgrArgs a) {
  __shared__ double red[NTH / 64];
  const rvpf_lrd_item it = load_item(a.items, blockIdx.x, a.nunits);
  const float w0 = a.w[it.unit * 2], w2 = a.w[it.unit * 2 + 1];
  const size_t base = (size_t)blockIdx.x * T * T;
  const float *f0 = a.flt + 2 * base, *f1 = f0 + T * T;
  double acc = 0.0;
  for (int k = threadIdx.x; k < T * T; k += NTH) {
    const int i = k >> 6, j = k & 63;
    float e = 0.f;
    if (i < it.h && j < it.w) {
      /* the weights' effect is small against the sample values: form the
       * residual in double, (u - src) first (both integers) */
      const float u16 = 16.f * a.u[base + k];
      const double ed = (double)(a.u[base + k] - a.src[base + k]) +
                        ((double)w0 * (double)(f0[k] - u16) + (double)w2 * (double)(f1[k] - u16)) * (1.0 / 2048.0);
      e = (float)ed;
      acc += ed * ed;
    }
    a.resid[base + k] = e;
  }
  const double s = block_sum(acc, red);
  if (threadIdx.x == 0) a.out[blockIdx.x] = s;
}

__global__ __launch_bounds__(NTH) void sgr_backward_kernel(SgrArgs a) {
  __shared__ double red[NTH / 64];
  const size_t base = (size_t)blockIdx.x * T * T;
  const float *f0 = a.flt + 2 * base, *f1 = f0 + T * T;
  double g0 = 0.0, g2 = 0.0;
  for (int k = threadIdx.x; k < T * T; k += NTH) {
    const float e2 = 2.f * a.resid[base + k] * (1.f / 2048.f); /* zero outside the item */
    const float u16 = 16.f * a.u[base + k];
    g0 += (double)(e2 * (f0[k] - u16));
    g2 += (double)(e2 * (f1[k] - u16));
  }
  const double s0 = block_sum(g0, red);
  const double s2 = block_sum(g2, red);
  if (threadIdx.x == 0) {
    a.out[(size_t)blockIdx.x * 2] = s0;
    a.out[(size_t)blockIdx.x * 2 + 1] = s2;
  }
}

/* Exact squared error of rectangles between two sample planes. */
struct RectSseArgs {
  const rvpf_lrd_item *items;
  int nunits;
  int plane_end_x[3], plane_end_y[3];
  PlanePtr a[3], b[3];
  double *out;
};

template <typename TS>
__global__ __launch_bounds__(NTH) void rect_sse_kernel(RectSseArgs g) {
  __shared__ double red[NTH / 64];
  const rvpf_lrd_item it = load_item(g.items, blockIdx.x, g.nunits);
  const int p = it.plane;
  const TS *pa = reinterpret_cast<const TS *>(g.a[p].data);
  const TS *pb = reinterpret_cast<const TS *>(g.b[p].data);
  uint64_t acc = 0;
  for (int k = threadIdx.x; k < T * T; k += NTH) {
    const int i = k >> 6, j = k & 63;
    const int y = it.y0 + i, x = it.x0 + j;
    if (i < it.h && j < it.w && y >= 0 && x >= 0 && y <= g.plane_end_y[p] && x <= g.plane_end_x[p]) {
      const int d = (int)pa[(int64_t)(it.y0 + i) * g.a[p].stride + it.x0 + j] -
                    (int)pb[(int64_t)(it.y0 + i) * g.b[p].stride + it.x0 + j];
      acc += (uint64_t)(d * d);
    }
  }
  const double s = block_sum((double)acc, red);
  if (threadIdx.x == 0) g.out[blockIdx.x] = s;
}

}  // namespace

int launch_lrd_gather(const PlanePtr deb[3], const PlanePtr cdef[3], const PlanePtr src[3], int bps,
                      const int pex[3], const int pey[3], const rvpf_lrd_item *items, int n, int nunits,
                      float *win, float *u, float *s, hipStream_t st) {
  GatherArgs a;
  a.nunits = nunits;
  for (int p = 0; p < 3; p++) {
    a.deb[p] = deb[p];
    a.cdef[p] = cdef[p];
    a.src[p] = src[p];
    a.plane_end_x[p] = pex[p];
    a.plane_end_y[p] = pey[p];
  }
  a.bps = bps;
  a.items = items;
  a.win = win;
  a.u = u;
  a.s = s;
  if (bps == 1)
    gather_kernel<uint8_t><<<n, NTH, 0, st>>>(a);
  else
    gather_kernel<uint16_t><<<n, NTH, 0, st>>>(a);
  return (int)hipGetLastError();
}

int launch_lrd_wiener(int backward, const rvpf_lrd_item *items, int n, int nunits, const float *win,
                      const float *src, const float *taps, float *resid, double *out, hipStream_t st) {
  WienerArgs a = {items, win, src, taps, nunits, resid, out};
  if (backward)
    wiener_backward_kernel<<<n, NTH, 0, st>>>(a);
  else
    wiener_forward_kernel<<<n, NTH, 0, st>>>(a);
  return (int)hipGetLastError();
}

int launch_lrd_sgr_filter(const rvpf_lrd_item *items, int n, int nunits, const float *win, const float *u,
                          const int32_t *sets, int bd, float *flt, hipStream_t st) {
  SgrFilterArgs a = {items, win, u, sets, nunits, bd, flt};
  sgr_filter_kernel<<<n, NTH, 0, st>>>(a);
  return (int)hipGetLastError();
}

int launch_lrd_sgr(int backward, const rvpf_lrd_item *items, int n, int nunits, const float *flt,
                   const float *u, const float *src, const float *w, float *resid, double *out, hipStream_t st) {
  SgrArgs a = {items, flt, u, src, w, nunits, resid, out};
  if (backward)
    sgr_backward_kernel<<<n, NTH, 0, st>>>(a);
  else
    sgr_forward_kernel<<<n, NTH, 0, st>>>(a);
  return (int)hipGetLastError();
}

int launch_lrd_rect_sse(const rvpf_lrd_item *items, int n, int nunits, const int pex[3], const int pey[3],
                        const PlanePtr a[3], const PlanePtr b[3], int bps, double *out, hipStream_t st) {
  RectSseArgs g;
  g.items = items;
  g.nunits = nunits;
  for (int p = 0; p < 3; p++) {
    g.plane_end_x[p] = pex[p];
    g.plane_end_y[p] = pey[p];
  }
  for (int p = 0; p < 3; p++) {
    g.a[p] = a[p];
    g.b[p] = b[p];
  }
  g.out = out;
  if (bps == 1)
    rect_sse_kernel<uint8_t><<<n, NTH, 0, st>>>(g);
  else
    rect_sse_kernel<uint16_t><<<n, NTH, 0, st>>>(g);
  return (int)hipGetLastError();
}

}  // namespace rvpf

using namespace rvpf;
using namespace rvpf::host;

namespace {

/* Plane extents for items; absent planes get plane 0's pointers and an
 * empty extent, so an item naming one reads nothing of it. */
void lrd_planes(const rvpf_frame_params *p, const rvpf_planes *pl, PlanePtr out[3], int pex[3], int pey[3]) {
  to_ptrs(pl, out);
  for (int i = 0; i < 3; i++) {
    if (i < nplanes_of(p)) {
      int w, h;
      crop(p, i, &w, &h);
      pex[i] = w - 1;
      pey[i] = h - 1;
    } else {
      out[i] = out[0];
      pex[i] = pey[i] = -1;
    }
  }
}

int lrd_check(rvpf_ctx *ctx, const void *items, int n, int nunits) {
  if (!ctx) return fail(nullptr, RVPF_E_INVALID, "ctx is NULL");
  if (n < 0 || nunits < 0 || (n > 0 && !items)) return fail(ctx, RVPF_E_INVALID, "items or counts invalid");
  return RVPF_OK;
}

}  // namespace

extern "C" {

int rvpf_lrd_gather(rvpf_ctx *ctx, const rvpf_frame_params *p, const rvpf_planes *deblocked,
                    const rvpf_planes *cdef, const rvpf_planes *source, const rvpf_lrd_item *items,
                    int32_t n, int32_t nunits, float *win, float *u, float *src, void *stream) {
  int rc;
  if ((rc = lrd_check(ctx, items, n, nunits))) return rc;
  if ((rc = validate_params(p, ctx->err)) || (rc = validate_planes(p, deblocked, "deblocked", 1, 0, ctx->err)) ||
      (rc = validate_planes(p, cdef, "cdef", 1, 0, ctx->err)) ||
      (rc = validate_planes(p, source, "source", 1, 0, ctx->err)))
    return rc;
  if (deblocked->bytes_per_sample != cdef->bytes_per_sample || source->bytes_per_sample != cdef->bytes_per_sample)
    return fail(ctx, RVPF_E_INVALID, "plane sample sizes differ");
  if (n == 0) return RVPF_OK;
  if (!win || !u || !src) return fail(ctx, RVPF_E_INVALID, "output buffers are NULL");
  PlanePtr d[3], c[3], s[3];
  int pex[3], pey[3];
  lrd_planes(p, deblocked, d, pex, pey);
  lrd_planes(p, cdef, c, pex, pey);
  lrd_planes(p, source, s, pex, pey);
  return check_launch(ctx, launch_lrd_gather(d, c, s, cdef->bytes_per_sample, pex, pey, items, n, nunits, win, u,
                                             src, (hipStream_t)stream),
                      "lrd gather");
}

int rvpf_lrd_wiener(rvpf_ctx *ctx, int backward, const rvpf_lrd_item *items, int32_t n, int32_t nunits,
                    const float *win, const float *src, const float *taps, float *resid, double *out,
                    void *stream) {
  int rc;
  if ((rc = lrd_check(ctx, items, n, nunits))) return rc;
  if (n == 0) return RVPF_OK;
  if (!win || !taps || !resid || !out || (!backward && !src)) return fail(ctx, RVPF_E_INVALID, "buffers are NULL");
  return check_launch(ctx, launch_lrd_wiener(backward, items, n, nunits, win, src, taps, resid, out,
                                             (hipStream_t)stream),
                      "lrd wiener");
}

int rvpf_lrd_sgr_filter(rvpf_ctx *ctx, int32_t bit_depth, const rvpf_lrd_item *items, int32_t n,
                        int32_t nunits, const float *win, const float *u, const int32_t *sets, float *flt,
                        void *stream) {
  int rc;
  if ((rc = lrd_check(ctx, items, n, nunits))) return rc;
  if (bit_depth != 8 && bit_depth != 10 && bit_depth != 12) return fail(ctx, RVPF_E_INVALID, "bit_depth");
  if (n == 0) return RVPF_OK;
  if (!win || !u || !sets || !flt) return fail(ctx, RVPF_E_INVALID, "buffers are NULL");
  return check_launch(ctx, launch_lrd_sgr_filter(items, n, nunits, win, u, sets, bit_depth, flt,
                                                 (hipStream_t)stream),
                      "lrd sgr filter");
}

int rvpf_lrd_sgr(rvpf_ctx *ctx, int backward, const rvpf_lrd_item *items, int32_t n, int32_t nunits,
                 const float *flt, const float *u, const float *src, const float *w, float *resid, double *out,
                 void *stream) {
  int rc;
  if ((rc = lrd_check(ctx, items, n, nunits))) return rc;
  if (n == 0) return RVPF_OK;
  if (!flt || !u || !resid || !out || (!backward && (!src || !w))) return fail(ctx, RVPF_E_INVALID, "buffers are NULL");
  return check_launch(ctx, launch_lrd_sgr(backward, items, n, nunits, flt, u, src, w, resid, out,
                                          (hipStream_t)stream),
                      "lrd sgr");
}

int rvpf_lrd_rect_sse(rvpf_ctx *ctx, const rvpf_frame_params *p, const rvpf_lrd_item *items, int32_t n,
                      int32_t nunits, const rvpf_planes *a, const rvpf_planes *b, double *out, void *stream) {
  int rc;
  if ((rc = lrd_check(ctx, items, n, nunits))) return rc;
  if ((rc = validate_params(p, ctx->err)) || (rc = validate_planes(p, a, "a", 1, 0, ctx->err)) ||
      (rc = validate_planes(p, b, "b", 1, 0, ctx->err)))
    return rc;
  if (a->bytes_per_sample != b->bytes_per_sample) return fail(ctx, RVPF_E_INVALID, "plane sample sizes differ");
  if (n == 0) return RVPF_OK;
  if (!out) return fail(ctx, RVPF_E_INVALID, "out is NULL");
  PlanePtr pa[3], pb[3];
  int pex[3], pey[3];
  lrd_planes(p, a, pa, pex, pey);
  lrd_planes(p, b, pb, pex, pey);
  return check_launch(ctx, launch_lrd_rect_sse(items, n, nunits, pex, pey, pa, pb, a->bytes_per_sample, out,
                                               (hipStream_t)stream),
                      "lrd rect sse");
}

int32_t rvpf_ac_q(int32_t bit_depth, int32_t qindex) {
  if ((bit_depth != 8 && bit_depth != 10 && bit_depth != 12) || qindex < 0 || qindex > 255) return -1;
  return AV1_AC_QLOOKUP[(bit_depth - 8) >> 1][qindex];
}

}  // extern "C"
