/*
 * noalloc.hip — rvpf_grain_noalloc: the output stage (crop + film grain)
 * with no device allocation at all, for callers that cannot allocate
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).
 *
 * Its kernels are those of rvpf_grain (grain.hip): the template build
 * (grain_prepare_kernel, with the scaling functions and block offsets) and
 * the noise blend. Their
 * scratch — 3 planes x 8 shifted copies of the 73 x 88 noise templates,
 * the 3 scaling functions (an entry per sample value) and one byte per
 * 32 x 32 noise block — lives in
 * the file-scope array `scratch` below, sized for frames up to
 * NOALLOC_MAX_W x NOALLOC_MAX_H. Every GPU the library runs on holds its
 * own copy until the process ends. The device address of that copy is
 * taken once per device (hipGetSymbolAddress, on the first call with that
 * device current) and passed to the kernels; the array's name is never
 * used as an address on the host.
 *
 * Every call on one GPU shares that GPU's scratch and leaves its contents
 * behind; nothing in it needs to start at zero (each call writes every
 * template sample, LUT entry and block offset its blend reads). Calls that
 * overlap on one GPU (on different streams, or from several threads) race
 * on it.
 */
#include <stdio.h>
#include <string.h>

#include <mutex>

#include "rvpf_host.h"

using namespace rvpf;
using namespace rvpf::host;

namespace {

constexpr int NOALLOC_MAX_W = 8192, NOALLOC_MAX_H = 8192;
/* block offsets table: noise stripes x blocks per stripe (grain_offsets_dims) */
constexpr int NOALLOC_STRIPES = ((NOALLOC_MAX_H + 1) / 2 + 15) / 16 + 1;
constexpr int NOALLOC_BLOCKS = ((NOALLOC_MAX_W + 1) / 2 + 15) / 16 + 1;
constexpr size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }
constexpr size_t OFF_LUT = align256(GRAIN_TMPL_BYTES);
constexpr size_t OFF_OFFSETS = OFF_LUT + align256(GRAIN_LUT_BYTES);
constexpr size_t SCRATCH_BYTES = OFF_OFFSETS + align256((size_t)NOALLOC_STRIPES * NOALLOC_BLOCKS);

__device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
std::mutex g_mu;
void *g_ptr[MAX_DEVICES]; /* device address of `scratch` on each device, once taken */

/* The device address of `scratch` on the current device (taken once per device). */
int scratch_ptr(void **out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(nullptr, e, "rvpf_grain_noalloc: hipGetDevice");
  if (dev < 0 || dev >= MAX_DEVICES) return fail(nullptr, RVPF_E_UNSUPPORTED, "rvpf_grain_noalloc: device index");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_ptr[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return hip_fail(nullptr, e != hipSuccess ? e : hipErrorInvalidSymbol, "rvpf_grain_noalloc: hipGetSymbolAddress");
    }
    g_ptr[dev] = p;
  }
  *out = g_ptr[dev];
  return RVPF_OK;
}

}  // namespace

extern "C" {

int rvpf_grain_noalloc(const rvpf_frame_params *p, const rvpf_planes *ref, rvpf_planes *display, void *stream) {
  char msg[256];
  int rc;
  if ((rc = validate_params(p, msg)) || (rc = validate_planes(p, ref, "ref", 1, 0, msg)) ||
      (rc = validate_planes(p, display, "display", 1, 0, msg)))
    return fail(nullptr, rc, msg);
  if (!display_bps_ok(p, ref->bytes_per_sample, display->bytes_per_sample))
    return fail(nullptr, RVPF_E_INVALID, "ref and display sample sizes differ (bytes from 16-bit samples only at 8 bits)");
  if (p->upscaled_width > NOALLOC_MAX_W || p->frame_height > NOALLOC_MAX_H) {
    snprintf(msg, sizeof msg, "rvpf_grain_noalloc: frame %dx%d exceeds the scratch's maximum of %dx%d",
             p->upscaled_width, p->frame_height, NOALLOC_MAX_W, NOALLOC_MAX_H);
    return fail(nullptr, RVPF_E_INVALID, msg);
  }
  hipStream_t st = (hipStream_t)stream;
  const Geom g = make_geom(p);
  PlanePtr in[3], out[3];
  to_ptrs(ref, in);
  to_ptrs(display, out);
  if (!grain_on(p)) {
    int w[3] = {0, 0, 0}, h[3] = {0, 0, 0};
    for (int i = 0; i < g.nplanes; i++) crop(p, i, &w[i], &h[i]);
    return check_launch(nullptr, launch_copy_planes(g, in, out, w, h, ref->bytes_per_sample, st,
                                                     display->bytes_per_sample),
                        "display copy");
  }
  void *base = nullptr;
  if ((rc = scratch_ptr(&base))) return rc;
  GrainScratch sc;
  sc.tmpl = reinterpret_cast<int16_t *>(static_cast<char *>(base));
  sc.lut = reinterpret_cast<int16_t *>(static_cast<char *>(base) + OFF_LUT);
  sc.offsets = reinterpret_cast<uint8_t *>(static_cast<char *>(base) + OFF_OFFSETS);
  sc.max_stripes = NOALLOC_STRIPES;
  sc.max_blocks = NOALLOC_BLOCKS;
  const GrainHdr gh = make_grain(p);
  if ((rc = check_launch(nullptr, launch_grain_prepare(gh, sc, st), "grain prepare"))) return rc;
  return check_launch(nullptr,
                      launch_grain_apply(gh, sc, in, out, ref->bytes_per_sample, st, display->bytes_per_sample),
                      "grain apply");
}

}  // extern "C"
