/*
 * rvpf_host.h — host-side state and helpers shared by the C ABI files
 * (rvpf_host.hip: per-frame calls; batch.hip: the batched path).
 */
#pragma once

#include <stdint.h>

#include <vector>

#include "rvpf_internal.h"

struct ProfMark {
  int stage;
  hipEvent_t a, b;
};

/* Tile choices of one batch shape, remembered by the context. */
struct TuneEntry {
  uint64_t key;
  rvpf_tile_config cfg[RVPF_TUNE_COUNT];
};

/* A scratch buffer shared by calls that may run on different streams.
 * While every call comes from one stream, stream order alone keeps them
 * apart and nothing is recorded (an event on the stream costs the queue a
 * few microseconds per call). The first call from another stream waits for
 * the device once; from then on `free` is recorded on the stream of each
 * call after its last kernel that reads the buffer, and a call from another
 * stream waits for it before its first kernel writes the buffer. */
struct ScratchGate {
  hipEvent_t free;
  int used;           /* a call has used the buffer */
  int multi;          /* calls have come from more than one stream */
  int recorded;       /* `free` holds a record */
  hipStream_t stream; /* the stream of the last call */
};

/* Film grain scratch slots: the templates of a frame are built on the side
 * stream while the previous frames still filter, into slot n mod
 * GRAIN_SLOTS for the n-th frame; its apply waits for them (ev_ready[slot]).
 * A slot is rebuilt only after the apply that last read it: the apply of
 * every GRAIN_DONE_EVERY-th frame records ev_done[] (alternately), and the
 * build of frame n waits for the record that follows frame n - GRAIN_SLOTS,
 * so the filter stream records an event once per GRAIN_DONE_EVERY frames. */
constexpr int GRAIN_SLOTS = 8, GRAIN_DONE_EVERY = 4;

/* The template slot of one frame's film grain. */
struct GrainTicket {
  int slot;
  int64_t frame;
  hipStream_t built_on;
};
struct rvpf_ctx {
  int device;
  int cus;                          /* compute units of the device */
  int profiling;
  std::vector<ProfMark> marks;      /* recorded, not yet read */
  std::vector<hipEvent_t> free_evs; /* timing events for reuse */
  hipStream_t side;
  hipEvent_t ev_ready[GRAIN_SLOTS], ev_done[2];
  int64_t grain_frames; /* templates built so far */
  int64_t done_frame[2]; /* the frame whose apply the latest record of ev_done[i] follows; -1: none */
  void *deb_buf; /* the chain's deblocked frame */
  size_t deb_bytes;
  void *cdef_buf;
  size_t cdef_bytes;
  void *up_buf; /* upscaled CDEF output and deblocked frame (super-resolution) */
  size_t up_bytes;
  ScratchGate chain_gate; /* deb_buf, cdef_buf and up_buf */
  void *search_buf; /* encoder searches: work planes, error tables */
  size_t search_bytes;
  ScratchGate search_gate;
  rvpf::GrainScratch grain[GRAIN_SLOTS];
  std::vector<TuneEntry> tune_cache;
  char err[256];
};

namespace rvpf {
namespace host {

int fail(rvpf_ctx *ctx, int code, const char *msg);
int hip_fail(rvpf_ctx *ctx, hipError_t e, const char *where);
int nplanes_of(const rvpf_frame_params *p);
bool lf_on(const rvpf_frame_params *p);
bool cdef_on(const rvpf_frame_params *p);
bool lr_plane_on(const rvpf_frame_params *p, int pl);
bool lr_on(const rvpf_frame_params *p);
bool grain_on(const rvpf_frame_params *p);
/* the mode-info area / the display crop of a plane */
void area(const rvpf_frame_params *p, int pl, int *w, int *h);
void crop(const rvpf_frame_params *p, int pl, int *w, int *h);
int validate_params(const rvpf_frame_params *p, char *msg);
int validate_side(const rvpf_frame_params *p, const rvpf_side_info *s, char *msg);
/* margin: extra rows/cols beyond the area the stage reads; pairs: the
 * stage stores two adjacent samples at once at even columns */
int validate_planes(const rvpf_frame_params *p, const rvpf_planes *pl, const char *name, int use_crop, int margin,
                    char *msg, bool pairs = false);
Geom make_geom(const rvpf_frame_params *p);
void to_ptrs(const rvpf_planes *pl, PlanePtr out[3]);
LfHdr make_lf(const rvpf_frame_params *p);
CdefHdr make_cdef(const rvpf_frame_params *p);
LrHdr make_lr(const rvpf_frame_params *p, const rvpf_side_info *s);
GrainHdr make_grain(const rvpf_frame_params *p);
int check_launch(rvpf_ctx *ctx, int rc, const char *what);
/* grows a scratch buffer of the context to at least `bytes` (waits for
 * the device before freeing the old one) */
int ensure_buf(rvpf_ctx *ctx, void **buf, size_t *cap, size_t bytes, const char *what);
/* orders a call's use of a shared scratch buffer on stream `st` after the
 * previous call's (acquire before the first write, release after the last
 * read) */
int scratch_acquire(rvpf_ctx *ctx, ScratchGate *g, hipStream_t st);
int scratch_release(rvpf_ctx *ctx, ScratchGate *g, hipStream_t st);
/* planes of w[i] x h[i] samples one after another from base (rows padded
 * to 64 samples, planes to 256 bytes); returns the bytes; base NULL only
 * counts */
size_t layout_planes(const int w[3], const int h[3], int bps, void *base, PlanePtr out[3]);
/* the display planes hold the reference planes' sample size, or bytes
 * from 16-bit planes at 8 bits */
bool display_bps_ok(const rvpf_frame_params *p, int ref_bps, int disp_bps);
/* bytes of the offsets table of a grain scratch for a w x h frame */
void grain_offsets_dims(int w, int h, int *stripes, int *blocks);

}  // namespace host
}  // namespace rvpf
