/*
 * rvpf.h — C ABI of the MI355X (gfx950) AV1 decoder post-filter chain.
 *
 * The chain is the AV1 in-loop and output filters run once a frame is
 * reconstructed: the deblocking loop filter (7.14), CDEF (7.15), loop
 * restoration (7.17, Wiener and self-guided) and film grain synthesis
 * (7.18.3). Every entry point here replaces a call inside rivet-av1's
 * decoder; each declaration names the reference function it stands in for
 * (paths relative to the rivet-av1 crate root).
 *
 * Conventions
 *   - Plain C: fixed-width integers, plain pointers and sizes, no HIP or
 *     torch types. A HIP stream is passed as `void *` (NULL = default
 *     stream). Sample planes and side-info arrays handed to the rvpf_*
 *     compute calls are DEVICE pointers (hipMalloc'd or torch tensors).
 *   - Every call returns an int status (RVPF_OK or a negative RVPF_E_*),
 *     never aborts. Work is enqueued on the stream; the call returns once
 *     it is enqueued (rvpf_ctx_sync waits).
 *   - Samples are uint16_t at every bit depth (the reference's PlaneBuf,
 *     src/decoder/state.rs:11-42), or uint8_t at 8 bits when
 *     rvpf_planes.bytes_per_sample == 1. Plane p sample (x, y) is at
 *     data[p][y * stride[p] + x]; there is no negative offset.
 *
 * Bit-exactness contract: for legal side info (transform blocks aligned to
 * their size inside blocks aligned to theirs, as every AV1 bitstream has)
 * the outputs equal rivet-av1's CPU filters sample for sample over
 *   - deblocking: the whole plane buffer;
 *   - CDEF / LR / chain reference planes: the mode-info-aligned area
 *     ((mi_cols*4) >> ssx) x ((mi_rows*4) >> ssy) of each plane; with
 *     super-resolution, the LR and chain reference planes over the
 *     upscaled_width x frame_height crop;
 *   - super-resolution upscaling: the upscaled_width x frame_height crop;
 *   - display planes: the cropped upscaled_width x frame_height picture.
 */
#ifndef RVPF_H
#define RVPF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes (the reference's Error kinds, src/lib.rs:60-73). */
#define RVPF_OK 0
#define RVPF_E_INVALID (-1)     /* geometry, strides or parameters out of range */
#define RVPF_E_UNSUPPORTED (-2) /* e.g. bytes_per_sample 1 above 8 bits, super-resolution frames in a batch */
#define RVPF_E_DEVICE (-3)      /* a HIP runtime error */
#define RVPF_E_NOMEM (-4)       /* device allocation failed */

/* Restoration types (src/consts.rs RESTORE_*). */
#define RVPF_RESTORE_NONE 0
#define RVPF_RESTORE_WIENER 1
#define RVPF_RESTORE_SGRPROJ 2

/*
 * Mode info of one 4x4 luma block: the fields of the reference's `Mi`
 * (src/decoder/state.rs:56-75) that the loop filter (postfilter.rs:107-264)
 * and CDEF's skip test (postfilter.rs:306-307) read. 8 bytes.
 */
typedef struct rvpf_mi {
  uint8_t mi_size;    /* BLOCK_SIZE, 0..21 (BLOCK_4X4 .. BLOCK_16X64) */
  uint8_t y_mode;     /* YMode; inter modes are >= NEARESTMV (14) */
  int8_t ref_frame0;  /* RefFrame[0]: -1 none, 0 intra, 1..7 inter */
  uint8_t seg_skip;   /* bit 0: skip (no residual); bits 1..3: segment_id */
  int8_t delta_lf[4]; /* DeltaLF[0..3] */
} rvpf_mi;

/*
 * One loop restoration unit (LrType / LrWiener / LrSgrSet / LrSgrXqd of
 * the reference's LrPlane, src/decoder/mod.rs:31-41). 12 bytes.
 */
typedef struct rvpf_lr_unit {
  uint8_t type;         /* RVPF_RESTORE_* */
  uint8_t sgr_set;      /* 0..15 */
  int8_t sgr_xqd[2];    /* [-96,31], [-32,95] */
  int8_t wiener[2][3];  /* [0] vertical, [1] horizontal taps c0..c2 */
  uint8_t pad[2];
} rvpf_lr_unit;

/* Film grain parameters (the reference's FilmGrainParams, src/header.rs:14-43). */
typedef struct rvpf_grain_params {
  int32_t apply_grain;
  uint32_t grain_seed;
  int32_t num_y_points; /* 0..14 */
  int32_t point_y_value[16];
  int32_t point_y_scaling[16];
  int32_t chroma_scaling_from_luma;
  int32_t num_cb_points; /* 0..10 */
  int32_t point_cb_value[16];
  int32_t point_cb_scaling[16];
  int32_t num_cr_points; /* 0..10 */
  int32_t point_cr_value[16];
  int32_t point_cr_scaling[16];
  int32_t grain_scaling_minus_8; /* 0..3 */
  int32_t ar_coeff_lag;          /* 0..3 */
  int32_t ar_coeffs_y_plus_128[24];
  int32_t ar_coeffs_cb_plus_128[25];
  int32_t ar_coeffs_cr_plus_128[25];
  int32_t ar_coeff_shift_minus_6; /* 0..3 */
  int32_t grain_scale_shift;      /* 0..3 */
  int32_t cb_mult, cb_luma_mult, cb_offset;
  int32_t cr_mult, cr_luma_mult, cr_offset;
  int32_t overlap_flag;
  int32_t clip_to_restricted_range;
} rvpf_grain_params;

/*
 * Everything scalar the chain reads from the sequence and frame headers
 * (the reference's FrameCtx / FrameHeader / SequenceHeader fields,
 * src/decoder/mod.rs:44-80, src/header.rs:89-177).
 */
typedef struct rvpf_frame_params {
  /* sequence */
  int32_t bit_depth;   /* 8, 10, 12 */
  int32_t ssx, ssy;    /* chroma subsampling: 4:2:0 = 1,1; 4:2:2 = 1,0; 4:4:4 = 0,0 */
  int32_t mono_chrome; /* 1 plane */
  int32_t matrix_coefficients;
  int32_t enable_cdef;
  int32_t film_grain_params_present;
  /* frame geometry */
  int32_t frame_width, frame_height;
  int32_t upscaled_width; /* frame_width, or with use_superres the width frame_width is upscaled to:
                           * frame_width = (upscaled_width * 8 + d / 2) / d for a denominator d in 9..16 */
  int32_t use_superres;
  int32_t mi_rows, mi_cols;
  int32_t coded_lossless, allow_intrabc;
  /* loop filter (7.14) */
  int32_t loop_filter_level[4];
  int32_t loop_filter_sharpness;
  int32_t loop_filter_delta_enabled;
  int32_t loop_filter_ref_deltas[8];
  int32_t loop_filter_mode_deltas[2];
  int32_t delta_lf_multi;
  int32_t segmentation_enabled;
  int32_t seg_feature_enabled[8][8]; /* [segment][SEG_LVL_*] */
  int32_t seg_feature_data[8][8];
  /* CDEF (7.15) */
  int32_t cdef_damping; /* 3..6 */
  int32_t cdef_y_pri_strength[8], cdef_y_sec_strength[8];   /* sec in {0,1,2,4} */
  int32_t cdef_uv_pri_strength[8], cdef_uv_sec_strength[8];
  /* loop restoration (7.17) */
  int32_t frame_restoration_type[3]; /* RVPF_RESTORE_* per plane */
  int32_t loop_restoration_size[3];  /* 32..256 */
  int32_t lr_unit_rows[3], lr_unit_cols[3]; /* count_units_in_frame(), src/decoder/mod.rs:1229 */
  /* film grain (7.18.3) */
  rvpf_grain_params grain;
} rvpf_frame_params;

/*
 * Per-block side info (device pointers). The 4x4 arrays have row stride
 * mi_stride >= mi_cols + 1 and at least mi_rows + 1 rows, like the
 * reference's `ms = mi_cols + 32` arrays (src/decoder/mod.rs:547-548):
 * blocks overhanging the frame write the extra row and column, and the
 * chroma filters read them.
 */
typedef struct rvpf_side_info {
  const rvpf_mi *mi;
  int32_t mi_stride;
  /* TX_SIZE per 4x4 of each plane (luma-MI stride for all planes; chroma
   * entry (row >> ssy) * mi_stride + (col >> ssx)), as lf_tx_sizes. */
  const uint8_t *tx_size[3];
  /* cdef_idx per 64x64 (-1: not filtered), stride cdef_stride >= (mi_cols+15)/16. */
  const int8_t *cdef_idx;
  int32_t cdef_stride;
  /* lr_unit_rows[p] * lr_unit_cols[p] units per plane with restoration. */
  const rvpf_lr_unit *lr_units[3];
} rvpf_side_info;

/* Sample planes (device pointers). */
typedef struct rvpf_planes {
  void *data[3];
  int64_t stride[3]; /* in samples */
  int32_t width[3];  /* allocated columns (<= stride) */
  int32_t height[3]; /* allocated rows */
  int32_t bytes_per_sample; /* 2, or 1 at 8 bits */
} rvpf_planes;

typedef struct rvpf_ctx rvpf_ctx;

/* A context owns device scratch, uploaded tables and graph instances, on
 * the device current at creation. Successive calls on one context may be
 * enqueued on different streams: the context orders its own scratch
 * buffers between them with events (a call's first kernel that writes a
 * scratch buffer waits for the previous call's last kernel that read it),
 * so each call's result is as if the calls ran one after another; the
 * caller orders its own buffers. Calls on one context are not reentrant
 * across host threads; use one context per thread. */
int rvpf_ctx_create(rvpf_ctx **out);
void rvpf_ctx_destroy(rvpf_ctx *ctx);
int rvpf_ctx_sync(rvpf_ctx *ctx, void *stream);
/* Message of the last error on this context (static storage, never NULL). */
const char *rvpf_ctx_last_error(const rvpf_ctx *ctx);

/*
 * Per-stage profiling (no counterpart in the reference, which has no
 * tracing; for callers that want stage timings without a profiler). While
 * enabled, every stage's kernels are bracketed by HIP timing events on the
 * stream they run on; rvpf_ctx_read_profile waits for them, adds the
 * elapsed times per stage and resets the counters.
 */
#define RVPF_STAGE_DEBLOCK 0
#define RVPF_STAGE_CDEF 1
#define RVPF_STAGE_LR 2
#define RVPF_STAGE_GRAIN_PREPARE 3
#define RVPF_STAGE_GRAIN_APPLY 4
#define RVPF_STAGE_COPY 5
#define RVPF_STAGE_UPSCALE 6
#define RVPF_STAGE_COUNT 7
typedef struct rvpf_profile {
  float ms[RVPF_STAGE_COUNT];       /* summed kernel time per stage */
  int32_t launches[RVPF_STAGE_COUNT]; /* stage executions measured */
} rvpf_profile;
int rvpf_ctx_set_profiling(rvpf_ctx *ctx, int enable);
int rvpf_ctx_read_profile(rvpf_ctx *ctx, rvpf_profile *out);

/* Validates params + geometry exactly as the compute calls do (host only). */
int rvpf_validate(const rvpf_frame_params *p, const rvpf_side_info *s);

/*
 * Deblocking loop filter, in place on `frame` (replaces
 * loop_filter_threads, src/decoder/postfilter.rs:67-104, with
 * edge_loop_filter :107-203, filter_strength :207-264 and
 * dsp::lf::filter_lanes, src/dsp/lf.rs:58-211). Vertical edges of a plane
 * are all filtered before its horizontal ones.
 */
int rvpf_deblock(rvpf_ctx *ctx, const rvpf_frame_params *p, const rvpf_side_info *s,
                 rvpf_planes *frame, void *stream);

/*
 * CDEF from the deblocked frame into `out` (replaces cdef,
 * src/decoder/postfilter.rs:273-301, with cdef_block :304-342,
 * cdef_direction :345-401, cdef_filter_block :449-513 and
 * dsp::cdef::filter_scalar, src/dsp/cdef.rs:48-90). Samples of blocks that
 * are not filtered are copied.
 */
int rvpf_cdef(rvpf_ctx *ctx, const rvpf_frame_params *p, const rvpf_side_info *s,
              const rvpf_planes *deblocked, rvpf_planes *out, void *stream);

/*
 * Super-resolution upscaling (7.16) of `in` (frame_width wide, the
 * mode-info area readable) into `out` (upscaled_width wide): rows
 * 0 .. frame_height - 1 and columns 0 .. upscaled_width - 1 of each plane
 * (subsampled for chroma) are written, nothing else (replaces upscale,
 * src/decoder/postfilter.rs:521-566, called on the CDEF output and on the
 * deblocked frame by decode_frame_wrapup, src/decoder/mod.rs:752-757).
 * Requires use_superres.
 */
int rvpf_upscale(rvpf_ctx *ctx, const rvpf_frame_params *p, const rvpf_planes *in, rvpf_planes *out,
                 void *stream);

/*
 * Loop restoration of the CDEF output into `out`, reading the deblocked
 * frame above and below each 64-row stripe (replaces loop_restoration,
 * src/decoder/postfilter.rs:608-636, with lr_stripe :640-661,
 * restore_stripe :677-716, restore_rect_from_window :781-853 and
 * box_filter_rect :901-991). Samples outside restored units are copied.
 * With use_superres the three frames are the upscaled ones (as the
 * reference's, from rvpf_upscale): their width is upscaled_width rounded
 * up to 8 (subsampled for chroma), and only the upscaled_width x
 * frame_height crop of `out` is defined.
 */
int rvpf_lr(rvpf_ctx *ctx, const rvpf_frame_params *p, const rvpf_side_info *s,
            const rvpf_planes *deblocked, const rvpf_planes *cdef, rvpf_planes *out,
            void *stream);

/*
 * Output stage: the upscaled_width x frame_height crop of `ref` into
 * `display`, with film grain added when film_grain_params_present &&
 * grain.apply_grain (replaces the grain branch of Decoder::output,
 * src/decoder/mod.rs:846-917, and grain::apply, src/decoder/grain.rs:24-354).
 */
int rvpf_grain(rvpf_ctx *ctx, const rvpf_frame_params *p, const rvpf_planes *ref,
               rvpf_planes *display, void *stream);

/*
 * rvpf_grain with no device allocation. No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). It takes
 * no context; the film grain scratch (noise templates, scaling LUTs, block
 * offsets, 399 360 bytes) is a static device array of the library, one per
 * GPU, held until the process ends. Frames up to 8192 x 8192 (upscaled_width
 * x frame_height); larger ones return RVPF_E_INVALID before anything is
 * enqueued. Calls that overlap on one GPU (on different streams or from
 * several threads) race on that scratch; calls on one stream do not. Errors
 * are read with rvpf_ctx_last_error(NULL).
 */
int rvpf_grain_noalloc(const rvpf_frame_params *p, const rvpf_planes *ref, rvpf_planes *display, void *stream);

/*
 * The packed picture Decoder::output returns (src/decoder/mod.rs:846-917,
 * the Frame layout of src/frame.rs:131-190): Y, then U, then V (Y alone
 * for monochrome), each tightly packed (stride = width), one byte per
 * sample at 8 bits, else little-endian uint16 with the value in the low
 * bits. Offsets are in bytes from the start of the picture.
 */
typedef struct rvpf_packed_layout {
  int64_t offset[3];
  int32_t width[3], height[3];
  int32_t num_planes;
  int32_t bytes_per_sample; /* 1 at 8 bits, else 2 */
  int64_t bytes;            /* the whole picture */
} rvpf_packed_layout;

/* Host only: the packed picture's layout for `p` (Frame::new, src/frame.rs:161-190). */
int rvpf_packed_layout_of(const rvpf_frame_params *p, rvpf_packed_layout *out);

/*
 * Output stage into the packed picture: the crop of `ref` (16-bit or, at
 * 8 bits, 8-bit sample planes), with film grain as rvpf_grain, written to
 * the device buffer `packed` of at least `bytes` bytes in the layout of
 * rvpf_packed_layout_of (replaces Decoder::output, src/decoder/mod.rs:846-917,
 * crop, grain and pack in one pass). The grain templates are built on the
 * context's side stream, so they overlap work queued before this call.
 */
int rvpf_output_packed(rvpf_ctx *ctx, const rvpf_frame_params *p, const rvpf_planes *ref, void *packed,
                       int64_t bytes, void *stream);

/*
 * The whole chain of Decoder::decode_frame_wrapup (src/decoder/mod.rs:747-816)
 * and Decoder::output: deblock (`recon` is read, not modified: the deblocked
 * frame is kept in the context's scratch), CDEF, super-resolution
 * upscaling when use_superres, loop restoration into `ref_out` (the
 * reference frame; upscaled with super-resolution, when only its
 * upscaled_width x frame_height crop is defined), then crop + grain into
 * `display` (may be NULL when the frame is not shown). At 8 bits the
 * display planes may hold bytes while the other planes hold 16-bit samples
 * (also for rvpf_grain).
 */
int rvpf_postfilter_frame(rvpf_ctx *ctx, const rvpf_frame_params *p, const rvpf_side_info *s,
                          rvpf_planes *recon, rvpf_planes *ref_out, rvpf_planes *display,
                          void *stream);

/*
 * Encoder filter searches: the reference's search_lf
 * (src/encoder/mod.rs:1460-1490) and cdef::search (src/encoder/cdef.rs:52-253),
 * which choose a frame's loop filter levels and CDEF strengths on the first
 * pass's reconstruction against the source. The filtering and error sums
 * run on the GPU with the decoder's kernels; the calls are synchronous
 * (they wait for `stream`) and return host results.
 */

/*
 * Loop filter levels: `recon` (not modified) deblocked with the start
 * levels p->loop_filter_level scaled by 0, 0.5, 0.75, 1, 1.25 and 1.5
 * (rounded half away from zero, at most 63), each plane's squared error
 * against `source` over the frame, and for each plane the first scale with
 * the least error; chroma is off when luma is. levels = {luma, luma, cb,
 * cr}. The other loop filter fields of p (sharpness, deltas, segment
 * features) are used as they are.
 */
int rvpf_lf_search(rvpf_ctx *ctx, const rvpf_frame_params *p, const rvpf_side_info *s,
                   const rvpf_planes *recon, const rvpf_planes *source, int32_t levels[4], void *stream);

/* The CDEF parameters cdef::search chooses (cdef_params() of the header). */
typedef struct rvpf_cdef_search_result {
  int32_t damping_minus_3; /* min(base_q_idx >> 6, 3) */
  int32_t bits;            /* cdef_bits: 1 << bits strength pairs */
  int32_t y_pri[8], y_sec[8]; /* as coded: secondary 3 means strength 4 */
  int32_t uv_pri[8], uv_sec[8];
} rvpf_cdef_search_result;

/*
 * CDEF strengths: every candidate (luma: primary {0,1,2,4,6,9,12}, or
 * {0,1,2,3,4,5,6,8,10,12,15} when `thorough`, x secondary {0,1,2,3};
 * chroma: {0,1,2,4,7} or {0,1,2,3,4,6,8,11} x {0,1,2}) filters every 8x8
 * block CDEF would filter in `deblocked`, with the decoder's direction
 * search and filter, and its squared error against `source` (inside the
 * frame) is summed per 64x64 block. Up to eight (luma, chroma) pairs are
 * then chosen greedily, for 0..3 bits, minimising the total error plus
 * lambda * bits per block, and each block takes the pair with its least
 * error. The 64x64 blocks that code an index are those with s->cdef_idx
 * >= 0 (device memory, layout (r >> 4) * cdef_stride + (c >> 4));
 * `cdef_idx_out` (host memory, the same layout and size) receives them
 * with their chosen index, the other entries as in s->cdef_idx.
 * p's CDEF fields are not read.
 */
int rvpf_cdef_search(rvpf_ctx *ctx, const rvpf_frame_params *p, const rvpf_side_info *s,
                     const rvpf_planes *deblocked, const rvpf_planes *source, int32_t base_q_idx,
                     double lambda, int32_t thorough, rvpf_cdef_search_result *out, int8_t *cdef_idx_out,
                     void *stream);

/*
 * Batched multi-stream path (no counterpart in the reference, whose
 * decoder filters one frame at a time; for servers and transcoders that
 * decode many streams at once).
 *
 * A batch is built once over a fixed set of frames — any mix of sizes,
 * bit depths, subsamplings and sample sizes — and their caller-owned
 * buffers. Every stage then runs as one launch over all frames (a fixed
 * grid of workgroups taking tiles from a device queue), the launches are
 * captured in one HIP graph, and rvpf_batch_run replays it. To filter new
 * content, write it into the same buffers and run again. The result of
 * every frame equals rvpf_postfilter_frame's on the same inputs.
 */
typedef struct rvpf_batch_frame {
  const rvpf_frame_params *params; /* host memory, read by rvpf_batch_create */
  rvpf_side_info side;             /* device pointers, as for the per-frame calls */
  rvpf_planes recon;               /* deblocked in place */
  rvpf_planes ref_out;             /* the reference frame */
  rvpf_planes display;             /* data[0] == NULL: the frame is not shown */
} rvpf_batch_frame;

typedef struct rvpf_batch rvpf_batch;

/* Launch shape of one batched stage. */
#define RVPF_TUNE_DEBLOCK 0
#define RVPF_TUNE_CDEF 1
#define RVPF_TUNE_LR 2
#define RVPF_TUNE_GRAIN_PREPARE 3
#define RVPF_TUNE_GRAIN_APPLY 4
#define RVPF_TUNE_COPY 5
#define RVPF_TUNE_COUNT 6
typedef struct rvpf_tile_config {
  int32_t shape;            /* tile shape: deblock 0..2 = 64x4, 32x8, 16x16 edges;
                               other stages 0 (one shape each) */
  int32_t workgroups_per_cu;/* persistent grid = CUs x this; 0 = one workgroup per grab */
  int32_t chunk;            /* tiles taken per queue grab, >= 1 */
  float us;                 /* measured stage time of this choice (output; 0 = not measured) */
} rvpf_tile_config;

/* Validates every frame as rvpf_postfilter_frame would (the message of a
 * failure names the frame) and allocates the batch's device scratch. */
int rvpf_batch_create(rvpf_ctx *ctx, const rvpf_batch_frame *frames, int32_t n, rvpf_batch **out);
void rvpf_batch_destroy(rvpf_batch *b);
/*
 * Autotuner: times each stage over the batch's own frames for every
 * candidate tile shape and grid, keeps the fastest, and remembers the
 * choice in the context for batches of the same shape (frame sizes, bit
 * depths, sample sizes and enabled filters), which then skip the timing.
 * force != 0 times again. The timing runs the stages on the batch's
 * buffers: deblocking works in place, so refill `recon` before the next
 * rvpf_batch_run. Synchronous.
 */
int rvpf_batch_autotune(rvpf_ctx *ctx, rvpf_batch *b, int force, void *stream);
int rvpf_batch_get_tiles(const rvpf_batch *b, rvpf_tile_config out[RVPF_TUNE_COUNT]);
int rvpf_batch_set_tiles(rvpf_batch *b, const rvpf_tile_config in[RVPF_TUNE_COUNT]);
/* Enqueues the whole batch (one graph launch; the graph is captured on
 * the first run after a tile change). */
int rvpf_batch_run(rvpf_ctx *ctx, rvpf_batch *b, void *stream);

/*
 * Differentiable loop restoration, for the encoder's restoration search
 * (the reference fits Wiener taps and self-guided weights by least squares
 * in src/encoder/lr.rs:226-430 and judges them with the exact filter,
 * :64-205). These calls give the squared error of float-relaxed filters
 * (no rounding, no clipping) and its gradient with respect to each unit's
 * taps, for gradient-based fitting; rvpf_lrd_rect_sse and rvpf_lr judge
 * the quantised result exactly.
 *
 * Work items are rectangles of at most 64 x 64 samples, one stripe of one
 * restoration unit each (the rectangles of lr.rs:82-127, split every 64
 * columns). Per item the caller-allocated float buffers hold a 70 x 70
 * window (the rectangle with a 3-sample border, 7.17.6's stripe rule) and
 * 64 x 64 tiles (row stride 64) of the CDEF output u, the source, the
 * residual and the self-guided filter outputs. All pointers are device
 * pointers; items with an out-of-range unit or plane count as empty.
 */
typedef struct rvpf_lrd_item {
  int32_t plane;
  int32_t x0, y0;                   /* first sample of the rectangle */
  int32_t w, h;                     /* 1..64 */
  int32_t stripe_start, stripe_end; /* StripeStartY / StripeEndY of its stripe */
  int32_t unit;                     /* index of its unit in the per-unit arrays */
} rvpf_lrd_item;

/* win[n][70][70], u[n][64][64], src[n][64][64] from the deblocked, CDEF
 * and source planes (LrCtx::sample, src/decoder/postfilter.rs:584-596). */
int rvpf_lrd_gather(rvpf_ctx *ctx, const rvpf_frame_params *p, const rvpf_planes *deblocked,
                    const rvpf_planes *cdef, const rvpf_planes *source, const rvpf_lrd_item *items,
                    int32_t n, int32_t nunits, float *win, float *u, float *src, void *stream);
/* Relaxed Wiener filter with taps[unit][2][3] (vertical, horizontal; c0..c2
 * as fractions of 128, the model of fit_wiener, lr.rs:240-325).
 * backward == 0: resid[n][64][64] = out - src, out[n] = squared error;
 * backward != 0: out[n][6] = d(squared error)/d(v0, v1, v2, h0, h1, h2)
 * from the forward residual. */
int rvpf_lrd_wiener(rvpf_ctx *ctx, int backward, const rvpf_lrd_item *items, int32_t n, int32_t nunits,
                    const float *win, const float *src, const float *taps, float *resid, double *out,
                    void *stream);
/* Relaxed self-guided filter outputs flt[n][2][64][64] (in 1/16 sample
 * units; 16 u for a pass whose radius is 0) for the set of each unit
 * (sets[unit], 0..15; box_filter_rect, postfilter.rs:901-991). */
int rvpf_lrd_sgr_filter(rvpf_ctx *ctx, int32_t bit_depth, const rvpf_lrd_item *items, int32_t n,
                        int32_t nunits, const float *win, const float *u, const int32_t *sets, float *flt,
                        void *stream);
/* Relaxed projection with w[unit][2] = (w0, w2) (7.17.2, fit_sgr's model,
 * lr.rs:332-430): forward / backward as rvpf_lrd_wiener, out[n][2] = the
 * gradient with respect to (w0, w2). */
int rvpf_lrd_sgr(rvpf_ctx *ctx, int backward, const rvpf_lrd_item *items, int32_t n, int32_t nunits,
                 const float *flt, const float *u, const float *src, const float *w, float *resid, double *out,
                 void *stream);
/* AC_QLOOKUP[(bit_depth - 8) >> 1][qindex] (src/tables.rs), from which
 * the search's Lagrange multiplier is derived (src/encoder/mod.rs:1449-1455:
 * lambda = 0.1 (AC_QLOOKUP / 8)^2). Returns -1 for arguments out of range. */
int32_t rvpf_ac_q(int32_t bit_depth, int32_t qindex);
/* Exact squared error between two frames' samples over each item:
 * out[n] (the judge's error, lr.rs:128-153). */
int rvpf_lrd_rect_sse(rvpf_ctx *ctx, const rvpf_frame_params *p, const rvpf_lrd_item *items, int32_t n,
                      int32_t nunits, const rvpf_planes *a, const rvpf_planes *b, double *out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* RVPF_H */
