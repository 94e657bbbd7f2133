/*
 * batch.hip — the batched multi-stream path of include/rvpf.h.
 *
 * A batch owns, for a fixed set of frames, one launch record per (frame,
 * stage kernel) in a device table, the prefix sums of their tile counts,
 * one work queue per kernel, and the per-frame scratch (CDEF planes that
 * loop restoration reads, film grain templates). Frames are grouped by
 * sample size (8-bit samples or 16-bit samples); each group's chain runs
 * on its own graph branch:
 *
 *   deblock V -> deblock H -> CDEF (or its copy) -> LR -> grain apply (or crop)
 *
 * with the film grain templates of every frame built on a third branch
 * that joins before the applies. Each kernel is one launch over all the
 * group's frames: a grid of workgroups (CUs x workgroups_per_cu) takes
 * `chunk` tiles at a time from its queue, so frames of any size share the
 * GPU without a launch per frame. The queues re-arm themselves at the end
 * of every launch (no memset node in the graph).
 *
 * Per frame the stages do exactly what rvpf_postfilter_frame does, with
 * the same kernels' tile functions, so the outputs are equal.
 */
#include <stdio.h>
#include <string.h>

#include <new>
#include <vector>

#include "rvpf_host.h"

using namespace rvpf;
using namespace rvpf::host;

namespace {

/* Kernels of a group's chain. K_LF stands for both deblocking passes: the
 * horizontal pass reads the same records with a queue of its own (K_LFH). */
enum Kernel { K_LF, K_CDEF, K_COPY_MID, K_LR, K_APPLY, K_COPY_OUT, K_LFH, K_N };

struct FrameIn {
  rvpf_frame_params p;
  rvpf_side_info s;
  rvpf_planes recon, ref, disp;
  bool shown;
  PlanePtr cdef[3]; /* where CDEF writes each plane (scratch if restored next) */
  GrainScratch gs;  /* film grain scratch (shown frames with grain) */
};

struct Work {
  int total = 0;         /* tiles */
  size_t rec_off = 0;    /* byte offset of the records in the table */
  size_t start_off = 0;  /* byte offset of the n + 1 prefix sums */
  int n = 0;             /* frames */
  int queue = 0;         /* index of the kernel's 2-int queue */
  int vec = 1;           /* CDEF: every frame's planes take the chunked staging */
};

struct Group {
  int bps = 2;
  std::vector<int> frames;
  Work w[K_N];
};

constexpr rvpf_tile_config DEFAULT_TILES[RVPF_TUNE_COUNT] = {
    {0, 8, 1, 0.f}, /* deblock */
    {0, 4, 1, 0.f}, /* CDEF */
    {0, 2, 1, 0.f}, /* LR */
    {0, 1, 1, 0.f}, /* grain prepare */
    {0, 8, 1, 0.f}, /* grain apply */
    {0, 8, 1, 0.f}, /* copy */
};

int shapes_of(int stage) {
  return stage == RVPF_TUNE_DEBLOCK ? LF_SHAPES : (stage == RVPF_TUNE_GRAIN_APPLY ? GRAIN_SHAPES : 1);
}

}  // namespace

struct rvpf_batch {
  rvpf_ctx *ctx = nullptr;
  int n = 0;
  std::vector<FrameIn> f;
  std::vector<Group> groups;
  Work prep; /* grain templates of every shown frame with grain */
  rvpf_tile_config cfg[RVPF_TUNE_COUNT];
  uint64_t key = 0;
  std::vector<char> host_table;
  void *table = nullptr;
  size_t table_cap = 0;
  int *queues = nullptr;
  int nqueues = 0;
  void *cdef_mem = nullptr;
  void *grain_mem = nullptr;
  hipStream_t cs[3] = {nullptr, nullptr, nullptr}; /* group 0 (+ capture origin), group 1, grain prepare */
  hipEvent_t ev_fork = nullptr, ev_group = nullptr, ev_prep = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

namespace {

size_t align_up(size_t v, size_t a) { return (v + a - 1) & ~(a - 1); }

size_t record_size(int k) {
  switch (k) {
    case K_LF:
    case K_LFH: return lf_record_size();
    case K_CDEF: return cdef_record_size();
    case K_LR: return lr_record_size();
    case K_APPLY: return grain_apply_record_size();
    default: return copy_record_size();
  }
}

/* Plans kernel k of frame x into rec; returns its tiles (0: nothing to do). */
int plan_frame(const rvpf_batch *b, int k, const FrameIn &x, void *rec) {
  const rvpf_frame_params *p = &x.p;
  const Geom g = make_geom(p);
  PlanePtr rec_[3], ref[3], disp[3];
  to_ptrs(&x.recon, rec_);
  to_ptrs(&x.ref, ref);
  to_ptrs(x.shown ? &x.disp : nullptr, disp);
  memset(rec, 0, record_size(k));
  switch (k) {
    case K_LF: {
      if (!lf_on(p)) return 0;
      LfSide side;
      side.mi = x.s.mi;
      side.mi_stride = x.s.mi_stride;
      for (int i = 0; i < 3; i++) side.tx[i] = x.s.tx_size[i];
      return lf_plan(g, make_lf(p), side, rec_, b->cfg[RVPF_TUNE_DEBLOCK].shape, rec);
    }
    case K_CDEF:
      if (!cdef_on(p)) return 0;
      return cdef_plan(g, make_cdef(p), x.s.mi, x.s.mi_stride, x.s.cdef_idx, x.s.cdef_stride, 1, rec_, x.cdef, rec);
    case K_COPY_MID: {
      /* CDEF off: planes not restored next go to the reference as they are */
      if (cdef_on(p)) return 0;
      int w[3] = {0, 0, 0}, h[3] = {0, 0, 0};
      for (int i = 0; i < g.nplanes; i++)
        if (!lr_plane_on(p, i)) area(p, i, &w[i], &h[i]);
      return copy_plan(g, rec_, ref, w, h, 0, rec);
    }
    case K_LR:
      if (!lr_on(p)) return 0;
      return lr_plan(g, make_lr(p, &x.s), rec_, x.cdef, ref, x.recon.bytes_per_sample, rec);
    case K_APPLY:
      if (!x.shown || !grain_on(p)) return 0;
      return grain_apply_plan(make_grain(p), x.gs, ref, disp, x.recon.bytes_per_sample, rec);
    case K_COPY_OUT: {
      if (!x.shown || grain_on(p)) return 0;
      int w[3] = {0, 0, 0}, h[3] = {0, 0, 0};
      for (int i = 0; i < g.nplanes; i++) crop(p, i, &w[i], &h[i]);
      return copy_plan(g, ref, disp, w, h, 0, rec);
    }
  }
  return 0;
}

/* Appends n records of `rs` bytes + n + 1 prefix sums to the host table. */
void place(std::vector<char> &t, Work &w, size_t rs, const std::vector<char> &recs, const std::vector<int> &start) {
  w.rec_off = align_up(t.size(), 256);
  t.resize(w.rec_off + recs.size());
  memcpy(t.data() + w.rec_off, recs.data(), recs.size());
  w.start_off = align_up(t.size(), 256);
  t.resize(w.start_off + start.size() * sizeof(int));
  memcpy(t.data() + w.start_off, start.data(), start.size() * sizeof(int));
  (void)rs;
}

/* Plans every kernel of every frame with the current tiles and uploads the
 * table (synchronously; nothing of the batch may be in flight). */
int build_tables(rvpf_batch *b) {
  rvpf_ctx *ctx = b->ctx;
  std::vector<char> &t = b->host_table;
  t.clear();
  int q = 0;
  for (Group &gr : b->groups) {
    const int nf = (int)gr.frames.size();
    for (int k = 0; k < K_N; k++) {
      Work &w = gr.w[k];
      w.n = nf;
      w.queue = q++;
      if (k == K_LFH) continue;
      const size_t rs = record_size(k);
      std::vector<char> recs(rs * nf);
      std::vector<int> start(nf + 1, 0);
      w.vec = 1;
      for (int i = 0; i < nf; i++) {
        const int items = plan_frame(b, k, b->f[gr.frames[i]], recs.data() + rs * i);
        start[i + 1] = start[i] + items;
        if (k == K_CDEF && items > 0 && !cdef_record_vec(recs.data() + rs * i)) w.vec = 0;
      }
      w.total = start[nf];
      place(t, w, rs, recs, start);
    }
    /* the horizontal pass shares the vertical pass's records and tiles */
    const int qh = gr.w[K_LFH].queue;
    gr.w[K_LFH] = gr.w[K_LF];
    gr.w[K_LFH].queue = qh;
  }
  {
    std::vector<int> idx;
    for (int i = 0; i < b->n; i++)
      if (b->f[i].shown && grain_on(&b->f[i].p)) idx.push_back(i);
    const size_t rs = grain_prep_record_size();
    std::vector<char> recs(rs * idx.size());
    std::vector<int> start(idx.size() + 1, 0);
    for (size_t i = 0; i < idx.size(); i++) {
      const FrameIn &x = b->f[idx[i]];
      start[i + 1] = start[i] + grain_prep_plan(make_grain(&x.p), x.gs, recs.data() + rs * i);
    }
    b->prep.n = (int)idx.size();
    b->prep.total = start[idx.size()];
    b->prep.queue = q++;
    place(t, b->prep, rs, recs, start);
  }
  hipError_t e = hipSuccess;
  if (t.size() > b->table_cap) {
    if (b->table) (void)hipFree(b->table);
    b->table = nullptr;
    b->table_cap = 0;
    e = hipMalloc(&b->table, t.size());
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return fail(ctx, RVPF_E_NOMEM, "cannot allocate the batch tables");
    }
    b->table_cap = t.size();
  }
  if (q > b->nqueues) {
    if (b->queues) (void)hipFree(b->queues);
    b->queues = nullptr;
    b->nqueues = 0;
    e = hipMalloc(&b->queues, (size_t)q * 2 * sizeof(int));
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return fail(ctx, RVPF_E_NOMEM, "cannot allocate the batch queues");
    }
    b->nqueues = q;
  }
  if ((e = hipMemcpy(b->table, t.data(), t.size(), hipMemcpyHostToDevice)) != hipSuccess)
    return hip_fail(ctx, e, "batch table upload");
  if ((e = hipMemset(b->queues, 0, (size_t)b->nqueues * 2 * sizeof(int))) != hipSuccess)
    return hip_fail(ctx, e, "batch queue reset");
  if ((e = hipDeviceSynchronize()) != hipSuccess) return hip_fail(ctx, e, "batch table upload");
  return RVPF_OK;
}

BatchWork batch_work(const rvpf_batch *b, const Work &w, int chunk) {
  BatchWork bw;
  bw.rec = static_cast<const char *>(b->table) + w.rec_off;
  bw.start = reinterpret_cast<const int *>(static_cast<const char *>(b->table) + w.start_off);
  bw.n = w.n;
  bw.total = w.total;
  bw.chunk = chunk;
  bw.queue = b->queues + 2 * w.queue;
  return bw;
}

StagePlan stage_plan(const rvpf_batch *b, int stage, int total) {
  const rvpf_tile_config &c = b->cfg[stage];
  StagePlan sp;
  sp.shape = c.shape;
  sp.chunk = c.chunk < 1 ? 1 : c.chunk;
  const int grabs = (total + sp.chunk - 1) / sp.chunk;
  int grid = c.workgroups_per_cu > 0 ? b->ctx->cus * c.workgroups_per_cu : grabs;
  if (grid > grabs) grid = grabs;
  sp.grid = grid < 1 ? 1 : grid;
  return sp;
}

/* Enqueues kernel k of group gr (both passes for K_LF). */
int launch_kernel(rvpf_batch *b, const Group &gr, int k, hipStream_t st) {
  const Work &w = gr.w[k];
  if (w.total == 0) return 0;
  switch (k) {
    case K_LF: {
      const StagePlan sp = stage_plan(b, RVPF_TUNE_DEBLOCK, w.total);
      return launch_deblock_batch(batch_work(b, w, sp.chunk), batch_work(b, gr.w[K_LFH], sp.chunk), sp, gr.bps, st);
    }
    case K_CDEF: {
      const StagePlan sp = stage_plan(b, RVPF_TUNE_CDEF, w.total);
      return launch_cdef_batch(batch_work(b, w, sp.chunk), sp, gr.bps, w.vec, st);
    }
    case K_LR: {
      const StagePlan sp = stage_plan(b, RVPF_TUNE_LR, w.total);
      return launch_lr_batch(batch_work(b, w, sp.chunk), sp, gr.bps, st);
    }
    case K_APPLY: {
      const StagePlan sp = stage_plan(b, RVPF_TUNE_GRAIN_APPLY, w.total);
      return launch_grain_apply_batch(batch_work(b, w, sp.chunk), sp, gr.bps, st);
    }
    case K_COPY_MID:
    case K_COPY_OUT: {
      const StagePlan sp = stage_plan(b, RVPF_TUNE_COPY, w.total);
      return launch_copy_batch(batch_work(b, w, sp.chunk), sp, gr.bps, st);
    }
  }
  return 0;
}

int launch_prep(rvpf_batch *b, hipStream_t st) {
  if (b->prep.total == 0) return 0;
  const StagePlan sp = stage_plan(b, RVPF_TUNE_GRAIN_PREPARE, b->prep.total);
  return launch_grain_prepare_batch(batch_work(b, b->prep, sp.chunk), sp, st);
}

/* Enqueues one tunable stage for every group (autotuner timing). */
int launch_stage(rvpf_batch *b, int stage, hipStream_t st) {
  int rc = 0;
  if (stage == RVPF_TUNE_GRAIN_PREPARE) return launch_prep(b, st);
  for (const Group &gr : b->groups) {
    switch (stage) {
      case RVPF_TUNE_DEBLOCK: rc = launch_kernel(b, gr, K_LF, st); break;
      case RVPF_TUNE_CDEF: rc = launch_kernel(b, gr, K_CDEF, st); break;
      case RVPF_TUNE_LR: rc = launch_kernel(b, gr, K_LR, st); break;
      case RVPF_TUNE_GRAIN_APPLY: rc = launch_kernel(b, gr, K_APPLY, st); break;
      case RVPF_TUNE_COPY:
        rc = launch_kernel(b, gr, K_COPY_MID, st);
        if (!rc) rc = launch_kernel(b, gr, K_COPY_OUT, st);
        break;
    }
    if (rc) return rc;
  }
  return rc;
}

bool stage_has_work(const rvpf_batch *b, int stage) {
  if (stage == RVPF_TUNE_GRAIN_PREPARE) return b->prep.total > 0;
  for (const Group &gr : b->groups) {
    switch (stage) {
      case RVPF_TUNE_DEBLOCK: if (gr.w[K_LF].total) return true; break;
      case RVPF_TUNE_CDEF: if (gr.w[K_CDEF].total) return true; break;
      case RVPF_TUNE_LR: if (gr.w[K_LR].total) return true; break;
      case RVPF_TUNE_GRAIN_APPLY: if (gr.w[K_APPLY].total) return true; break;
      case RVPF_TUNE_COPY: if (gr.w[K_COPY_MID].total || gr.w[K_COPY_OUT].total) return true; break;
    }
  }
  return false;
}

void drop_graph(rvpf_batch *b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

/* Captures the whole batch into one graph. */
int capture(rvpf_batch *b) {
  rvpf_ctx *ctx = b->ctx;
  hipStream_t s0 = b->cs[0];
  hipError_t e = hipStreamBeginCapture(s0, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return hip_fail(ctx, e, "graph capture");
  int rc = 0;
  const bool two = b->groups.size() > 1;
  const bool prep = b->prep.total > 0;
  e = hipEventRecord(b->ev_fork, s0);
  if (e == hipSuccess && prep) e = hipStreamWaitEvent(b->cs[2], b->ev_fork, 0);
  if (e == hipSuccess && prep) rc = launch_prep(b, b->cs[2]);
  if (e == hipSuccess && !rc && prep) e = hipEventRecord(b->ev_prep, b->cs[2]);
  if (e == hipSuccess && !rc && two) e = hipStreamWaitEvent(b->cs[1], b->ev_fork, 0);
  for (size_t g = 0; g < b->groups.size() && e == hipSuccess && !rc; g++) {
    const Group &gr = b->groups[g];
    hipStream_t st = g == 0 ? s0 : b->cs[1];
    const int pre[4] = {K_LF, K_CDEF, K_COPY_MID, K_LR};
    for (int i = 0; i < 4 && !rc; i++) rc = launch_kernel(b, gr, pre[i], st);
    if (!rc && prep) e = hipStreamWaitEvent(st, b->ev_prep, 0);
    if (!rc && e == hipSuccess) rc = launch_kernel(b, gr, K_APPLY, st);
    if (!rc && e == hipSuccess) rc = launch_kernel(b, gr, K_COPY_OUT, st);
  }
  if (e == hipSuccess && !rc && two) e = hipEventRecord(b->ev_group, b->cs[1]);
  if (e == hipSuccess && !rc && two) e = hipStreamWaitEvent(s0, b->ev_group, 0);
  hipGraph_t graph = nullptr;
  hipError_t e2 = hipStreamEndCapture(s0, &graph);
  if (e != hipSuccess || rc || e2 != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    if (rc) return hip_fail(ctx, (hipError_t)rc, "batch kernel launch during capture");
    return hip_fail(ctx, e != hipSuccess ? e : e2, "graph capture");
  }
  b->graph = graph;
  e = hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    drop_graph(b);
    return hip_fail(ctx, e, "graph instantiate");
  }
  return RVPF_OK;
}

uint64_t fnv(uint64_t h, int64_t v) {
  for (int i = 0; i < 8; i++) {
    h ^= (uint64_t)((v >> (8 * i)) & 0xff);
    h *= 1099511628211ull;
  }
  return h;
}

/* What the best tiles depend on: the frames' shapes and enabled filters. */
uint64_t batch_key(const rvpf_batch *b) {
  uint64_t h = 1469598103934665603ull;
  h = fnv(h, b->n);
  for (const FrameIn &x : b->f) {
    const rvpf_frame_params &p = x.p;
    h = fnv(h, p.frame_width);
    h = fnv(h, p.frame_height);
    h = fnv(h, p.bit_depth);
    h = fnv(h, p.ssx * 4 + p.ssy * 2 + p.mono_chrome);
    h = fnv(h, x.recon.bytes_per_sample);
    h = fnv(h, lf_on(&p) * 8 + cdef_on(&p) * 4 + (x.shown && grain_on(&p)) * 2 + x.shown);
    for (int i = 0; i < 3; i++) h = fnv(h, p.frame_restoration_type[i] * 1024 + p.loop_restoration_size[i]);
  }
  return h;
}

int check_tiles(rvpf_ctx *ctx, const rvpf_tile_config *c) {
  for (int s = 0; s < RVPF_TUNE_COUNT; s++) {
    if (c[s].shape < 0 || c[s].shape >= shapes_of(s) || c[s].workgroups_per_cu < 0 || c[s].workgroups_per_cu > 32 ||
        c[s].chunk < 1 || c[s].chunk > 64) {
      char msg[96];
      snprintf(msg, sizeof msg, "tile config of stage %d out of range", s);
      return fail(ctx, RVPF_E_INVALID, msg);
    }
  }
  return RVPF_OK;
}

}  // namespace

extern "C" {

int rvpf_batch_create(rvpf_ctx *ctx, const rvpf_batch_frame *frames, int32_t n, rvpf_batch **out) {
  if (!ctx) return fail(nullptr, RVPF_E_INVALID, "ctx is NULL");
  if (!out || !frames || n < 1) return fail(ctx, RVPF_E_INVALID, "batch: frames, n or out invalid");
  *out = nullptr;
  rvpf_batch *b = new (std::nothrow) rvpf_batch();
  if (!b) return fail(ctx, RVPF_E_NOMEM, "cannot allocate the batch");
  b->ctx = ctx;
  b->n = n;
  memcpy(b->cfg, DEFAULT_TILES, sizeof b->cfg);
  int rc = RVPF_OK;
  char msg[256];
  size_t cdef_bytes = 0, grain_bytes = 0;
  std::vector<size_t> cdef_off(n), grain_off(n);
  try {
    b->f.resize(n);
  } catch (...) {
    delete b;
    return fail(ctx, RVPF_E_NOMEM, "cannot allocate the batch");
  }
  for (int i = 0; i < n && rc == RVPF_OK; i++) {
    const rvpf_batch_frame &in = frames[i];
    FrameIn &x = b->f[i];
    if (!in.params) {
      rc = fail(ctx, RVPF_E_INVALID, "batch: params is NULL");
      break;
    }
    x.p = *in.params;
    x.s = in.side;
    x.recon = in.recon;
    x.ref = in.ref_out;
    x.disp = in.display;
    x.shown = in.display.data[0] != nullptr;
    const rvpf_frame_params *p = &x.p;
    int r;
    if (!(r = validate_params(p, msg)) && p->use_superres) {
      snprintf(msg, sizeof msg, "super-resolution frames are not batched (use rvpf_postfilter_frame)");
      r = RVPF_E_UNSUPPORTED;
    }
    if (r || (r = validate_side(p, &x.s, msg)) ||
        (r = validate_planes(p, &x.recon, "recon", 0, lf_on(p) ? 8 : 0, msg)) ||
        (r = validate_planes(p, &x.ref, "ref_out", 0, 0, msg, true)) ||
        (x.shown && (r = validate_planes(p, &x.disp, "display", 1, 0, msg)))) {
      char full[256];
      snprintf(full, sizeof full, "batch frame %d: %.200s", i, msg);
      rc = fail(ctx, r, full);
      break;
    }
    const int bps = x.recon.bytes_per_sample;
    if (x.ref.bytes_per_sample != bps || (x.shown && x.disp.bytes_per_sample != bps)) {
      snprintf(msg, sizeof msg, "batch frame %d: plane sample sizes differ", i);
      rc = fail(ctx, RVPF_E_INVALID, msg);
      break;
    }
    /* CDEF scratch for the planes loop restoration reads */
    cdef_off[i] = cdef_bytes;
    if (cdef_on(p) && lr_on(p))
      for (int pl = 0; pl < nplanes_of(p); pl++) {
        if (!lr_plane_on(p, pl)) continue;
        int w, h;
        area(p, pl, &w, &h);
        cdef_bytes += align_up((size_t)((w + 63) & ~63) * h * bps, 256);
      }
    grain_off[i] = grain_bytes;
    if (x.shown && grain_on(p)) {
      int stripes, blocks;
      grain_offsets_dims(p->upscaled_width, p->frame_height, &stripes, &blocks);
      grain_bytes += align_up(GRAIN_TMPL_BYTES, 256) + align_up(GRAIN_LUT_BYTES, 256) +
                     align_up((size_t)stripes * blocks, 256);
    }
  }
  hipError_t e = hipSuccess;
  if (rc == RVPF_OK && cdef_bytes) e = hipMalloc(&b->cdef_mem, cdef_bytes);
  if (rc == RVPF_OK && e == hipSuccess && grain_bytes) e = hipMalloc(&b->grain_mem, grain_bytes);
  if (rc == RVPF_OK && e != hipSuccess) {
    (void)hipGetLastError();
    rc = fail(ctx, RVPF_E_NOMEM, "cannot allocate the batch scratch");
  }
  if (rc == RVPF_OK) {
    for (int i = 0; i < n; i++) {
      FrameIn &x = b->f[i];
      const rvpf_frame_params *p = &x.p;
      const int bps = x.recon.bytes_per_sample;
      PlanePtr ref[3];
      to_ptrs(&x.ref, ref);
      size_t off = cdef_off[i];
      for (int pl = 0; pl < 3; pl++) {
        x.cdef[pl] = ref[pl];
        if (pl < nplanes_of(p) && cdef_on(p) && lr_plane_on(p, pl)) {
          int w, h;
          area(p, pl, &w, &h);
          x.cdef[pl].data = static_cast<char *>(b->cdef_mem) + off;
          x.cdef[pl].stride = (w + 63) & ~63;
          off += align_up((size_t)x.cdef[pl].stride * h * bps, 256);
        } else if (pl < nplanes_of(p) && !cdef_on(p)) {
          /* CDEF is the identity: loop restoration reads the deblocked planes */
          x.cdef[pl].data = x.recon.data[pl];
          x.cdef[pl].stride = x.recon.stride[pl];
        }
      }
      memset(&x.gs, 0, sizeof x.gs);
      if (x.shown && grain_on(p)) {
        int stripes, blocks;
        grain_offsets_dims(p->upscaled_width, p->frame_height, &stripes, &blocks);
        char *base = static_cast<char *>(b->grain_mem) + grain_off[i];
        x.gs.tmpl = reinterpret_cast<int16_t *>(base);
        base += align_up(GRAIN_TMPL_BYTES, 256);
        x.gs.lut = reinterpret_cast<int16_t *>(base);
        base += align_up(GRAIN_LUT_BYTES, 256);
        x.gs.offsets = reinterpret_cast<uint8_t *>(base);
        x.gs.max_stripes = stripes;
        x.gs.max_blocks = blocks;
      }
    }
    /* groups by sample size */
    for (int bps = 2; bps >= 1; bps--) {
      Group gr;
      gr.bps = bps;
      for (int i = 0; i < n; i++)
        if (b->f[i].recon.bytes_per_sample == bps) gr.frames.push_back(i);
      if (!gr.frames.empty()) b->groups.push_back(gr);
    }
    b->key = batch_key(b);
    for (int i = 0; i < 3 && e == hipSuccess; i++) e = hipStreamCreateWithFlags(&b->cs[i], hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_group, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->ev_prep, hipEventDisableTiming);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      rc = hip_fail(ctx, e, "batch streams");
    }
  }
  if (rc == RVPF_OK) {
    for (const TuneEntry &t : ctx->tune_cache)
      if (t.key == b->key) memcpy(b->cfg, t.cfg, sizeof b->cfg);
    rc = build_tables(b);
  }
  if (rc != RVPF_OK) {
    rvpf_batch_destroy(b);
    return rc;
  }
  *out = b;
  return RVPF_OK;
}

void rvpf_batch_destroy(rvpf_batch *b) {
  if (!b) return;
  (void)hipDeviceSynchronize();
  drop_graph(b);
  for (int i = 0; i < 3; i++)
    if (b->cs[i]) (void)hipStreamDestroy(b->cs[i]);
  if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
  if (b->ev_group) (void)hipEventDestroy(b->ev_group);
  if (b->ev_prep) (void)hipEventDestroy(b->ev_prep);
  if (b->table) (void)hipFree(b->table);
  if (b->queues) (void)hipFree(b->queues);
  if (b->cdef_mem) (void)hipFree(b->cdef_mem);
  if (b->grain_mem) (void)hipFree(b->grain_mem);
  delete b;
}

int rvpf_batch_get_tiles(const rvpf_batch *b, rvpf_tile_config out[RVPF_TUNE_COUNT]) {
  if (!b || !out) return fail(nullptr, RVPF_E_INVALID, "batch or out is NULL");
  memcpy(out, b->cfg, sizeof b->cfg);
  return RVPF_OK;
}

int rvpf_batch_set_tiles(rvpf_batch *b, const rvpf_tile_config in[RVPF_TUNE_COUNT]) {
  if (!b || !in) return fail(nullptr, RVPF_E_INVALID, "batch or in is NULL");
  int rc = check_tiles(b->ctx, in);
  if (rc) return rc;
  hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(b->ctx, e, "synchronize");
  drop_graph(b);
  memcpy(b->cfg, in, sizeof b->cfg);
  return build_tables(b);
}

int rvpf_batch_autotune(rvpf_ctx *ctx, rvpf_batch *b, int force, void *stream) {
  if (!ctx || !b) return fail(ctx, RVPF_E_INVALID, "ctx or batch is NULL");
  hipError_t e = hipStreamSynchronize((hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "synchronize");
  if (!force)
    for (const TuneEntry &t : ctx->tune_cache)
      if (t.key == b->key) return rvpf_batch_set_tiles(b, t.cfg);
  if ((e = hipDeviceSynchronize()) != hipSuccess) return hip_fail(ctx, e, "synchronize");
  drop_graph(b);
  hipStream_t st = b->cs[0];
  hipEvent_t t0 = nullptr, t1 = nullptr;
  e = hipEventCreate(&t0);
  if (e == hipSuccess) e = hipEventCreate(&t1);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    if (t0) (void)hipEventDestroy(t0);
    return hip_fail(ctx, e, "autotune events");
  }
  static const int WPC[] = {0, 1, 2, 4, 8};
  static const int CHUNK[] = {1, 2, 4};
  int rc = RVPF_OK;
  for (int s = 0; s < RVPF_TUNE_COUNT && rc == RVPF_OK; s++) {
    if (!stage_has_work(b, s)) continue;
    rvpf_tile_config best = b->cfg[s];
    best.us = -1.f;
    for (int shape = 0; shape < shapes_of(s) && rc == RVPF_OK; shape++)
      for (int wi = 0; wi < 5 && rc == RVPF_OK; wi++)
        for (int ci = 0; ci < 3 && rc == RVPF_OK; ci++) {
          if (WPC[wi] == 0 && CHUNK[ci] != 1) continue;
          if (s == RVPF_TUNE_GRAIN_PREPARE && (CHUNK[ci] != 1 || WPC[wi] > 2)) continue;
          rvpf_tile_config c = {shape, WPC[wi], CHUNK[ci], 0.f};
          const bool reshape = c.shape != b->cfg[s].shape;
          b->cfg[s] = c;
          if (reshape && (rc = build_tables(b))) break;
          /* one launch to warm up, then the mean of three */
          rc = launch_stage(b, s, st);
          if (!rc) e = hipEventRecord(t0, st);
          for (int r = 0; r < 3 && !rc && e == hipSuccess; r++) rc = launch_stage(b, s, st);
          if (!rc && e == hipSuccess) e = hipEventRecord(t1, st);
          if (!rc && e == hipSuccess) e = hipEventSynchronize(t1);
          float ms = 0.f;
          if (!rc && e == hipSuccess) e = hipEventElapsedTime(&ms, t0, t1);
          if (rc) rc = hip_fail(ctx, (hipError_t)rc, "autotune launch");
          else if (e != hipSuccess) rc = hip_fail(ctx, e, "autotune timing");
          if (rc) break;
          c.us = ms * 1000.f / 3.f;
          if (best.us < 0.f || c.us < best.us) best = c;
        }
    if (rc) break;
    const bool reshape = best.shape != b->cfg[s].shape;
    b->cfg[s] = best;
    if (reshape) rc = build_tables(b);
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  if (rc) return rc;
  if ((e = hipDeviceSynchronize()) != hipSuccess) return hip_fail(ctx, e, "synchronize");
  bool stored = false;
  for (TuneEntry &t : ctx->tune_cache)
    if (t.key == b->key) {
      memcpy(t.cfg, b->cfg, sizeof b->cfg);
      stored = true;
    }
  if (!stored) {
    TuneEntry t;
    t.key = b->key;
    memcpy(t.cfg, b->cfg, sizeof b->cfg);
    ctx->tune_cache.push_back(t);
  }
  return RVPF_OK;
}

int rvpf_batch_run(rvpf_ctx *ctx, rvpf_batch *b, void *stream) {
  if (!ctx || !b) return fail(ctx, RVPF_E_INVALID, "ctx or batch is NULL");
  int rc;
  if (!b->exec && (rc = capture(b))) return rc;
  hipError_t e = hipGraphLaunch(b->exec, (hipStream_t)stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "graph launch");
  return RVPF_OK;
}

}  // extern "C"
