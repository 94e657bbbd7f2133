#!/usr/bin/env python3
"""Benchmark: AV1 post-filter chain (deblock + CDEF + LR + film grain) on
one MI355X, 4K 10-bit 4:2:0, bit-exact.

A step is one pass of the whole chain (rvpf_postfilter_frame: deblock in
place, CDEF, loop restoration into the reference frame, crop + film grain
into the display frame) over a batch of 4 synthetic frames of a GOP
(qindex 64 / 96 / 160 / 200, global pan 1.5 px/frame), every filter on,
restoration units of 64 mixing Wiener / self-guided / none. Inputs are
resident in HBM before the timed region (one fresh reconstruction per
step and frame, since deblocking works in place).

Prints one JSON line: Mpix/s (luma pixels of the displayed frames per
second), the dominant kernel's HBM roofline fraction (algorithmic bytes
per launch / its average duration, timed with HIP events on the launch
stream), and the CPU oracle's throughput on the host as the baseline.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "rivet-av1_amd"))

HBM_PEAK_GBS = 8000.0
QINDEX = [64, 96, 160, 200]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


CHROMA = {"420": dict(ssx=1, ssy=1), "422": dict(ssx=1, ssy=0), "444": dict(ssx=0, ssy=0), "mono": dict(mono=True)}
CHROMA_NAME = {"420": "4:2:0", "422": "4:2:2", "444": "4:4:4", "mono": "monochrome"}


def make_frames(n, w, h, bd, seed, superres_denom=0, chroma="420"):
    from rvpf import workload
    frames = []
    for i in range(n):
        frames.append(workload.Frame(w, h, bit_depth=bd, **CHROMA[chroma], sb_size=128, qindex=QINDEX[i % 4],
                                     seed=seed, frame_index=i, pan_px=1.5, lf=True, cdef=True, cdef_bits=3,
                                     lr=2, lr_unit_size=64, grain=True, grain_lag=3, grain_y_points=8,
                                     with_source=False, superres_denom=superres_denom))
    return frames


# HBM bytes per launch of each stage's kernels on the default workload,
# from the calibrated FETCH_SIZE / WRITE_SIZE counter passes summarised here
PMC_SUMMARY = "profiles/chain_4k10b_pmc.md"
STAGE_KERNELS = {"deblock": ["lf_fused_kernel<u16>"], "cdef": ["cdef_kernel<u16, true>"],
                 "lr": ["lr_kernel<u16>"], "grain_apply": ["grain_apply_kernel<u16>"],
                 "upscale": ["upscale_kernel<u16>"]}


def measured_traffic(stage):
    """Calibrated HBM bytes per launch of `stage` (None if not recorded)."""
    path = os.path.join(ROOT, PMC_SUMMARY)
    if not os.path.exists(path):
        return None
    rows = {}
    for line in open(path):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 5 and cells[0].startswith("`"):
            try:
                rows[cells[0].strip("`")] = int(cells[4])
            except ValueError:
                pass
    names = STAGE_KERNELS.get(stage, [])
    if not names or any(n not in rows for n in names):
        return None
    return sum(rows[n] for n in names)


def samples_per_frame(f):
    return sum(f.crop_shape(p)[0] * f.crop_shape(p)[1] for p in range(f.num_planes))


def coded_samples_per_frame(f):
    """Samples of the coded frame (before super-resolution upscaling)."""
    n = 0
    for p in range(f.num_planes):
        sx = f.params.ssx if p else 0
        n += f.crop_shape(p)[0] * ((f.params.frame_width + sx) >> sx)
    return n


def frame_layout(f):
    """Chroma layout of a frame, as a name."""
    if f.num_planes == 1:
        return "monochrome"
    return {(1, 1): "4:2:0", (1, 0): "4:2:2", (0, 0): "4:4:4"}[(f.params.ssx, f.params.ssy)]


def host_threads():
    """Host cores this process may run on (nproc)."""
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def _oracle_rate(frames, seconds, threads, pyoracle):
    """Mpix/s of the oracle chain over the frames in turn for about `seconds`."""
    pyoracle.postfilter_frame(frames[0], threads=threads)  # warm-up
    n, t0 = 0, time.perf_counter()
    while True:
        pyoracle.postfilter_frame(frames[n % len(frames)], threads=threads)
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 400:
            break
    px = sum(frames[i % len(frames)].luma_pixels() for i in range(n))
    return px / el / 1e6, n, el


def cpu_baseline(frames, seconds):
    """The oracle chain on the host, threaded like the reference's decoder
    (set_threads(nproc), src/decoder/mod.rs:323-329), and on one thread."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    threads = host_threads()
    what = "%dx%d %d-bit %s chain%s" % (frames[0].params.upscaled_width, frames[0].params.frame_height,
                                        frames[0].params.bit_depth, frame_layout(frames[0]),
                                        " with super-resolution" if frames[0].params.use_superres else "")
    rate, n, el = _oracle_rate(frames, seconds, threads, pyoracle)
    rate1, n1, el1 = _oracle_rate(frames, seconds, 1, pyoracle)
    return {"value": round(rate, 2), "unit": "Mpix/s", "cores": threads, "kind": "port",
            "sample": "%d frames of the same %s (%.1f s), C oracle, %d threads (nproc)" % (n, what, el, threads),
            "cpu_model": cpu_model(),
            "one_thread": {"value": round(rate1, 2), "unit": "Mpix/s", "cores": 1,
                           "sample": "%d frames (%.1f s), 1 thread" % (n1, el1)}}


def batch_streams(n_streams, frames_per_stream, seed):
    """The 32-stream mixed workload (rvpf.workload.batch_stream_configs):
    [(bit depth, [Frame])]."""
    from rvpf import workload
    cfgs = workload.batch_stream_configs(n_streams, frames_per_stream, seed)
    flat = workload.make_frames([kw for _, kws in cfgs for kw in kws])
    out, k = [], 0
    for bd, kws in cfgs:
        out.append((bd, flat[k:k + len(kws)]))
        k += len(kws)
    return out


def run_batch(args, steps, warmup, check_all):
    """Batched path vs the plain per-frame drop-in on the same frames of
    the 32-stream workload; returns the figures as a dict. check_all: every
    frame of the batched output against the oracle (else the smallest)."""
    import numpy as np
    import torch
    from rvpf import api, workload
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    log("generating %d streams x %d frames" % (args.streams, args.stream_frames))
    streams = batch_streams(args.streams, args.stream_frames, args.seed)
    ctx = api.Context()
    items = []
    for bd, frames in streams:
        for f in frames:
            df = api.DeviceFrame(f, bytes_per_sample=1 if bd == 8 else 2)
            pristine = df.upload_planes(f.rec)
            items.append({"f": f, "df": df, "pristine": pristine,
                          "rec": [None if t is None else t.clone() for t in pristine],
                          "ref": df.alloc_like_recon(), "disp": df.alloc_display()})
    luma_px = sum(it["f"].luma_pixels() for it in items)
    log("%d frames, %.1f Mpix" % (len(items), luma_px / 1e6))
    stream = torch.cuda.current_stream()

    def refill():
        for it in items:
            for t, s_ in zip(it["rec"], it["pristine"]):
                if t is not None:
                    t.copy_(s_)

    def timed(fn, reps):
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        total = 0.0
        for _ in range(reps):
            refill()
            ev0.record(stream)
            fn()
            ev1.record(stream)
            torch.cuda.synchronize()
            total += ev0.elapsed_time(ev1)
        return total / reps

    def plain():
        for it in items:
            api.postfilter_frame(ctx, it["df"], it["rec"], it["ref"], it["disp"], stream)

    for _ in range(warmup):
        refill()
        plain()
    ms_plain = timed(plain, steps)
    plain_out = [[t.clone() if t is not None else None for t in it["ref"] + it["disp"]] for it in items]
    log("plain per-frame drop-in: %.3f ms per batch, %.1f Mpix/s" % (ms_plain, luma_px / ms_plain / 1e3))

    batch = api.Batch(ctx, [(it["df"], it["rec"], it["ref"], it["disp"]) for it in items])
    t0 = time.perf_counter()
    batch.autotune(force=True)
    tune_s = time.perf_counter() - t0
    tiles = batch.get_tiles()
    log("autotuned in %.1f s:" % tune_s, tiles)
    for it in items:
        for t in it["ref"] + it["disp"]:
            if t is not None:
                t.zero_()
    for _ in range(warmup):
        refill()
        batch.run(stream)
    ms_batch = timed(lambda: batch.run(stream), steps)
    log("batched graph launch: %.3f ms per batch, %.1f Mpix/s" % (ms_batch, luma_px / ms_batch / 1e3))
    same = True
    for it, want in zip(items, plain_out):
        for t, w in zip(it["ref"] + it["disp"], want):
            if t is not None:
                same &= bool(torch.equal(t, w))
    del plain_out
    log("batched outputs equal the per-frame outputs:", same)
    bitexact, checked = None, 0
    if not args.no_check:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        from concurrent.futures import ThreadPoolExecutor
        order = range(len(items)) if check_all else \
            [min(range(len(items)), key=lambda i: items[i]["f"].luma_pixels())]

        def check(i):
            it = items[i]
            f = it["f"]
            deb, ref, disp = pyoracle.postfilter_frame(f, threads=4)
            got_ref, got_disp = api.to_host(it["ref"]), api.to_host(it["disp"])
            ok = True
            for p in range(f.num_planes):
                r, c = f.ref_area(p)
                ok &= bool(np.array_equal(got_ref[p][:r, :c], ref[p][:r, :c]))
                ok &= bool(np.array_equal(got_disp[p], disp[p]))
            return ok
        workers = max(1, host_threads() // 4)
        with ThreadPoolExecutor(max_workers=workers) as ex:
            oks = list(ex.map(check, order))
        bitexact, checked = all(oks), len(oks)
        log("bit-exact vs oracle (%d frames):" % checked, bitexact)
    res = {
        "workload": "%d streams x %d frames, sizes %s, half 8-bit (u8 planes) / half 10-bit, 4:2:0 + four "
                    "4:4:4 streams, qindex U[48,200], all filters + film grain" %
                    (args.streams, args.stream_frames, "/".join("%dx%d" % s for s in workload.BATCH_SIZES)),
        "frames": len(items),
        "mpix_per_batch": round(luma_px / 1e6, 2),
        "steps": steps,
        "plain_ms": round(ms_plain, 4),
        "plain_mpix_s": round(luma_px / ms_plain / 1e3, 1),
        "batched_ms": round(ms_batch, 4),
        "batched_mpix_s": round(luma_px / ms_batch / 1e3, 1),
        "batched_over_plain": round(ms_plain / ms_batch, 3),
        "batched_equals_plain": same,
        "bitexact_vs_oracle": bitexact,
        "frames_checked_vs_oracle": checked,
        "autotune_s": round(tune_s, 2),
        "tiles": {k: {"shape": v[0], "workgroups_per_cu": v[1], "chunk": v[2], "us": round(v[3], 1)}
                  for k, v in tiles.items()},
    }
    batch.close()
    ctx.close()
    del items
    torch.cuda.empty_cache()
    return res


def bench_batch(args):
    """--workload batch32: the batched path's own JSON line."""
    r = run_batch(args, args.steps, args.warmup, check_all=True)
    out = {
        "metric": "batched post-filter Mpix/s, 32-stream mixed workload (one graph launch vs per-frame drop-in)",
        "value": r["batched_mpix_s"],
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["batched_ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8/u16",
        "data": "synthetic",
        "config": r,
    }
    print(json.dumps(out), flush=True)


def bench_lrdiff(args):
    """The encoder's restoration search on the GPU: 1080p 10-bit frames at
    qindex 160 (deblocked + CDEF'd reconstruction), 64-sample units; the
    relaxed forward + adjoint kernels, the fits and the exact judge."""
    import numpy as np
    import torch
    from rvpf import api, lrdiff, workload
    torch.cuda.set_device(0)
    nf = args.lr_frames
    log("generating %d frames 1920x1080 10-bit" % nf)
    frames = [workload.Frame(1920, 1080, bit_depth=10, qindex=160, sb_size=128, seed=args.seed + i, frame_index=i,
                             pan_px=1.5, lr=2, lr_unit_size=64) for i in range(nf)]
    ctx = api.Context()
    stream = torch.cuda.current_stream()
    entries = []
    for f in frames:
        df = api.DeviceFrame(f)
        rec = df.upload_planes(f.rec)
        api.deblock(ctx, df, rec, stream)
        cd = df.alloc_like_recon()
        api.cdef(ctx, df, rec, cd, stream)
        entries.append((df, rec, cd, df.upload_planes(f.src)))
    search = lrdiff.Search(ctx, entries, stream=stream)
    torch.cuda.synchronize()
    units = search.U
    luma_units = int((search.units[:, 0] == 0).sum())
    px = sum(f.luma_pixels() for f in frames)
    lam = lrdiff.rd_lambda(10, 160)

    def timed(fn, reps):
        ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        fn()
        torch.cuda.synchronize()
        ev0.record(stream)
        for _ in range(reps):
            fn()
        ev1.record(stream)
        torch.cuda.synchronize()
        return ev0.elapsed_time(ev1) / reps

    taps = torch.full((search.U, 2, 3), 0.02, device="cuda") * search.tap_mask
    ms_fb = timed(lambda: search.wiener_value_grad(taps), args.steps)
    log("relaxed Wiener forward+adjoint: %.3f ms for %d frames (%.1f Mpix/s)" % (ms_fb, nf, px / ms_fb / 1e3))
    ms_search = timed(lambda: search.search(lam), max(1, args.steps // 5))
    log("full search: %.3f ms for %d frames (%.2f ms/frame)" % (ms_search, nf, ms_search / nf))
    # checks: finite differences (all frames) and the reference's least squares (frame 0)
    t0 = ((torch.rand((search.U, 2, 3), generator=torch.Generator().manual_seed(7)) - 0.5) * 0.3).double().cuda()
    luma = search.plane_of_unit == 0
    _, g = search.wiener_value_grad(t0)
    worst = 0.0
    for d in range(2):
        for k in range(3):
            tp, tm = t0.clone(), t0.clone()
            tp[:, d, k] += 1e-3
            tm[:, d, k] -= 1e-3
            fd = (search.wiener_value_grad(tp)[0] - search.wiener_value_grad(tm)[0]) / 2e-3
            scale = torch.maximum(fd.abs(), 1e-2 * g.view(-1, 6).abs().amax(dim=1))
            ok = luma & (scale > 0)
            worst = max(worst, ((g[:, d, k] - fd).abs()[ok] / scale[ok]).max().item())
    log("adjoint vs central differences (h=1e-3): max rel err %.2e" % worst)
    f0 = frames[0]
    df0, rec0, cd0, _ = entries[0]
    ls = None
    if not args.no_check:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        host = pyoracle.LrSearchFrame(f0, api.to_host(rec0), api.to_host(cd0), f0.src)
        taps3 = search.fit_wiener(3).cpu()
        q3 = search.quantise_wiener(taps3)
        taps3 = taps3.numpy()
        u0, u1 = search.frame_units[0]
        sample = list(range(u0, u1, max(1, (u1 - u0) // 64)))
        dmax, same = 0.0, 0
        for u in sample:
            p, ur, uc = (int(x) for x in search.units[u])
            qr, fr = host.fit_wiener(p, ur, uc)
            dmax = max(dmax, float(np.abs(taps3[u] - fr).max()))
            same += int(np.array_equal(q3[u], qr))
        ls = {"units": len(sample), "max_abs_diff_fraction": dmax, "quantised_equal": same}
        log("least-squares taps vs the oracle's fit_wiener on %d units: max |diff| %.2e, %d quantised equal"
            % (len(sample), dmax, same))
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        host = pyoracle.LrSearchFrame(f0, api.to_host(rec0), api.to_host(cd0), f0.src)
        t_start, n = time.perf_counter(), 0
        u1 = search.frame_units[0][1]
        while time.perf_counter() - t_start < args.cpu_seconds and n < u1:
            p, ur, uc = (int(x) for x in search.units[n])
            host.search_unit(p, ur, uc, lam, lrdiff.SGR_SETS_NORMAL)
            n += 1
        el = time.perf_counter() - t_start
        cpu = {"value": round(n / el, 2), "unit": "units/s", "cores": 1, "kind": "port",
               "sample": "%d units of a 1080p 10-bit frame (%.1f s), oracle search (fit_wiener + 6 SGR sets + "
                         "exact judge), 1 thread" % (n, el)}
        log("cpu:", cpu)
    out = {
        "metric": "encoder LR search units/s (relaxed forward+adjoint fits + exact judge), 1080p 10-bit",
        "value": round(units / (ms_search / 1e3), 1),
        "unit": "units/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": 1,
        "ms_per_step": round(ms_search, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (relaxed fits), u16 (judge)",
        "data": "synthetic",
        "config": {
            "workload": "%d frames 1920x1080 10-bit 4:2:0, qindex 160, deblocked+CDEF'd reconstruction, "
                        "64-sample units (%d luma units/frame), Wiener + SGR sets %s" %
                        (nf, luma_units // nf, list(lrdiff.SGR_SETS_NORMAL)),
            "units_per_step": units,
            "forward_adjoint_ms": round(ms_fb, 4),
            "forward_adjoint_mpix_s": round(px / ms_fb / 1e3, 1),
            "search_ms_per_frame": round(ms_search / nf, 3),
            "fd_max_rel_err": worst,
            "least_squares_check": ls,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def bench_grain(args):
    """The film grain output stage alone (rvpf_grain: template build, then
    the noise blend into the display planes, both on the caller's stream)
    on the GOP's 4K frames; a kernel trace of this run gives the template
    build's own duration, with nothing else on the GPU."""
    import torch
    from rvpf import api
    torch.cuda.set_device(0)
    frames = make_frames(args.frames, args.width, args.height, args.bit_depth, args.seed, chroma=args.chroma)
    ctx = api.Context()
    stream = torch.cuda.current_stream()
    items = []
    for f in frames:
        df = api.DeviceFrame(f)
        items.append((df, df.upload_planes(f.rec), df.alloc_display()))
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        for df, ref, disp in items:
            api.grain(ctx, df, ref, disp, stream)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(args.steps):
        for df, ref, disp in items:
            api.grain(ctx, df, ref, disp, stream)
    e1.record(stream)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / args.steps
    mpix = sum(f.luma_pixels() for f in frames) / 1e6
    ctx.close()
    print(json.dumps({"metric": "film grain output stage Mpix/s (template build + noise blend, serial)",
                      "value": round(mpix / (ms / 1e3), 1), "unit": "Mpix/s", "n_gpus": 1, "steps": args.steps,
                      "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True,
                      "dtype": "u16", "data": "synthetic",
                      "config": {"workload": "%dx%d %d-bit %s, %d frames, film grain lag 3" % (
                          args.width, args.height, args.bit_depth, CHROMA_NAME[args.chroma], args.frames)}}))


def bench_encsearch(args):
    """The encoder's loop filter level search and CDEF strength search on
    the GPU (search_lf + cdef::search, normal effort) on 4K 10-bit frames:
    the searches on each frame's first-pass reconstruction, wall clock per
    frame (the searches return host results, so each call waits)."""
    import numpy as np
    import torch
    from rvpf import api, workload
    torch.cuda.set_device(0)
    nf = args.frames
    w, h, bd = args.width, args.height, args.bit_depth
    log("generating %d frames %dx%d %d-bit" % (nf, w, h, bd))
    frames = [workload.Frame(w, h, bit_depth=bd, ssx=1, ssy=1, sb_size=128, qindex=QINDEX[i % 4], seed=args.seed,
                             frame_index=i, pan_px=1.5, lr=0, lf_extras=True) for i in range(nf)]
    ctx = api.Context()
    stream = torch.cuda.current_stream()
    items = []
    for f in frames:
        df = api.DeviceFrame(f)
        rec = df.upload_planes(f.rec)
        src = df.upload_planes(f.src)
        deb = [None if t is None else t.clone() for t in rec]
        api.deblock(ctx, df, deb, stream)
        items.append((f, df, rec, deb, src))
    torch.cuda.synchronize()
    lams = {q: 0.1 * (ac_q(bd, q) / 8.0) ** 2 for q in set(QINDEX)}

    def run_all():
        out = []
        for f, df, rec, deb, src in items:
            lv = api.lf_search(ctx, df, rec, src, stream)
            res, table = api.cdef_search(ctx, df, deb, src, f.cfg.qindex, lams[f.cfg.qindex], False, stream)
            out.append((lv, res, table))
        return out

    for _ in range(args.warmup):
        run_all()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        results = run_all()
    torch.cuda.synchronize()
    ms_step = (time.perf_counter() - t0) * 1e3 / args.steps
    px = sum(f.luma_pixels() for f in frames)
    log("searches: %.3f ms per step (%d frames), %.2f ms/frame" % (ms_step, nf, ms_step / nf))
    # per-search split on frame 0
    f, df, rec, deb, src = items[0]
    t1 = time.perf_counter()
    for _ in range(args.steps):
        api.lf_search(ctx, df, rec, src, stream)
    t2 = time.perf_counter()
    for _ in range(args.steps):
        api.cdef_search(ctx, df, deb, src, f.cfg.qindex, lams[f.cfg.qindex], False, stream)
    t3 = time.perf_counter()
    lf_ms, cdef_ms = (t2 - t1) * 1e3 / args.steps, (t3 - t2) * 1e3 / args.steps
    log("frame 0: lf search %.3f ms, cdef search %.3f ms" % (lf_ms, cdef_ms))
    check = None
    cpu = None
    if not args.no_check or not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        threads = host_threads()
        host_deb = api.to_host(deb)
        tc = time.perf_counter()
        want_lv, _ = pyoracle.lf_search(f, [None if a is None else a.copy() for a in f.rec], f.src, threads=threads)
        want_res, want_table, _, _ = pyoracle.cdef_search(f, host_deb, f.src, f.cfg.qindex, lams[f.cfg.qindex],
                                                          False, threads=threads)
        el = time.perf_counter() - tc
        lv, res, table = results[0]
        g = f.geom
        n = ((g.mi_rows + 15) // 16) * g.cdef_stride
        check = bool(lv == want_lv and res.as_tuple() == want_res.as_tuple()
                     and np.array_equal(table[:n], want_table[:n]))
        log("equal to the oracle's choices (frame 0):", check)
        if not args.no_cpu:
            cpu = {"value": round(f.luma_pixels() / el / 1e6, 3), "unit": "Mpix/s", "cores": threads, "kind": "port",
                   "sample": "frame 0 (%dx%d %d-bit): oracle search_lf + cdef::search (%.1f s), %d threads"
                             % (w, h, bd, el, threads)}
            log("cpu:", cpu)
    out = {
        "metric": "encoder filter search Mpix/s (search_lf + cdef::search, normal effort), %dx%d %d-bit" % (w, h, bd),
        "value": round(px / (ms_step / 1e3) / 1e6, 1),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16",
        "data": "synthetic",
        "config": {
            "workload": "%d frames %dx%d %d-bit 4:2:0 (qindex %s): loop filter level search (6 scales) + CDEF "
                        "strength search (28 luma x 15 chroma candidates, greedy <= 8 pairs), wall clock"
                        % (nf, w, h, bd, "/".join(str(q) for q in QINDEX[:nf])),
            "frames_per_step": nf,
            "lf_search_ms_frame0": round(lf_ms, 3),
            "cdef_search_ms_frame0": round(cdef_ms, 3),
            "equal_to_oracle_frame0": check,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


def ac_q(bd, q):
    """AC_QLOOKUP[(bd - 8) >> 1][q] through the library (rvpf_ac_q)."""
    from rvpf import api
    return api.lib().rvpf_ac_q(bd, q)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=["gop4k", "batch32", "lrdiff", "encsearch", "grain"], default="gop4k",
                    help="gop4k: the per-frame chain on the 4K 10-bit GOP (default); "
                         "batch32: the batched path against the per-frame drop-in on 32 mixed streams; "
                         "lrdiff: the encoder's restoration search on 1080p 10-bit frames; "
                         "encsearch: the encoder's loop filter and CDEF searches on the GOP's frames; "
                         "grain: the film grain output stage alone on the GOP's frames")
    ap.add_argument("--lr-frames", type=int, default=16)
    ap.add_argument("--streams", type=int, default=32)
    ap.add_argument("--stream-frames", type=int, default=4)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--frames", type=int, default=4)
    ap.add_argument("--width", type=int, default=3840)
    ap.add_argument("--height", type=int, default=2160)
    ap.add_argument("--bit-depth", type=int, default=10)
    ap.add_argument("--chroma", choices=sorted(CHROMA), default="420",
                    help="chroma layout of the gop4k frames (4:2:0 by default)")
    ap.add_argument("--superres-denom", type=int, default=0,
                    help="code the frames at width*8/d and upscale (super-resolution), d in 9..16")
    ap.add_argument("--seed", type=int, default=2160)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-stages", action="store_true")
    ap.add_argument("--no-check", action="store_true")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the 32-stream batched-vs-per-frame leg of the default run")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="first run a plain device copy of this many MiB (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the post-filter chain shards by stream/frame with no exchange; "
              "multi-GPU runs are not part of this build (use --gpus 1)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch32":
        bench_batch(args)
        return
    if args.workload == "lrdiff":
        bench_lrdiff(args)
        return
    if args.workload == "encsearch":
        bench_encsearch(args)
        return
    if args.workload == "grain":
        bench_grain(args)
        return

    import numpy as np
    import torch
    from rvpf import api

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    if args.calib_copy_mib:
        # a plain copy of known size: reads and writes exactly n bytes each
        n = args.calib_copy_mib << 20
        a_ = torch.ones(n // 2, dtype=torch.int16, device=dev)
        b_ = torch.empty_like(a_)
        torch.cuda.synchronize()
        b_.copy_(a_)
        torch.cuda.synchronize()
        del a_, b_
    log("generating %d frames %dx%d %d-bit" % (args.frames, args.width, args.height, args.bit_depth))
    frames = make_frames(args.frames, args.width, args.height, args.bit_depth, args.seed, args.superres_denom,
                         args.chroma)
    ctx = api.Context()
    dfs = [api.DeviceFrame(f) for f in frames]
    nsteps = args.steps + args.warmup
    # fresh reconstructions for every step (deblocking works in place)
    pristine = [df.upload_planes(f.rec) for df, f in zip(dfs, frames)]
    recs = [[[None if t is None else t.clone() for t in pristine[i]] for i in range(len(frames))]
            for _ in range(nsteps)]
    refs = [df.alloc_ref() for df in dfs]
    disps = [df.alloc_display() for df in dfs]
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()

    def step(s):
        for i, df in enumerate(dfs):
            api.postfilter_frame(ctx, df, recs[s][i], refs[i], disps[i], stream)

    log("warmup %d steps" % args.warmup)
    for s in range(args.warmup):
        step(s)
    torch.cuda.synchronize()
    log("timing %d steps" % args.steps)
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ev0.record(stream)
    for s in range(args.warmup, nsteps):
        step(s)
    ev1.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    gpu_ms = ev0.elapsed_time(ev1)
    ms_per_step = gpu_ms / args.steps
    luma_px = sum(f.luma_pixels() for f in frames)
    mpix = luma_px * args.steps / (gpu_ms / 1e3) / 1e6
    log("chain: %.3f ms/step (%d frames), %.1f Mpix/s, wall %.3f s" % (ms_per_step, len(frames), mpix, wall))

    bitexact = None
    if not args.no_check:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import pyoracle
        f = frames[0]
        deb, ref, disp = pyoracle.postfilter_frame(f, threads=host_threads())
        got_ref, got_disp = api.to_host(refs[0]), api.to_host(disps[0])
        bitexact = True
        for p in range(f.num_planes):
            r, c = f.ref_area(p)
            bitexact &= bool(np.array_equal(got_ref[p][:r, :c], ref[p][:r, :c]))
            bitexact &= bool(np.array_equal(got_disp[p], disp[p]))
        log("bit-exact vs oracle (frame 0):", bitexact)

    # Per-stage kernel time from the library's profiling mode: HIP timing
    # events bracket each stage's kernels on the stream they run on, over
    # a second pass of the same timed chain.
    b = 1 if args.bit_depth == 8 else 2
    stages = {}
    if not args.no_stages:
        spf = samples_per_frame(frames[0])
        cpf = coded_samples_per_frame(frames[0])
        # bytes per frame: deblock and CDEF run on the coded frame; the
        # upscaling reads coded and writes upscaled samples twice (CDEF
        # output and deblocked frame, every plane being restored here)
        alg = {"deblock": 2 * b * cpf, "cdef": 2 * b * cpf, "lr": (2 * b + b * 4 / 64) * spf,
               "grain_apply": 2 * b * spf, "upscale": 2 * b * (cpf + spf)}
        for s_ in range(nsteps):
            for i in range(len(frames)):
                for t, src in zip(recs[s_][i], pristine[i]):
                    if t is not None:
                        t.copy_(src)
        torch.cuda.synchronize()
        ctx.read_profile()
        ctx.set_profiling(True)
        for s_ in range(args.warmup, nsteps):
            step(s_)
        torch.cuda.synchronize()
        prof = ctx.read_profile()
        ctx.set_profiling(False)
        for k, (ms, n) in prof.items():
            if n == 0:
                continue
            avg = ms / n
            st = {"ms": round(avg, 4), "launches": n}
            if k in alg:
                gbs = alg[k] / (avg / 1e3) / 1e9
                st.update({"alg_bytes": int(alg[k]), "GB_s": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)})
                log("stage %-13s %.4f ms  %.1f GB/s algorithmic (%.1f%% of HBM peak)" % (k, avg, gbs, 100 * gbs / HBM_PEAK_GBS))
            else:
                log("stage %-13s %.4f ms (side stream, overlapped)" % (k, avg))
            stages[k] = st

    roofline = None
    # the roofline names one kernel: that of the longest main-stream stage
    main = [k for k in stages if "GB_s" in stages[k] and len(STAGE_KERNELS.get(k, [k])) == 1]
    if main:
        dom = max(main, key=lambda k: stages[k]["ms"])
        s = stages[dom]
        default_workload = (args.width, args.height, args.bit_depth, args.frames, args.superres_denom,
                            args.chroma) == (3840, 2160, 10, 4, 0, "420")
        roofline = {"bound": "hbm", "achieved": s["GB_s"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": s["frac"], "traffic": measured_traffic(dom) if default_workload else None,
                    "kernel": dom, "alg_bytes_per_launch": s["alg_bytes"], "ms_per_launch": s["ms"],
                    "traffic_source": PMC_SUMMARY if default_workload else None}
    spf, cpf = samples_per_frame(frames[0]), coded_samples_per_frame(frames[0])
    chain_alg = (4 * b * cpf + (4 * b + b * 4 / 64) * spf + (2 * b * (cpf + spf) if cpf != spf else 0)) * len(frames)
    chain_gbs = chain_alg / (ms_per_step / 1e3) / 1e9

    cpu = None
    if not args.no_cpu:
        log("cpu baseline (%.0f s budget per thread count)" % args.cpu_seconds)
        cpu = cpu_baseline(frames, args.cpu_seconds)
        log("cpu:", cpu)

    batch32 = None
    if not args.no_batch:
        # the batched path against the per-frame drop-in on the 32-stream workload
        del recs, refs, disps, pristine
        torch.cuda.empty_cache()
        batch32 = run_batch(args, steps=min(args.steps, 5), warmup=min(args.warmup, 2), check_all=True)

    out = {
        "metric": "post-filter Mpix/s (deblock+CDEF+LR+grain) 4K 10-bit, bit-exact; % HBM roofline",
        "value": round(mpix, 1),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u16" if b == 2 else "u8",
        "data": "synthetic",
        "config": {
            "workload": "%dx%d %d-bit %s, %d-frame GOP (qindex %s), deblock+CDEF+%sLR(Wiener/SGR, 64-px units)"
                        "+film grain, per-frame drop-in" % (
                            args.width, args.height, args.bit_depth, CHROMA_NAME[args.chroma], len(frames),
                            "/".join(str(q) for q in QINDEX[:len(frames)]),
                            "super-resolution from width %d (denominator %d)+" % (
                                frames[0].params.frame_width, args.superres_denom) if args.superres_denom else ""),
            "frames_per_step": len(frames),
            "bitexact_frame0": bitexact,
            "chain_alg_GB_s": round(chain_gbs, 1),
            "chain_hbm_frac": round(chain_gbs / HBM_PEAK_GBS, 4),
            "stages": stages,
            "batch32": batch32,
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
