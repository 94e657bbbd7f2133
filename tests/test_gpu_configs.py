"""GPU parity at the measurement configurations' own sizes.

Every frame of each workload goes through the HIP library (C ABI) and is
compared with the oracle sample for sample (reference frame, display
frame; the batched path's in-place deblocked frame; the per-frame chain
leaves the reconstruction unmodified), as the reference decoder runs its
chain per frame (src/decoder/mod.rs:747-917):

- 352x288 8-bit 4:2:0 key frame: SB 64, qindex 128, deblock + CDEF (8
  strengths) + Wiener restoration on luma (64-sample units);
- 1920x1080 10-bit 4:2:0 GOP of 8 frames: SB 128, qindex 96, Wiener and
  self-guided mixed per unit, film grain (lag 3, 8 luma points, a seed per
  frame), 1.5 px/frame pan;
- 3840x2160 10-bit 4:2:0 and 4:4:4, frames at qindex 64, 96, 128, 160 and
  200, all filters; and a 4K 12-bit monochrome GOP with film grain;
- the 32-stream mixed batch (4 frames each) through one graph launch
  (rvpf_batch_run, autotuned tiles), every frame;
- the encoder restoration search on 16 frames of 1080p 10-bit at qindex
  160: adjoint against central differences (h = 1e-3, relative error below
  1e-3) over every unit, the least-squares Wiener taps and the exact judge
  against the oracle over every unit of a frame, and the unit choices
  against the oracle's search (src/encoder/lr.rs:64-430).
"""
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import common
import pyoracle
from rvpf import api, lrdiff, workload

pytestmark = pytest.mark.gpu

BENCH_SEED = 2160


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def ctx():
    c = api.Context()
    yield c
    c.close()


def _gop(w, h, bd, n, qindex, seed, **kw):
    q = qindex if isinstance(qindex, (list, tuple)) else [qindex] * n
    return [dict(width=w, height=h, bit_depth=bd, qindex=q[i], seed=seed, frame_index=i, pan_px=1.5,
                 with_source=False, **kw) for i in range(n)]


def _run_chain_and_check(ctx, cfgs, name, bps=2):
    """Per-frame drop-in over the frames in order on one stream, then every
    frame against the oracle (computed on a pool of threads)."""
    import torch
    _say(name, ": generating", len(cfgs), "frames")
    frames = workload.make_frames(cfgs)
    jobs = []
    for f in frames:
        df = api.DeviceFrame(f, bytes_per_sample=bps)
        rec = df.upload_planes(f.rec)
        jobs.append((f, df, rec, df.alloc_ref(), df.alloc_display()))
    _say(name, ": chain")
    for f, df, rec, ref, disp in jobs:
        api.postfilter_frame(ctx, df, rec, ref, disp)
    ctx.sync()
    torch.cuda.synchronize()
    _say(name, ": oracle")
    with ThreadPoolExecutor(max_workers=4) as ex:
        wants = list(ex.map(lambda f: pyoracle.postfilter_frame(f, threads=4), frames))
    for i, ((f, df, rec, ref, disp), (deb_w, ref_w, disp_w)) in enumerate(zip(jobs, wants)):
        what = "%s frame %d (qindex %d)" % (name, i, f.cfg.qindex)
        _say("check", what)
        common.assert_recon_kept(api.to_host(rec), f, what)
        common.assert_planes(api.to_host(ref), ref_w, f, "ref", what + " reference")
        common.assert_planes(api.to_host(disp), disp_w, f, "crop", what + " display")


def test_c1_352x288_8bit_key_frame(ctx):
    cfg = dict(width=352, height=288, bit_depth=8, sb_size=64, qindex=128, seed=BENCH_SEED, lf=True, cdef=True,
               cdef_bits=3, lr=1, lr_unit_size=64, with_source=False)
    _run_chain_and_check(ctx, [cfg], "352x288 8-bit")
    _run_chain_and_check(ctx, [cfg], "352x288 8-bit u8 planes", bps=1)


def test_c2_1080p_10bit_gop8(ctx):
    cfgs = _gop(1920, 1080, 10, 8, 96, BENCH_SEED, sb_size=128, lr=2, lr_unit_size=64, grain=True, grain_lag=3,
                grain_y_points=8)
    _run_chain_and_check(ctx, cfgs, "1080p 10-bit GOP")


QSWEEP = [64, 96, 128, 160, 200]


@pytest.mark.parametrize("chroma", ["420", "444"])
def test_c3_4k_10bit_qindex_sweep(ctx, chroma):
    ss = dict(ssx=1, ssy=1) if chroma == "420" else dict(ssx=0, ssy=0)
    cfgs = _gop(3840, 2160, 10, len(QSWEEP), QSWEEP, BENCH_SEED, sb_size=128, lr=2, lr_unit_size=64, grain=True,
                grain_lag=3, grain_y_points=8, **ss)
    _run_chain_and_check(ctx, cfgs, "4K 10-bit " + chroma)


def test_c3_4k_12bit_mono_grain(ctx):
    cfgs = _gop(3840, 2160, 12, 4, [64, 96, 160, 200], BENCH_SEED + 1, mono=True, sb_size=128, lr=2,
                lr_unit_size=64, grain=True, grain_lag=3, grain_y_points=8)
    _run_chain_and_check(ctx, cfgs, "4K 12-bit mono")


def test_c4_batch32_every_frame(ctx):
    """The 32-stream mixed workload of the batched path (4 frames per
    stream, 8-bit streams in 8-bit sample planes): one autotuned graph
    launch, every frame against the oracle, and equal to 128 plain per-frame
    calls on the same inputs."""
    import torch
    cfgs = workload.batch_stream_configs(32, 4, BENCH_SEED)
    flat = [(bd, kw) for bd, kws in cfgs for kw in kws]
    _say("batch32: generating", len(flat), "frames")
    frames = workload.make_frames([kw for _, kw in flat])
    items = []
    for (bd, _), f in zip(flat, frames):
        df = api.DeviceFrame(f, bytes_per_sample=1 if bd == 8 else 2)
        pristine = df.upload_planes(f.rec)
        items.append(dict(f=f, df=df, pristine=pristine, rec=[None if t is None else t.clone() for t in pristine],
                          ref=df.alloc_like_recon(), disp=df.alloc_display()))
    _say("batch32:", sum(f.luma_pixels() for f in frames) / 1e6, "Mpix")

    def refill():
        for it in items:
            for t, s in zip(it["rec"], it["pristine"]):
                if t is not None:
                    t.copy_(s)
            for t in it["ref"] + it["disp"]:
                if t is not None:
                    t.zero_()

    batch = api.Batch(ctx, [(it["df"], it["rec"], it["ref"], it["disp"]) for it in items])
    try:
        _say("batch32: autotune")
        batch.autotune(force=True)
        refill()
        _say("batch32: run")
        batch.run()
        ctx.sync()
        torch.cuda.synchronize()
        got = [(api.to_host(it["rec"]), api.to_host(it["ref"]), api.to_host(it["disp"])) for it in items]
        batched_dev = [[t.clone() for t in it["ref"] + it["disp"] if t is not None] for it in items]
    finally:
        batch.close()
    _say("batch32: plain per-frame calls")
    refill()
    for it in items:
        api.postfilter_frame(ctx, it["df"], it["rec"], it["ref"], it["disp"])
    ctx.sync()
    torch.cuda.synchronize()
    for k, it in enumerate(items):
        plain = [t for t in it["ref"] + it["disp"] if t is not None]
        assert all(torch.equal(a, b) for a, b in zip(plain, batched_dev[k])), "frame %d: batched != plain" % k
    _say("batch32: oracle")

    def check(k):
        it = items[k]
        f = it["f"]
        deb_w, ref_w, disp_w = pyoracle.postfilter_frame(f, threads=4)
        rec, ref, disp = got[k]
        what = "batch32 frame %d (%dx%d %d-bit bps %d)" % (k, f.params.frame_width, f.params.frame_height,
                                                          f.bit_depth, it["df"].bps)
        common.assert_planes(rec, deb_w, f, "full", what + " deblocked")
        common.assert_planes(ref, ref_w, f, "mi", what + " reference")
        common.assert_planes(disp, disp_w, f, "crop", what + " display")
        return k

    with ThreadPoolExecutor(max_workers=4) as ex:
        for k in ex.map(check, range(len(items))):
            if k % 16 == 15:
                _say("checked", k + 1, "frames")


def test_c5_lr_search_16x1080p(ctx):
    import torch
    nf = 16
    cfgs = [dict(width=1920, height=1080, bit_depth=10, qindex=160, sb_size=128, seed=BENCH_SEED + i, frame_index=i,
                 pan_px=1.5, lr=2, lr_unit_size=64) for i in range(nf)]
    _say("lr search: generating", nf, "frames")
    frames = workload.make_frames(cfgs)
    entries = []
    for f in frames:
        df = api.DeviceFrame(f)
        rec = df.upload_planes(f.rec)
        api.deblock(ctx, df, rec)
        cd = df.alloc_like_recon()
        api.cdef(ctx, df, rec, cd)
        entries.append((df, rec, cd, df.upload_planes(f.src)))
    s = lrdiff.Search(ctx, entries)
    torch.cuda.synchronize()
    luma_units = int((s.units[:, 0] == 0).sum())
    assert luma_units == nf * 510, luma_units  # 30 x 17 units of 64 per 1080p frame
    # adjoint vs central differences over every unit of every frame
    _say("lr search: finite differences over", s.U, "units")
    t0 = ((torch.rand((s.U, 2, 3), generator=torch.Generator().manual_seed(7)) - 0.5) * 0.3).double().cuda()
    t0 = t0 * s.tap_mask.double()
    _, g = s.wiener_value_grad(t0)
    free = s.tap_mask > 0
    worst = 0.0
    for d in range(2):
        for k in range(3):
            tp, tm = t0.clone(), t0.clone()
            tp[:, d, k] += 1e-3
            tm[:, d, k] -= 1e-3
            fd = (s.wiener_value_grad(tp)[0] - s.wiener_value_grad(tm)[0]) / 2e-3
            scale = torch.maximum(fd.abs(), 1e-2 * g.view(-1, 6).abs().amax(dim=1))
            ok = free[:, d, k] & (scale > 0)
            worst = max(worst, ((g[:, d, k] - fd).abs()[ok] / scale[ok]).max().item())
    _say("lr search: worst relative error", worst)
    assert worst < 1e-3, worst
    # least-squares taps, exact judge and search choice against the oracle on every unit of frame 0
    f0 = frames[0]
    df0, rec0, cd0, _ = entries[0]
    host = pyoracle.LrSearchFrame(f0, api.to_host(rec0), api.to_host(cd0), f0.src)
    taps3 = s.fit_wiener(3)
    q3 = s.quantise_wiener(taps3)
    taps3 = taps3.cpu().numpy()
    u0, u1 = s.frame_units[0]
    units = [tuple(int(x) for x in s.units[u]) for u in range(u0, u1)]
    _say("lr search: least squares on", len(units), "units")

    def fit(u):
        p, ur, uc = units[u - u0]
        return host.fit_wiener(p, ur, uc)
    with ThreadPoolExecutor(max_workers=8) as ex:
        fits = list(ex.map(fit, range(u0, u1)))
    dmax = max(float(np.abs(taps3[u] - fits[u - u0][1]).max()) for u in range(u0, u1))
    same = sum(int(np.array_equal(q3[u], fits[u - u0][0])) for u in range(u0, u1))
    _say("lr search: max |tap - LS| %.2e, quantised equal %d of %d" % (dmax, same, u1 - u0))
    assert dmax < 2e-3, dmax
    assert same >= 0.95 * (u1 - u0), same
    recs = s.unit_records("wiener", q3)
    got = s.exact_sse(recs).cpu().numpy()

    def judge(u):
        p, ur, uc = units[u - u0]
        return host.unit_sse(p, ur, uc, recs[u])
    with ThreadPoolExecutor(max_workers=8) as ex:
        want = list(ex.map(judge, range(u0, u1)))
    assert [int(got[u]) for u in range(u0, u1)] == [int(w) for w in want]
    lam = lrdiff.rd_lambda(10, 160)
    choice, cost, _ = s.search(lam)
    cost = cost.cpu().numpy()

    def ref_search(u):
        p, ur, uc = units[u - u0]
        return host.search_unit(p, ur, uc, lam, lrdiff.SGR_SETS_NORMAL)
    _say("lr search: oracle search on", u1 - u0, "units")
    with ThreadPoolExecutor(max_workers=8) as ex:
        refs = list(ex.map(ref_search, range(u0, u1)))
    same = sum(int(np.array_equal(refs[u - u0][0][:10], choice[u, :10])) for u in range(u0, u1))
    total_ref = sum(r[1] for r in refs)
    total = float(cost[u0:u1].sum())
    _say("lr search: same choice on %d of %d units, cost %.6g vs oracle %.6g" % (same, u1 - u0, total, total_ref))
    assert same >= 0.9 * (u1 - u0)
    assert total <= total_ref * 1.001
