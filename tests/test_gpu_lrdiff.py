"""GPU tests of the differentiable restoration search (rvpf_lrd_* and
rvpf.lrdiff), against the oracle's restatement of the reference encoder
(src/encoder/lr.rs:64-430) and against finite differences.

- The adjoint gradients of the relaxed Wiener and self-guided errors agree
  with central finite differences of the forward kernels (h = 1e-3,
  relative error < 1e-3).
- The relaxed forward equals a float64 torch evaluation of the same model.
- Three alternating exact steps reproduce the reference's least-squares
  Wiener taps (fit_wiener), and the relaxed error's vertical gradient
  vanishes there (the last step solves the vertical taps exactly).
- The fitted self-guided weights are a stationary point of the relaxed error.
- The exact judge (rvpf_lr + rvpf_lrd_rect_sse) equals the oracle's exact
  error for the same unit filters, sample for sample in the sums.
- The whole search picks the reference's choice for most units and costs no
  more in total.
"""
import numpy as np
import pytest

import common
import pyoracle
from rvpf import abi, api, lrdiff

pytestmark = pytest.mark.gpu

FRAMES = [
    dict(width=200, height=136, bit_depth=10, qindex=160, lr=2, seed=5),
    dict(width=176, height=120, bit_depth=8, qindex=96, lr=2, seed=31, ssx=1, ssy=1),
    dict(width=130, height=90, bit_depth=12, qindex=200, lr=2, seed=32, ssx=0, ssy=0, lr_unit_size=64),
]


def _say(*a):
    print(*a, flush=True)


class Case:
    def __init__(self, ctx, kw):
        import torch
        self.f = common.make_frame(**kw)
        f = self.f
        deb = pyoracle.loop_filter(f, [None if a is None else a.copy() for a in f.rec])
        cd = pyoracle.cdef(f, deb)
        self.host = pyoracle.LrSearchFrame(f, deb, cd, f.src)
        self.df = api.DeviceFrame(f)
        self.deb, self.cdef, self.src = (self.df.upload_planes(x) for x in (deb, cd, f.src))
        self.s = lrdiff.Search(ctx, self.df, self.deb, self.cdef, self.src)
        torch.cuda.synchronize()


@pytest.fixture(scope="module")
def ctx():
    c = api.Context()
    yield c
    c.close()


@pytest.fixture(scope="module")
def cases(ctx):
    return [Case(ctx, kw) for kw in FRAMES]


def _relaxed_wiener_f64(s, taps):
    """The relaxed Wiener model in float64 torch, per item."""
    import torch
    win = s.win.double()
    t = taps.double()[s.unit_of_item]
    v, h = t[:, 0, :], t[:, 1, :]
    c = win[:, :, 3:3 + 64]
    H = c.clone()
    for k in range(3):
        d = 3 - k
        H = H + h[:, k, None, None] * (win[:, :, 3 - d:3 - d + 64] + win[:, :, 3 + d:3 + d + 64] - 2 * c)
    ctr = H[:, 3:3 + 64, :]
    out = ctr.clone()
    for k in range(3):
        d = 3 - k
        out = out + v[:, k, None, None] * (H[:, 3 - d:3 - d + 64, :] + H[:, 3 + d:3 + d + 64, :] - 2 * ctr)
    it = s.items.long()
    mask = (torch.arange(64, device=win.device)[None, :, None] < it[:, 4, None, None]) & \
           (torch.arange(64, device=win.device)[None, None, :] < it[:, 3, None, None])
    e = torch.where(mask, out - s.src.double(), torch.zeros_like(out))
    return (e * e).sum(dim=(1, 2))


def test_wiener_forward_matches_f64(cases):
    import torch
    for c in cases:
        _say("forward", c.f.params.frame_width, c.f.params.frame_height)
        g = torch.Generator().manual_seed(1)
        taps = ((torch.rand((c.s.U, 2, 3), generator=g) - 0.5) * 0.3).cuda() * c.s.tap_mask
        sse, _ = c.s.wiener_forward(taps)
        want = _relaxed_wiener_f64(c.s, taps)
        rel = ((sse - want).abs() / want.clamp_min(1.0)).max().item()
        assert rel < 1e-5, rel


def _fd_check(value, x, h=1e-3, tol=1e-3):
    """Central differences of value(x) [U] along every coordinate of x [U, K]
    against the analytic gradient; returns the largest relative error."""
    import torch
    f0, g = value(x)
    K = g.shape[1]
    worst = 0.0
    for k in range(K):
        xp, xm = x.clone(), x.clone()
        xp[:, k] += h
        xm[:, k] -= h
        fp, _ = value(xp)
        fm, _ = value(xm)
        fd = (fp - fm) / (2 * h)
        scale = torch.maximum(fd.abs(), 1e-2 * g.abs().amax(dim=1))
        ok = scale > 0
        rel = ((g[:, k] - fd).abs()[ok] / scale[ok]).max().item() if ok.any() else 0.0
        worst = max(worst, rel)
    return worst


def test_wiener_gradient_finite_differences(cases):
    import torch
    for c in cases:
        _say("wiener fd", c.f.params.frame_width)
        g = torch.Generator().manual_seed(2)
        taps = ((torch.rand((c.s.U, 2, 3), generator=g) - 0.5) * 0.3).double().cuda()
        luma = c.s.plane_of_unit == 0  # chroma units have no outermost tap

        def value(x):
            sse, grad = c.s.wiener_value_grad(x.view(-1, 2, 3))
            return sse[luma], grad.view(-1, 6)[luma]
        rel = _fd_check(value, taps.view(-1, 6))
        assert rel < 1e-3, rel


def test_wiener_autograd_function(cases):
    import torch
    c = cases[0]
    taps = (torch.full((c.s.U, 2, 3), 0.05, device="cuda") * c.s.tap_mask).requires_grad_(True)
    loss = lrdiff.WienerSSE.apply(taps, c.s).sum()
    loss.backward()
    _, want = c.s.wiener_value_grad(taps.detach())
    assert torch.allclose(taps.grad.double(), want, rtol=1e-5, atol=1e-3)


def test_sgr_autograd_function(cases):
    """SgrSSE through autograd: the forward is the relaxed error per unit,
    and backward scales the adjoint gradient by the incoming gradient."""
    import torch
    c = cases[0]
    g = torch.Generator().manual_seed(11)
    weights = torch.rand(c.s.U, generator=g).cuda()
    for set_ in (0, 10, 14):
        flt = c.s.sgr_filter(set_)
        w = ((torch.rand((c.s.U, 2), generator=g) - 0.5) * 60).cuda().requires_grad_(True)
        sse = lrdiff.SgrSSE.apply(w, c.s, flt)
        want_sse, want_grad = c.s.sgr_value_grad(w.detach().double(), flt)
        assert torch.allclose(sse.double(), want_sse.double(), rtol=1e-5, atol=1e-3), set_
        (sse * weights).sum().backward()
        assert torch.allclose(w.grad.double(), want_grad.double() * weights.double()[:, None], rtol=1e-4,
                              atol=1e-2), set_


def test_sgr_gradient_finite_differences(cases):
    import torch
    for c in cases:
        for set_ in (0, 10, 14):
            _say("sgr fd", c.f.params.frame_width, "set", set_)
            flt = c.s.sgr_filter(set_)
            g = torch.Generator().manual_seed(3 + set_)
            w = ((torch.rand((c.s.U, 2), generator=g) - 0.5) * 100).double().cuda()
            r0, _, r1, _ = lrdiff.SGR_PARAMS[set_]
            cols = [k for k, r in ((0, r0), (1, r1)) if r]

            def value(x):
                full = torch.zeros((c.s.U, 2), dtype=torch.float64, device="cuda")
                full[:, cols] = x
                sse, grad = c.s.sgr_value_grad(full, flt)
                return sse, grad[:, cols]
            rel = _fd_check(value, w[:, cols].contiguous())
            assert rel < 1e-3, (set_, rel)


def test_fit_wiener_is_the_reference_least_squares(cases):
    import torch
    for c in cases:
        _say("fit wiener", c.f.params.frame_width)
        taps = c.s.fit_wiener(3)
        q = c.s.quantise_wiener(taps)
        got = taps.cpu().numpy()
        same, diffs = 0, []
        for u, (p, ur, uc) in enumerate(c.s.units):
            qr, fr = c.host.fit_wiener(int(p), int(ur), int(uc))
            diffs.append(np.abs(got[u] - fr).max())
            same += int(np.array_equal(q[u], qr))
        assert max(diffs) < 2e-3, max(diffs)
        assert same >= 0.95 * c.s.U, (same, c.s.U)
        # stationarity of the vertical taps at the least-squares solution
        _, g_fit = c.s.wiener_value_grad(taps)
        _, g_zero = c.s.wiener_value_grad(torch.zeros_like(taps))
        free = c.s.tap_mask[:, 0, :] > 0  # chroma units keep their outermost tap at 0
        ratio = (g_fit[:, 0, :].abs()[free].amax() / g_zero[:, 0, :].abs()[free].amax()).item()
        assert ratio < 1e-3, ratio


def test_more_steps_lower_the_relaxed_error(cases):
    for c in cases:
        t3 = c.s.fit_wiener(3)
        t9 = c.s.fit_wiener(6, taps=t3)
        e3, _ = c.s.wiener_value_grad(t3)
        e9, _ = c.s.wiener_value_grad(t9)
        assert (e9 <= e3 * (1 + 1e-6) + 1e-3).all()


def test_fit_sgr_is_stationary(cases):
    for c in cases:
        for set_ in lrdiff.SGR_SETS_NORMAL:
            w, flt = c.s.fit_sgr(set_)
            _, g = c.s.sgr_value_grad(w, flt)
            _, g0 = c.s.sgr_value_grad(w * 0, flt)
            ratio = (g.abs().amax() / g0.abs().amax().clamp_min(1e-30)).item()
            assert ratio < 1e-3, (set_, ratio)


def test_exact_judge_matches_oracle(cases):
    for c in cases:
        _say("judge", c.f.params.frame_width)
        rw = c.s.unit_records("wiener", c.s.quantise_wiener(c.s.fit_wiener(3)))
        w, _ = c.s.fit_sgr(3)
        rs = c.s.unit_records("sgr", c.s.quantise_sgr(w, 3), 3)
        for recs in (None, rw, rs):
            got = c.s.exact_sse(recs).cpu().numpy()
            for u, (p, ur, uc) in enumerate(c.s.units):
                want = c.host.unit_sse(int(p), int(ur), int(uc), None if recs is None else recs[u])
                assert got[u] == want, (u, got[u], want)


def test_search_matches_reference(cases):
    for c in cases:
        lam = lrdiff.rd_lambda(c.f.bit_depth, c.f.cfg.qindex)
        choice, cost, ftype = c.s.search(lam)
        cost = cost.cpu().numpy()
        same, total_ref = 0, 0.0
        for u, (p, ur, uc) in enumerate(c.s.units):
            ref, rc = c.host.search_unit(int(p), int(ur), int(uc), lam, lrdiff.SGR_SETS_NORMAL)
            total_ref += rc
            same += int(np.array_equal(ref[:10], choice[u, :10]))
        _say("search", c.f.params.frame_width, "same", same, "of", c.s.U, "cost", cost.sum(), "ref", total_ref)
        assert same >= 0.9 * c.s.U
        assert cost.sum() <= total_ref * 1.001


def test_multi_frame_search_equals_per_frame(ctx, cases):
    """One Search over several frames (one launch per kernel) chooses as the
    per-frame searches do."""
    multi = lrdiff.Search(ctx, [(c.df, c.deb, c.cdef, c.src) for c in cases])
    lams = [lrdiff.rd_lambda(c.f.bit_depth, c.f.cfg.qindex) for c in cases]
    choice, cost, ftypes = multi.search(lams)
    for k, c in enumerate(cases):
        ch1, co1, ft1 = c.s.search(lams[k])
        u0, u1 = multi.frame_units[k]
        assert np.array_equal(choice[u0:u1], ch1)
        assert ftypes[k] == ft1


ODD_FRAMES = [  # frame sizes of the reference's odd_sizes test (tests/encode.rs:119), whose encoder runs this search
    dict(width=1, height=1, bit_depth=8, qindex=160, lr=2, seed=41),
    dict(width=7, height=5, bit_depth=10, qindex=160, lr=2, seed=42, ssx=0, ssy=0),
    dict(width=16, height=200, bit_depth=12, qindex=160, lr=2, seed=43, ssx=1, ssy=0),
    dict(width=130, height=34, bit_depth=8, qindex=160, lr=2, seed=44),
]


def test_odd_sizes_exact_judge_and_search(ctx):
    """Units smaller than one 8x8 block: the exact judge of the unrestored,
    Wiener and self-guided records equals the oracle's, and the search
    chooses a valid type for every unit at no more than the unrestored
    unit's cost, SSE + lambda (the unrestored unit is one of its
    candidates)."""
    for kw in ODD_FRAMES:
        c = Case(ctx, kw)
        _say("odd judge", kw["width"], kw["height"], "units", c.s.U)
        rw = c.s.unit_records("wiener", c.s.quantise_wiener(c.s.fit_wiener(3)))
        w, _ = c.s.fit_sgr(3)
        rs = c.s.unit_records("sgr", c.s.quantise_sgr(w, 3), 3)
        for recs in (None, rw, rs):
            got = c.s.exact_sse(recs).cpu().numpy()
            for u, (p, ur, uc) in enumerate(c.s.units):
                want = c.host.unit_sse(int(p), int(ur), int(uc), None if recs is None else recs[u])
                assert got[u] == want, (kw, u, got[u], want)
        lam = lrdiff.rd_lambda(c.f.bit_depth, c.f.cfg.qindex)
        choice, cost, ftype = c.s.search(lam)
        cost = cost.cpu().numpy()
        none_sse = c.s.exact_sse(None).cpu().numpy()
        for u in range(c.s.U):
            assert int(choice[u, 0]) in (0, 1, 2)  # none, Wiener, self-guided (include/rvpf.h RVPF_RESTORE_*)
            none_cost = float(none_sse[u]) + lam  # cost of the unrestored unit (lr.rs:154-186)
            assert cost[u] <= none_cost * (1 + 1e-9), (kw, u, cost[u], none_cost)
