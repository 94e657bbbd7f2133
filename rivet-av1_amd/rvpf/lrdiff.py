"""Gradient-based loop restoration search on the GPU.

This is the encoder's restoration search (rivet-av1 src/encoder/lr.rs:64-430).
For every restoration unit it fits Wiener taps and self-guided projection
weights, then keeps, per unit, the cheapest of: no restoration, the fitted
Wiener filter, or the fitted self-guided filter of each parameter set. The
cost is the exact squared error plus a rate term.

The fitting works on float-relaxed filters (no rounding, no clipping; HIP
kernels behind the C ABI). Their squared error against the source and its
gradient with respect to each unit's taps come from a hand-written adjoint
kernel. `WienerSSE` and `SgrSSE` expose them as torch.autograd Functions. The
relaxed error is quadratic in one filter direction while the other is held
fixed. Newton steps whose Hessian columns come from gradient differences
therefore solve each direction exactly. Three steps (vertical, horizontal,
vertical, starting from zero taps) are the reference's fit_wiener
(lr.rs:240-325). More steps keep lowering the relaxed error. The fitted taps
are always quantised to the coded ranges, and the choice between filters is
made with the exact integer filter (rvpf_lr) plus the rate, as in lr.rs:153-186.
"""
import ctypes as C

import numpy as np

from . import abi, api

T = 64
WS = T + 6

# Coded ranges (src/tables.rs:541-547, 646-652; spec 5.11.58).
WIENER_TAPS_MIN = (-5, -23, -17)
WIENER_TAPS_MAX = (10, 8, 46)
SGRPROJ_XQD_MIN = (-96, -32)
SGRPROJ_XQD_MAX = (31, 95)
# SGR_PARAMS (src/tables.rs:538): r0, e0, r1, e1 per set.
SGR_PARAMS = [(2, 12, 1, 4), (2, 15, 1, 6), (2, 18, 1, 8), (2, 21, 1, 9), (2, 24, 1, 10), (2, 29, 1, 11),
              (2, 36, 1, 12), (2, 45, 1, 13), (2, 56, 1, 14), (2, 68, 1, 15), (0, 0, 1, 5), (0, 0, 1, 8),
              (0, 0, 1, 11), (0, 0, 1, 14), (2, 30, 0, 0), (2, 75, 0, 0)]
# The self-guided sets tried per effort (lr.rs:40-41).
SGR_SETS_NORMAL = (0, 3, 6, 10, 13, 14)
SGR_SETS_FAST = (3, 10)
# Rate of each choice in units of lambda (lr.rs:154-176).
RATE_NONE, RATE_WIENER_LUMA, RATE_WIENER_CHROMA, RATE_SGR = 1.0, 40.0, 28.0, 20.0


def _torch():
    import torch
    return torch


def rd_lambda(bit_depth, qindex):
    """The encoder's Lagrange multiplier (src/encoder/mod.rs:1449-1455)."""
    q = api.lib().rvpf_ac_q(bit_depth, qindex)
    if q < 0:
        raise ValueError("bit depth or qindex out of range")
    return 0.1 * (q / 8.0) ** 2


def round_half_away(x):
    """Rust's f64::round / C round: halves away from zero."""
    return np.sign(x) * np.floor(np.abs(x) + 0.5)


def unit_layout(params):
    """The units of every plane and their rectangles.

    Returns (units, items): units[u] = (plane, unit row, unit column);
    items[i] = (plane, x0, y0, w, h, stripe_start, stripe_end, unit), one
    per stripe of a unit (lr_stripe, src/decoder/postfilter.rs:640-661; the
    rectangles of lr.rs:82-127), split every 64 columns."""
    nplanes = 1 if params.mono_chrome else 3
    units, items = [], []
    stripes = (params.frame_height + 8 + 63) // 64
    for p in range(nplanes):
        sx = params.ssx if p else 0
        sy = params.ssy if p else 0
        pex = ((params.upscaled_width + sx) >> sx) - 1
        pey = ((params.frame_height + sy) >> sy) - 1
        size = params.loop_restoration_size[p]
        rows, cols = params.lr_unit_rows[p], params.lr_unit_cols[p]
        base = len(units)
        for ur in range(rows):
            for uc in range(cols):
                units.append((p, ur, uc))
        for st in range(stripes):
            ss = (-8 + 64 * st) >> sy
            se = ss + (64 >> sy) - 1
            y0, y1 = max(ss, 0), min(se, pey)
            if y0 > y1:
                continue
            luma_row = max(64 * st - 8, 0)
            ur = min(rows - 1, ((luma_row + 8) >> sy) // size)
            for uc in range(cols):
                x0 = uc * size
                x1 = pex if uc + 1 == cols else min((uc + 1) * size - 1, pex)
                if x0 > pex:
                    continue
                for xs in range(x0, x1 + 1, T):
                    items.append((p, xs, y0, min(T, x1 - xs + 1), y1 - y0 + 1, ss, se, base + ur * cols + uc))
    return np.array(units, dtype=np.int32).reshape(-1, 3), np.array(items, dtype=np.int32).reshape(-1, 8)


def _ptr(t):
    return C.c_void_p(t.data_ptr())


class WienerSSE:
    """torch.autograd.Function: per-unit relaxed Wiener squared error of
    taps [U, 2, 3] (vertical, horizontal; fractions of 128)."""
    _fn = None

    @classmethod
    def apply(cls, taps, search):
        if cls._fn is None:
            torch = _torch()

            class _F(torch.autograd.Function):
                @staticmethod
                def forward(ctx_, taps_, s_):
                    sse, resid = s_.wiener_forward(taps_)
                    ctx_.s = s_
                    ctx_.resid = resid
                    ctx_.save_for_backward(taps_)
                    return s_.per_unit(sse)

                @staticmethod
                def backward(ctx_, g):
                    (taps_,) = ctx_.saved_tensors
                    gi = ctx_.s.per_unit(ctx_.s.wiener_backward(taps_, ctx_.resid))
                    return (gi * g.to(gi.dtype)[:, None]).to(taps_.dtype).view_as(taps_), None

            cls._fn = _F
        return cls._fn.apply(taps, search)


class SgrSSE:
    """torch.autograd.Function: per-unit relaxed self-guided squared error
    of weights w [U, 2] = (w0, w2) (units of 1/128), given the filter
    outputs of each unit's set."""
    _fn = None

    @classmethod
    def apply(cls, w, search, flt):
        if cls._fn is None:
            torch = _torch()

            class _F(torch.autograd.Function):
                @staticmethod
                def forward(ctx_, w_, s_, flt_):
                    sse, resid = s_.sgr_forward(w_, flt_)
                    ctx_.s, ctx_.flt, ctx_.resid = s_, flt_, resid
                    return s_.per_unit(sse)

                @staticmethod
                def backward(ctx_, g):
                    gi = ctx_.s.per_unit(ctx_.s.sgr_backward(ctx_.flt, ctx_.resid))
                    return (gi * g.to(gi.dtype)[:, None]).float(), None, None

            cls._fn = _F
        return cls._fn.apply(w, search, flt)


class Search:
    """The restoration search state of one or more frames on the GPU.

    ctx: api.Context; entries: list of (df, deblocked, cdef, source), one
    per frame, where df is an api.DeviceFrame whose frame.params give the
    unit sizes and counts, and the planes are device planes of the frame's
    sample type. (Search(ctx, df, deblocked, cdef, source) is one frame.)
    Units of all frames are numbered together; every kernel runs once over
    all frames' items."""

    def __init__(self, ctx, entries, deblocked=None, cdef=None, source=None, stream=None):
        torch = _torch()
        self.torch = torch
        if not isinstance(entries, (list, tuple)):
            entries = [(entries, deblocked, cdef, source)]
        self.ctx, self.stream, self.entries = ctx, stream, list(entries)
        units, items, self.frame_units, self.frame_items = [], [], [], []
        u0 = n0 = 0
        for df, _, _, _ in self.entries:
            fu, fi = unit_layout(df.frame.params)
            fi = fi.copy()
            fi[:, 7] += u0
            units.append(fu)
            items.append(fi)
            self.frame_units.append((u0, u0 + len(fu)))
            self.frame_items.append((n0, n0 + len(fi)))
            u0 += len(fu)
            n0 += len(fi)
        self.units = np.concatenate(units).astype(np.int32)
        self.frame_of_unit = np.concatenate([np.full(len(fu), k, np.int32) for k, fu in enumerate(units)])
        items = np.concatenate(items).astype(np.int32)
        self.U, self.n = len(self.units), len(items)
        dev = self.entries[0][1][0].device
        self.items_host = items
        self.items = torch.from_numpy(items).to(dev)
        self.unit_of_item = self.items[:, 7].long()
        self.plane_of_unit = torch.from_numpy(self.units[:, 0]).to(dev).long()
        self.win = torch.empty((self.n, WS, WS), dtype=torch.float32, device=dev)
        self.u = torch.empty((self.n, T, T), dtype=torch.float32, device=dev)
        self.src = torch.empty((self.n, T, T), dtype=torch.float32, device=dev)
        for k, (df, deb, cd, so) in enumerate(self.entries):
            i0, i1 = self.frame_items[k]
            if i1 == i0:
                continue
            pd, pc, ps = (api.planes_struct(x, df.bps) for x in (deb, cd, so))
            ctx.check(api.lib().rvpf_lrd_gather(ctx.h, C.byref(df.frame.params), C.byref(pd), C.byref(pc),
                                                C.byref(ps), _ptr(self.items[i0:]), i1 - i0, self.U,
                                                _ptr(self.win[i0:]), _ptr(self.u[i0:]), _ptr(self.src[i0:]),
                                                api._stream_handle(stream)))
        # chroma Wiener filters have no outermost tap (lr.rs:247)
        self.tap_mask = torch.ones((self.U, 2, 3), dtype=torch.float32, device=dev)
        self.tap_mask[self.plane_of_unit > 0, :, 0] = 0.0

    # single-frame conveniences
    @property
    def df(self):
        return self.entries[0][0]

    @property
    def params(self):
        return self.entries[0][0].frame.params

    # -- kernels ---------------------------------------------------------
    def per_unit(self, per_item):
        torch = self.torch
        shape = (self.U,) + tuple(per_item.shape[1:])
        return torch.zeros(shape, dtype=per_item.dtype, device=per_item.device).index_add_(
            0, self.unit_of_item, per_item)

    def wiener_forward(self, taps):
        torch = self.torch
        taps = taps.detach().contiguous().float()
        resid = torch.empty_like(self.src)
        sse = torch.empty(self.n, dtype=torch.float64, device=self.src.device)
        self.ctx.check(api.lib().rvpf_lrd_wiener(self.ctx.h, 0, _ptr(self.items), self.n, self.U, _ptr(self.win),
                                                 _ptr(self.src), _ptr(taps), _ptr(resid), _ptr(sse),
                                                 api._stream_handle(self.stream)))
        return sse, resid

    def wiener_backward(self, taps, resid):
        torch = self.torch
        taps = taps.detach().contiguous().float()
        g = torch.empty((self.n, 6), dtype=torch.float64, device=self.src.device)
        self.ctx.check(api.lib().rvpf_lrd_wiener(self.ctx.h, 1, _ptr(self.items), self.n, self.U, _ptr(self.win),
                                                 None, _ptr(taps), _ptr(resid), _ptr(g),
                                                 api._stream_handle(self.stream)))
        return g

    def wiener_value_grad(self, taps):
        """(per-unit relaxed squared error [U], gradient [U, 2, 3]) in float64."""
        sse, resid = self.wiener_forward(taps)
        g = self.per_unit(self.wiener_backward(taps, resid)).view(self.U, 2, 3)
        return self.per_unit(sse), g

    def sgr_filter(self, sets):
        """Relaxed filter outputs [n, 2, 64, 64] for sets (one int, or [U])."""
        torch = self.torch
        sets = torch.as_tensor(sets, dtype=torch.int32, device=self.src.device).expand(self.U).contiguous()
        flt = torch.empty((self.n, 2, T, T), dtype=torch.float32, device=self.src.device)
        for k, (df, _, _, _) in enumerate(self.entries):
            i0, i1 = self.frame_items[k]
            if i1 == i0:
                continue
            self.ctx.check(api.lib().rvpf_lrd_sgr_filter(
                self.ctx.h, df.frame.params.bit_depth, _ptr(self.items[i0:]), i1 - i0, self.U, _ptr(self.win[i0:]),
                _ptr(self.u[i0:]), _ptr(sets), _ptr(flt[i0:]), api._stream_handle(self.stream)))
        return flt

    def sgr_forward(self, w, flt):
        torch = self.torch
        w = w.detach().contiguous().float()
        resid = torch.empty_like(self.src)
        sse = torch.empty(self.n, dtype=torch.float64, device=self.src.device)
        self.ctx.check(api.lib().rvpf_lrd_sgr(self.ctx.h, 0, _ptr(self.items), self.n, self.U, _ptr(flt),
                                              _ptr(self.u), _ptr(self.src), _ptr(w), _ptr(resid), _ptr(sse),
                                              api._stream_handle(self.stream)))
        return sse, resid

    def sgr_backward(self, flt, resid):
        torch = self.torch
        g = torch.empty((self.n, 2), dtype=torch.float64, device=self.src.device)
        self.ctx.check(api.lib().rvpf_lrd_sgr(self.ctx.h, 1, _ptr(self.items), self.n, self.U, _ptr(flt),
                                              _ptr(self.u), None, None, _ptr(resid), _ptr(g),
                                              api._stream_handle(self.stream)))
        return g

    def sgr_value_grad(self, w, flt):
        sse, resid = self.sgr_forward(w, flt)
        return self.per_unit(sse), self.per_unit(self.sgr_backward(flt, resid))

    # -- fitting ---------------------------------------------------------
    def fit_wiener(self, steps=3, taps=None):
        """Alternating exact Newton steps on the relaxed error: step s solves
        the vertical taps (s even) or the horizontal ones (s odd) with the
        other direction fixed. steps=3 from zero taps is fit_wiener's
        solution (lr.rs:240-325). Returns float taps [U, 2, 3] (fractions)."""
        torch = self.torch
        dev = self.src.device
        taps = torch.zeros((self.U, 2, 3), dtype=torch.float64, device=dev) if taps is None else taps.double()
        chroma = self.plane_of_unit > 0
        eye = torch.eye(3, dtype=torch.float64, device=dev)
        for s in range(steps):
            d = s % 2 == 1  # 0: vertical, 1: horizontal
            _, g0 = self.wiener_value_grad(taps)
            g0 = g0[:, int(d), :]
            H = torch.empty((self.U, 3, 3), dtype=torch.float64, device=dev)
            for k in range(3):
                t2 = taps.clone()
                t2[:, int(d), k] += 1.0
                _, gk = self.wiener_value_grad(t2)
                H[:, :, k] = gk[:, int(d), :] - g0
            H = 0.5 * (H + H.transpose(1, 2))
            # chroma: the outermost tap stays 0
            H[chroma, 0, :] = 0.0
            H[chroma, :, 0] = 0.0
            H[chroma, 0, 0] = 1.0
            g0 = g0.clone()
            g0[chroma, 0] = 0.0
            # singular systems keep zero taps, as solve() does (lr.rs:209-234)
            det = torch.linalg.det(H)
            scale = H.diagonal(dim1=1, dim2=2).abs().amax(dim=1).clamp_min(1e-30)
            ok = det.abs() > 1e-12 * scale ** 3
            Hs = torch.where(ok[:, None, None], H, eye.expand_as(H))
            sol = taps[:, int(d), :] - torch.linalg.solve(Hs, g0.unsqueeze(-1)).squeeze(-1)
            taps[:, int(d), :] = torch.where(ok[:, None], sol, torch.zeros_like(sol))
        return taps

    def fit_sgr(self, set_, flt=None):
        """Least-squares (w0, w2) of the relaxed projection for one set, from
        gradients (the relaxed error is quadratic in the weights; fit_sgr,
        lr.rs:332-430). Returns (w [U, 2] float64, flt)."""
        torch = self.torch
        dev = self.src.device
        if flt is None:
            flt = self.sgr_filter(set_)
        r0, _, r1, _ = SGR_PARAMS[set_]
        w = torch.zeros((self.U, 2), dtype=torch.float64, device=dev)
        _, g0 = self.sgr_value_grad(w, flt)
        H = torch.zeros((self.U, 2, 2), dtype=torch.float64, device=dev)
        for k in range(2):
            w2 = w.clone()
            w2[:, k] = 1.0
            _, gk = self.sgr_value_grad(w2, flt)
            H[:, :, k] = gk - g0
        H = 0.5 * (H + H.transpose(1, 2))
        if r0 and r1:
            det = torch.linalg.det(H)
            ok = det.abs() > 1e-12 * H.diagonal(dim1=1, dim2=2).abs().amax(dim=1).clamp_min(1e-30) ** 2
            eye = torch.eye(2, dtype=torch.float64, device=dev).expand_as(H)
            sol = -torch.linalg.solve(torch.where(ok[:, None, None], H, eye), g0.unsqueeze(-1)).squeeze(-1)
            w = torch.where(ok[:, None], sol, torch.zeros_like(sol))
        elif r0:
            h = H[:, 0, 0]
            w[:, 0] = torch.where(h > 0, -g0[:, 0] / h.clamp_min(1e-30), torch.zeros_like(h))
        else:
            h = H[:, 1, 1]
            w[:, 1] = torch.where(h > 0, -g0[:, 1] / h.clamp_min(1e-30), torch.zeros_like(h))
        return w, flt

    # -- quantisation ----------------------------------------------------
    def quantise_wiener(self, taps):
        """round(c * 128) clamped to the coded ranges (lr.rs:311-322)."""
        t = round_half_away(taps.detach().cpu().numpy() * 128.0)
        q = np.clip(t, WIENER_TAPS_MIN, WIENER_TAPS_MAX).astype(np.int32)
        q[self.units[:, 0] > 0, :, 0] = 0
        return q

    @staticmethod
    def quantise_sgr(w, set_):
        """xqd from (w0, w2) as fit_sgr does (lr.rs:404-413)."""
        r0, _, r1, _ = SGR_PARAMS[set_]
        w = w.detach().cpu().numpy()
        x0 = np.clip(round_half_away(w[:, 0]), SGRPROJ_XQD_MIN[0], SGRPROJ_XQD_MAX[0]) if r0 else np.zeros(len(w))
        if r1:
            x1 = np.clip(128 - x0 - round_half_away(w[:, 1]), SGRPROJ_XQD_MIN[1], SGRPROJ_XQD_MAX[1])
        else:
            x1 = np.clip(128 - x0, SGRPROJ_XQD_MIN[1], SGRPROJ_XQD_MAX[1])
        return np.stack([x0, x1], axis=1).astype(np.int32)

    # -- exact judge -----------------------------------------------------
    def unit_records(self, kind, q=None, set_=None):
        """Per-plane arrays of rvpf_lr_unit bytes giving every unit the same
        kind of filter: "wiener" with taps q[U, 2, 3], "sgr" with set_ and
        xqd q[U, 2]."""
        recs = np.zeros((self.U, 12), dtype=np.uint8)
        if kind == "wiener":
            recs[:, 0] = abi.RESTORE_WIENER
            recs[:, 4:10] = q.reshape(self.U, 6).astype(np.int8).view(np.uint8)
        else:
            recs[:, 0] = abi.RESTORE_SGRPROJ
            recs[:, 1] = set_
            recs[:, 2:4] = q.astype(np.int8).view(np.uint8)
        return recs

    def exact_sse(self, recs):
        """Exact (integer) squared error per unit with the units' filters
        recs[U, 12] (rvpf_lr over each frame, then per-rectangle errors);
        recs None: the CDEF output itself."""
        torch = self.torch
        dev = self.src.device
        out = torch.empty(self.n, dtype=torch.float64, device=dev)
        for k, (df, deb, cd, so) in enumerate(self.entries):
            i0, i1 = self.frame_items[k]
            u0, u1 = self.frame_units[k]
            if i1 == i0:
                continue
            if recs is None:
                planes = cd
            else:
                params = abi.RvpfFrameParams.from_buffer_copy(df.frame.params)
                side = abi.RvpfSideInfo.from_buffer_copy(df.side)
                keep = []
                nplanes = 1 if params.mono_chrome else 3
                fu = self.units[u0:u1]
                for p in range(nplanes):
                    params.frame_restoration_type[p] = abi.RESTORE_SWITCHABLE
                    t = torch.from_numpy(np.ascontiguousarray(recs[u0:u1][fu[:, 0] == p])).to(dev)
                    keep.append(t)
                    side.lr_units[p] = t.data_ptr()
                planes = [None if x is None else torch.empty_like(x) for x in cd]
                pd, pc, po = (api.planes_struct(x, df.bps) for x in (deb, cd, planes))
                self.ctx.check(api.lib().rvpf_lr(self.ctx.h, C.byref(params), C.byref(side), C.byref(pd),
                                                 C.byref(pc), C.byref(po), api._stream_handle(self.stream)))
            pa, pb = api.planes_struct(planes, df.bps), api.planes_struct(so, df.bps)
            self.ctx.check(api.lib().rvpf_lrd_rect_sse(self.ctx.h, C.byref(df.frame.params), _ptr(self.items[i0:]),
                                                       i1 - i0, self.U, C.byref(pa), C.byref(pb), _ptr(out[i0:]),
                                                       api._stream_handle(self.stream)))
        return self.per_unit(out)

    def search(self, lam, sets=SGR_SETS_NORMAL, wiener_steps=3):
        """The per-unit choice of lr.rs:153-186, with lam one Lagrange
        multiplier or one per frame: returns (recs [U, 12] of the chosen
        rvpf_lr_unit bytes, cost [U], frame types [frame][plane])."""
        torch = self.torch
        dev = self.src.device
        lam_u = torch.as_tensor(np.asarray(lam, dtype=np.float64), device=dev)
        if lam_u.dim() == 1:
            lam_u = lam_u[torch.from_numpy(self.frame_of_unit).to(dev).long()]
        cands = [np.zeros((self.U, 12), dtype=np.uint8)]
        costs = [self.exact_sse(None) + lam_u * RATE_NONE]
        taps = self.fit_wiener(wiener_steps)
        rw = self.unit_records("wiener", self.quantise_wiener(taps))
        cands.append(rw)
        costs.append(self.exact_sse(rw) + lam_u * torch.where(self.plane_of_unit == 0, RATE_WIENER_LUMA,
                                                              RATE_WIENER_CHROMA))
        for s_ in sets:
            w, _ = self.fit_sgr(s_)
            rs = self.unit_records("sgr", self.quantise_sgr(w, s_), s_)
            cands.append(rs)
            costs.append(self.exact_sse(rs) + lam_u * RATE_SGR)
        # the first candidate of least cost (strict <, in the reference's order)
        c = torch.stack(costs, dim=1).cpu().numpy()
        pick = np.argmin(c, axis=1)
        choice = np.stack(cands, axis=0)[pick, np.arange(self.U)]
        best = torch.from_numpy(c[np.arange(self.U), pick]).to(dev)
        ftypes = []
        for k, (df, _, _, _) in enumerate(self.entries):
            u0, u1 = self.frame_units[k]
            fu, fc = self.units[u0:u1], choice[u0:u1]
            nplanes = 1 if df.frame.params.mono_chrome else 3
            ft = []
            for p in range(nplanes):
                t = fc[fu[:, 0] == p, 0]
                w_, s2 = bool((t == abi.RESTORE_WIENER).any()), bool((t == abi.RESTORE_SGRPROJ).any())
                ft.append(abi.RESTORE_SWITCHABLE if w_ and s2 else abi.RESTORE_WIENER if w_ else
                          abi.RESTORE_SGRPROJ if s2 else abi.RESTORE_NONE)
            ftypes.append(ft)
        return choice, best, (ftypes[0] if len(ftypes) == 1 else ftypes)
