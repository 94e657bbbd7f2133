"""Device-side binding of librvpf.so (the HIP post-filter chain).

Device memory and streams come from PyTorch (ROCm); every compute call
goes through the C ABI of include/rvpf.h. There is no CPU fallback: if the
library is missing or no GPU is present, the calls raise.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_lib = None


class RvpfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("rvpf error %d: %s" % (code, msg))
        self.code = code


def lib():
    """Loads librvpf.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        path = abi.lib_path("librvpf.so")
        if not os.path.exists(path):
            raise RuntimeError("librvpf.so is not built (run __graft_entry__.build()): %s" % path)
        L = C.CDLL(path)
        P = C.POINTER
        L.rvpf_ctx_create.argtypes = [P(C.c_void_p)]
        L.rvpf_ctx_destroy.argtypes = [C.c_void_p]
        L.rvpf_ctx_destroy.restype = None
        L.rvpf_ctx_sync.argtypes = [C.c_void_p, C.c_void_p]
        L.rvpf_ctx_last_error.argtypes = [C.c_void_p]
        L.rvpf_ctx_last_error.restype = C.c_char_p
        L.rvpf_ctx_set_profiling.argtypes = [C.c_void_p, C.c_int]
        L.rvpf_ctx_read_profile.argtypes = [C.c_void_p, P(abi.RvpfProfile)]
        L.rvpf_validate.argtypes = [P(abi.RvpfFrameParams), P(abi.RvpfSideInfo)]
        L.rvpf_deblock.argtypes = [C.c_void_p, P(abi.RvpfFrameParams), P(abi.RvpfSideInfo), P(abi.RvpfPlanes), C.c_void_p]
        L.rvpf_cdef.argtypes = [C.c_void_p, P(abi.RvpfFrameParams), P(abi.RvpfSideInfo), P(abi.RvpfPlanes),
                                P(abi.RvpfPlanes), C.c_void_p]
        L.rvpf_lr.argtypes = [C.c_void_p, P(abi.RvpfFrameParams), P(abi.RvpfSideInfo), P(abi.RvpfPlanes),
                              P(abi.RvpfPlanes), P(abi.RvpfPlanes), C.c_void_p]
        L.rvpf_grain.argtypes = [C.c_void_p, P(abi.RvpfFrameParams), P(abi.RvpfPlanes), P(abi.RvpfPlanes), C.c_void_p]
        L.rvpf_upscale.argtypes = [C.c_void_p, P(abi.RvpfFrameParams), P(abi.RvpfPlanes), P(abi.RvpfPlanes), C.c_void_p]
        L.rvpf_grain_noalloc.argtypes = [P(abi.RvpfFrameParams), P(abi.RvpfPlanes), P(abi.RvpfPlanes), C.c_void_p]
        L.rvpf_packed_layout_of.argtypes = [P(abi.RvpfFrameParams), P(abi.RvpfPackedLayout)]
        L.rvpf_lf_search.argtypes = [C.c_void_p, P(abi.RvpfFrameParams), P(abi.RvpfSideInfo), P(abi.RvpfPlanes),
                                     P(abi.RvpfPlanes), C.c_void_p, C.c_void_p]
        L.rvpf_cdef_search.argtypes = [C.c_void_p, P(abi.RvpfFrameParams), P(abi.RvpfSideInfo), P(abi.RvpfPlanes),
                                       P(abi.RvpfPlanes), C.c_int32, C.c_double, C.c_int32,
                                       P(abi.RvpfCdefSearchResult), C.c_void_p, C.c_void_p]
        L.rvpf_output_packed.argtypes = [C.c_void_p, P(abi.RvpfFrameParams), P(abi.RvpfPlanes), C.c_void_p,
                                         C.c_int64, C.c_void_p]
        L.rvpf_postfilter_frame.argtypes = [C.c_void_p, P(abi.RvpfFrameParams), P(abi.RvpfSideInfo),
                                            P(abi.RvpfPlanes), P(abi.RvpfPlanes), P(abi.RvpfPlanes), C.c_void_p]
        L.rvpf_batch_create.argtypes = [C.c_void_p, P(abi.RvpfBatchFrame), C.c_int32, P(C.c_void_p)]
        L.rvpf_batch_destroy.argtypes = [C.c_void_p]
        L.rvpf_batch_destroy.restype = None
        L.rvpf_batch_autotune.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        L.rvpf_batch_get_tiles.argtypes = [C.c_void_p, P(abi.RvpfTileConfig)]
        L.rvpf_batch_set_tiles.argtypes = [C.c_void_p, P(abi.RvpfTileConfig)]
        L.rvpf_batch_run.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        V = C.c_void_p
        L.rvpf_lrd_gather.argtypes = [V, P(abi.RvpfFrameParams), P(abi.RvpfPlanes), P(abi.RvpfPlanes),
                                      P(abi.RvpfPlanes), V, C.c_int32, C.c_int32, V, V, V, V]
        L.rvpf_lrd_wiener.argtypes = [V, C.c_int, V, C.c_int32, C.c_int32, V, V, V, V, V, V]
        L.rvpf_lrd_sgr_filter.argtypes = [V, C.c_int32, V, C.c_int32, C.c_int32, V, V, V, V, V]
        L.rvpf_lrd_sgr.argtypes = [V, C.c_int, V, C.c_int32, C.c_int32, V, V, V, V, V, V, V]
        L.rvpf_lrd_rect_sse.argtypes = [V, P(abi.RvpfFrameParams), V, C.c_int32, C.c_int32, P(abi.RvpfPlanes),
                                        P(abi.RvpfPlanes), V, V]
        L.rvpf_ac_q.argtypes = [C.c_int32, C.c_int32]
        L.rvpf_ac_q.restype = C.c_int32
        _lib = L
    return _lib


def validate(params, side=None):
    """Host-side validation (no GPU needed). Returns (code, message)."""
    rc = lib().rvpf_validate(C.byref(params), C.byref(side) if side is not None else None)
    return rc, lib().rvpf_ctx_last_error(None).decode()


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("rvpf needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


def _stream_handle(stream):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


class Context:
    """rvpf_ctx: owns device scratch; one per host thread."""

    def __init__(self):
        _torch()
        h = C.c_void_p()
        rc = lib().rvpf_ctx_create(C.byref(h))
        if rc != 0:
            raise RvpfError(rc, lib().rvpf_ctx_last_error(None).decode())
        self.h = h

    def close(self):
        if self.h:
            lib().rvpf_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def check(self, rc):
        if rc != 0:
            raise RvpfError(rc, lib().rvpf_ctx_last_error(self.h).decode())

    def set_profiling(self, enable):
        self.check(lib().rvpf_ctx_set_profiling(self.h, int(bool(enable))))

    def read_profile(self):
        """{stage: (total ms, launches)} since the last read."""
        pr = abi.RvpfProfile()
        self.check(lib().rvpf_ctx_read_profile(self.h, C.byref(pr)))
        return {name: (pr.ms[i], pr.launches[i]) for i, name in enumerate(abi.STAGES)}

    def sync(self, stream=None):
        self.check(lib().rvpf_ctx_sync(self.h, _stream_handle(stream)))


def planes_struct(tensors, bytes_per_sample=2):
    """rvpf_planes over up to three 2-D device tensors (None = absent)."""
    s = abi.RvpfPlanes()
    s.bytes_per_sample = bytes_per_sample
    for p, t in enumerate(tensors):
        if t is None:
            continue
        assert t.dim() == 2 and t.stride(1) == 1
        s.data[p] = t.data_ptr()
        s.stride[p] = t.stride(0)
        s.width[p] = t.shape[1]
        s.height[p] = t.shape[0]
    return s


class DeviceFrame:
    """A workload.Frame's side info and reconstruction in device memory."""

    def __init__(self, frame, device="cuda", bytes_per_sample=2):
        torch = _torch()
        self.frame = frame
        self.bps = bytes_per_sample
        dt = torch.uint8 if bytes_per_sample == 1 else torch.int16

        def dev(a):
            return torch.from_numpy(np.ascontiguousarray(a)).to(device)

        self.mi = dev(frame.mi)
        self.tx = [dev(t) for t in frame.tx]
        self.cdef_idx = dev(frame.cdef_idx)
        self.lr_units = [dev(u) for u in frame.lr_units]
        self.side = abi.RvpfSideInfo()
        self.side.mi = self.mi.data_ptr()
        self.side.mi_stride = frame.geom.mi_stride
        for p in range(3):
            self.side.tx_size[p] = self.tx[p].data_ptr()
            self.side.lr_units[p] = self.lr_units[p].data_ptr()
        self.side.cdef_idx = self.cdef_idx.data_ptr()
        self.side.cdef_stride = frame.geom.cdef_stride
        self.dtype = dt
        self.device = device
        self.rec_host = frame.rec

    def upload_planes(self, arrs):
        """Device copies of host uint16 planes in this frame's sample type."""
        torch = _torch()
        out = []
        for a in arrs:
            if a is None:
                out.append(None)
                continue
            if self.bps == 1:
                out.append(torch.from_numpy(a.astype(np.uint8)).to(self.device))
            else:
                out.append(torch.from_numpy(a.view(np.int16)).to(self.device))
        return out

    def alloc_like_recon(self):
        torch = _torch()
        return [None if a is None else torch.zeros(a.shape, dtype=self.dtype, device=self.device)
                for a in self.rec_host]

    def alloc_ref(self):
        """Reference-frame planes (upscaled with super-resolution)."""
        torch = _torch()
        return [torch.zeros(self.frame.ref_shape(p), dtype=self.dtype, device=self.device)
                if p < self.frame.num_planes else None for p in range(3)]

    def alloc_display(self):
        torch = _torch()
        return [torch.zeros(self.frame.crop_shape(p), dtype=self.dtype, device=self.device)
                if p < self.frame.num_planes else None for p in range(3)]


def to_host(tensors):
    """Device planes -> host uint16 numpy arrays."""
    out = []
    for t in tensors:
        if t is None:
            out.append(None)
            continue
        a = t.cpu().numpy()
        out.append(a.astype(np.uint16) if a.dtype == np.uint8 else a.view(np.uint16))
    return out


def deblock(ctx, df, planes, stream=None):
    ps = planes_struct(planes, df.bps)
    ctx.check(lib().rvpf_deblock(ctx.h, C.byref(df.frame.params), C.byref(df.side), C.byref(ps),
                                 _stream_handle(stream)))


def cdef(ctx, df, deblocked, out, stream=None):
    pi, po = planes_struct(deblocked, df.bps), planes_struct(out, df.bps)
    ctx.check(lib().rvpf_cdef(ctx.h, C.byref(df.frame.params), C.byref(df.side), C.byref(pi), C.byref(po),
                              _stream_handle(stream)))


def lr(ctx, df, deblocked, cdef_planes, out, stream=None):
    pd, pc, po = planes_struct(deblocked, df.bps), planes_struct(cdef_planes, df.bps), planes_struct(out, df.bps)
    ctx.check(lib().rvpf_lr(ctx.h, C.byref(df.frame.params), C.byref(df.side), C.byref(pd), C.byref(pc),
                            C.byref(po), _stream_handle(stream)))


def upscale(ctx, df, planes, out, stream=None):
    """Super-resolution upscaling of `planes` (the coded width) into `out`
    (the upscaled width); the frame must have use_superres."""
    pi, po = planes_struct(planes, df.bps), planes_struct(out, df.bps)
    ctx.check(lib().rvpf_upscale(ctx.h, C.byref(df.frame.params), C.byref(pi), C.byref(po),
                                 _stream_handle(stream)))


def grain(ctx, df, ref, display, stream=None, display_bps=None):
    """display_bps: 1 for byte display planes from 16-bit planes (8 bits)."""
    pr, pdsp = planes_struct(ref, df.bps), planes_struct(display, display_bps or df.bps)
    ctx.check(lib().rvpf_grain(ctx.h, C.byref(df.frame.params), C.byref(pr), C.byref(pdsp),
                               _stream_handle(stream)))


def grain_noalloc(df, ref, display, stream=None, display_bps=None):
    """rvpf_grain_noalloc: as grain(), with no device allocation (the
    library's static scratch; frames up to 8192x8192). Raises RvpfError."""
    pr, pdsp = planes_struct(ref, df.bps), planes_struct(display, display_bps or df.bps)
    rc = lib().rvpf_grain_noalloc(C.byref(df.frame.params), C.byref(pr), C.byref(pdsp), _stream_handle(stream))
    if rc != 0:
        raise RvpfError(rc, lib().rvpf_ctx_last_error(None).decode())


def packed_layout(params):
    """rvpf_packed_layout_of: where each plane sits in the packed picture
    of Decoder::output (host only)."""
    lay = abi.RvpfPackedLayout()
    rc = lib().rvpf_packed_layout_of(C.byref(params), C.byref(lay))
    if rc != 0:
        raise RvpfError(rc, lib().rvpf_ctx_last_error(None).decode())
    return lay


def alloc_packed(df):
    """A device byte buffer for the frame's packed picture."""
    torch = _torch()
    return torch.zeros(packed_layout(df.frame.params).bytes, dtype=torch.uint8, device=df.device)


def output_packed(ctx, df, ref, packed, stream=None):
    """Crop + film grain of the reference planes into the packed picture
    (a 1-D uint8 device tensor)."""
    pr = planes_struct(ref, df.bps)
    assert packed.dim() == 1 and packed.is_contiguous()
    ctx.check(lib().rvpf_output_packed(ctx.h, C.byref(df.frame.params), C.byref(pr), C.c_void_p(packed.data_ptr()),
                                       packed.numel() * packed.element_size(), _stream_handle(stream)))


def frame_md5(packed, params):
    """MD5 of a packed picture (bytes) as the AOM vectors' .md5 files hash
    it: the packed planes, and for a monochrome frame two neutral (mid-grey)
    4:2:0 chroma planes after them (tests/common/mod.rs:101-119 of the
    reference)."""
    import hashlib
    m = hashlib.md5()
    m.update(bytes(packed))
    if params.mono_chrome:
        cw, ch = (params.upscaled_width + 1) // 2, (params.frame_height + 1) // 2
        mid = 1 << (params.bit_depth - 1)
        plane = (mid.to_bytes(2, "little") if params.bit_depth > 8 else bytes([mid])) * (cw * ch)
        m.update(plane)
        m.update(plane)
    return m.hexdigest()


def lf_search(ctx, df, recon, source, stream=None):
    """search_lf on the GPU: the loop filter levels [4] chosen from the
    frame's start levels (recon is not modified)."""
    pr, ps = planes_struct(recon, df.bps), planes_struct(source, df.bps)
    out = (C.c_int32 * 4)()
    ctx.check(lib().rvpf_lf_search(ctx.h, C.byref(df.frame.params), C.byref(df.side), C.byref(pr), C.byref(ps),
                                   out, _stream_handle(stream)))
    return list(out)


def cdef_search(ctx, df, deblocked, source, base_q_idx, lam, thorough=False, stream=None):
    """cdef::search on the GPU: (RvpfCdefSearchResult, cdef_idx table as
    an int8 numpy array of the frame's cdef_idx size; entries past the
    64x64 grid's rows are the frame's own)."""
    pd, ps = planes_struct(deblocked, df.bps), planes_struct(source, df.bps)
    res = abi.RvpfCdefSearchResult()
    table = df.frame.cdef_idx.copy()
    ctx.check(lib().rvpf_cdef_search(ctx.h, C.byref(df.frame.params), C.byref(df.side), C.byref(pd), C.byref(ps),
                                     int(base_q_idx), float(lam), int(bool(thorough)), C.byref(res),
                                     table.ctypes.data, _stream_handle(stream)))
    return res, table


def postfilter_frame(ctx, df, recon, ref_out, display, stream=None, display_bps=None):
    pr, po = planes_struct(recon, df.bps), planes_struct(ref_out, df.bps)
    pdsp = planes_struct(display, display_bps or df.bps) if display is not None else None
    ctx.check(lib().rvpf_postfilter_frame(ctx.h, C.byref(df.frame.params), C.byref(df.side), C.byref(pr),
                                          C.byref(po), C.byref(pdsp) if pdsp is not None else None,
                                          _stream_handle(stream)))


class Batch:
    """rvpf_batch: many frames' buffers, filtered by one graph launch.

    entries: list of (DeviceFrame, recon planes, ref_out planes, display
    planes or None). The tensors must stay alive as long as the batch."""

    def __init__(self, ctx, entries):
        arr = (abi.RvpfBatchFrame * len(entries))()
        self._keep = []
        for i, (df, recon, ref, disp) in enumerate(entries):
            arr[i].params = C.addressof(df.frame.params)
            arr[i].side = df.side
            arr[i].recon = planes_struct(recon, df.bps)
            arr[i].ref_out = planes_struct(ref, df.bps)
            if disp is not None:
                arr[i].display = planes_struct(disp, df.bps)
            self._keep.append((df, recon, ref, disp))
        self.ctx = ctx
        h = C.c_void_p()
        ctx.check(lib().rvpf_batch_create(ctx.h, arr, len(entries), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            lib().rvpf_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def autotune(self, force=False, stream=None):
        self.ctx.check(lib().rvpf_batch_autotune(self.ctx.h, self.h, int(bool(force)), _stream_handle(stream)))

    def get_tiles(self):
        arr = (abi.RvpfTileConfig * len(abi.TUNE_STAGES))()
        self.ctx.check(lib().rvpf_batch_get_tiles(self.h, arr))
        return {n: (arr[i].shape, arr[i].workgroups_per_cu, arr[i].chunk, arr[i].us)
                for i, n in enumerate(abi.TUNE_STAGES)}

    def set_tiles(self, tiles):
        """tiles: {stage: (shape, workgroups_per_cu, chunk)} (others unchanged)."""
        arr = (abi.RvpfTileConfig * len(abi.TUNE_STAGES))()
        self.ctx.check(lib().rvpf_batch_get_tiles(self.h, arr))
        for i, n in enumerate(abi.TUNE_STAGES):
            if n in tiles:
                arr[i].shape, arr[i].workgroups_per_cu, arr[i].chunk = tiles[n][:3]
        self.ctx.check(lib().rvpf_batch_set_tiles(self.h, arr))

    def run(self, stream=None):
        self.ctx.check(lib().rvpf_batch_run(self.ctx.h, self.h, _stream_handle(stream)))
