"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, never by the product package.
"""
import ctypes as C
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, os.path.join(ROOT, "rivet-av1_amd"))
from rvpf import abi  # noqa: E402

_lib = None


class OrcPlanes(C.Structure):
    _fields_ = [("data", C.c_void_p * 3), ("stride", C.c_int64 * 3)]


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        P = C.POINTER
        L.orc_lf_filter_lanes.argtypes = [C.c_void_p] + [C.c_int] * 6
        L.orc_filter_strength.argtypes = [P(abi.RvpfFrameParams), P(abi.RvpfSideInfo), C.c_int, C.c_int,
                                          C.c_int, C.c_int, C.c_void_p]
        L.orc_loop_filter.argtypes = [P(abi.RvpfFrameParams), P(abi.RvpfSideInfo), P(OrcPlanes), C.c_int]
        L.orc_cdef_direction.argtypes = [C.c_void_p, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.orc_cdef_direction.restype = C.c_int
        L.orc_cdef_filter_window.argtypes = [C.c_void_p] + [C.c_int] * 7 + [C.c_void_p]
        L.orc_cdef.argtypes = [P(abi.RvpfFrameParams), P(abi.RvpfSideInfo), P(OrcPlanes), P(OrcPlanes), C.c_int]
        L.orc_lr.argtypes = [P(abi.RvpfFrameParams), P(abi.RvpfSideInfo), P(OrcPlanes), P(OrcPlanes),
                             P(OrcPlanes), C.c_int]
        L.orc_grain_templates.argtypes = [P(abi.RvpfFrameParams), C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_grain_scaling_luts.argtypes = [P(abi.RvpfFrameParams), C.c_void_p]
        L.orc_grain_apply.argtypes = [P(abi.RvpfFrameParams), C.c_int, C.c_int, C.c_void_p * 3]
        L.orc_postfilter_frame.argtypes = [P(abi.RvpfFrameParams), P(abi.RvpfSideInfo), P(OrcPlanes),
                                           P(OrcPlanes), C.c_void_p]
        L.orc_upscale.argtypes = [P(abi.RvpfFrameParams), P(OrcPlanes), P(OrcPlanes)]
        L.orc_grain_rng_step.argtypes = [C.c_uint32]
        L.orc_grain_rng_step.restype = C.c_uint32
        OP = P(OrcPlanes)
        L.orc_fit_wiener.argtypes = [P(abi.RvpfFrameParams), OP, OP, OP, C.c_int, C.c_int, C.c_int, C.c_int,
                                     C.c_void_p, C.c_void_p]
        L.orc_fit_sgr.argtypes = [P(abi.RvpfFrameParams), OP, OP, OP, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_unit_sse.argtypes = [P(abi.RvpfFrameParams), OP, OP, OP, C.c_int, C.c_int, C.c_int, C.c_void_p]
        L.orc_unit_sse.restype = C.c_uint64
        L.orc_lr_search_unit.argtypes = [P(abi.RvpfFrameParams), OP, OP, OP, C.c_int, C.c_int, C.c_int, C.c_double,
                                         C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p]
        L.orc_cdef_search.argtypes = [P(abi.RvpfFrameParams), P(abi.RvpfSideInfo), OP, OP, C.c_int, C.c_double,
                                      C.c_int, C.c_int, P(abi.RvpfCdefSearchResult), C.c_void_p, C.c_void_p,
                                      C.c_void_p, C.c_void_p]
        L.orc_cdef_candidates.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p]
        L.orc_lf_search.argtypes = [P(abi.RvpfFrameParams), P(abi.RvpfSideInfo), OP, OP, C.c_int, C.c_void_p,
                                    C.c_void_p]
        _lib = L
    return _lib


def planes_struct(arrs):
    s = OrcPlanes()
    for p, a in enumerate(arrs):
        if a is None:
            continue
        assert a.dtype == np.uint16 and a.flags.c_contiguous
        s.data[p] = a.ctypes.data
        s.stride[p] = a.shape[1]
    return s


def lf_filter_lanes(v, limit, blimit, thresh, filter_size, chroma, bd):
    """v: int32 array [14, 4]; filtered in place."""
    assert v.dtype == np.int32 and v.shape == (14, 4) and v.flags.c_contiguous
    lib().orc_lf_filter_lanes(v.ctypes.data, limit, blimit, thresh, filter_size, int(chroma), bd)
    return v


def cdef_direction(luma, x0, y0, bd):
    var = C.c_int32()
    d = lib().orc_cdef_direction(luma.ctypes.data, luma.shape[1], x0, y0, bd, C.byref(var))
    return d, var.value


def cdef_filter_window(win, w, h, pri, sec, damping, direction, coeff_shift):
    """win: int32 [144] (INT32_MIN = unavailable); returns int32 [h, w]."""
    assert win.dtype == np.int32 and win.size == 144 and win.flags.c_contiguous
    out = np.zeros(64, dtype=np.int32)
    lib().orc_cdef_filter_window(win.ctypes.data, w, h, pri, sec, damping, direction, coeff_shift, out.ctypes.data)
    return out[:w * h].reshape(h, w)


def loop_filter(frame, planes, threads=1):
    s = frame.host_side_info()
    ps = planes_struct(planes)
    lib().orc_loop_filter(C.byref(frame.params), C.byref(s), C.byref(ps), threads)
    return planes


def cdef(frame, deblocked, threads=1):
    out = [None if a is None else np.zeros_like(a) for a in deblocked]
    s = frame.host_side_info()
    pi, po = planes_struct(deblocked), planes_struct(out)
    lib().orc_cdef(C.byref(frame.params), C.byref(s), C.byref(pi), C.byref(po), threads)
    return out


def upscale(frame, planes):
    """Super-resolution upscaling (7.16) of frame-width planes into zeroed
    planes of the reference-frame shape."""
    out = [None if a is None else np.zeros(frame.ref_shape(p), dtype=np.uint16) for p, a in enumerate(planes)]
    pi, po = planes_struct(planes), planes_struct(out)
    lib().orc_upscale(C.byref(frame.params), C.byref(pi), C.byref(po))
    return out


def lr(frame, deblocked, cdef_planes, threads=1):
    out = [None if a is None else np.zeros_like(a) for a in cdef_planes]
    s = frame.host_side_info()
    pd, pc, po = planes_struct(deblocked), planes_struct(cdef_planes), planes_struct(out)
    if lib().orc_lr(C.byref(frame.params), C.byref(s), C.byref(pd), C.byref(pc), C.byref(po), threads) != 0:
        raise MemoryError("orc_lr: cannot allocate its job list")
    return out


def grain_templates(params):
    t = [np.zeros(73 * 82, dtype=np.int32) for _ in range(3)]
    lib().orc_grain_templates(C.byref(params), t[0].ctypes.data, t[1].ctypes.data, t[2].ctypes.data)
    return [a.reshape(73, 82) for a in t]


def grain_scaling_luts(params):
    lut = np.zeros((3, 256), dtype=np.int32)
    lib().orc_grain_scaling_luts(C.byref(params), lut.ctypes.data)
    return lut


def grain_apply(params, w, h, planes):
    """planes: tight uint16 arrays (luma h x w, chroma rounded up); in place."""
    arr = (C.c_void_p * 3)()
    for p, a in enumerate(planes):
        if a is not None:
            assert a.dtype == np.uint16 and a.flags.c_contiguous
            arr[p] = a.ctypes.data
    lib().orc_grain_apply(C.byref(params), w, h, arr)
    return planes


def rng_step(state):
    return lib().orc_grain_rng_step(state)


def postfilter_frame(frame, threads=1, recon=None):
    """Runs the chain on copies of frame.rec. Returns (deblocked, ref, display)."""
    rec = [None if a is None else a.copy() for a in (recon or frame.rec)]
    ref = [None if a is None else np.zeros(frame.ref_shape(p), dtype=np.uint16) for p, a in enumerate(rec)]
    disp = []
    arr = (C.c_void_p * 3)()
    for p in range(3):
        if p < frame.num_planes:
            d = np.zeros(frame.crop_shape(p), dtype=np.uint16)
            arr[p] = d.ctypes.data
            disp.append(d)
        else:
            disp.append(None)
    s = frame.host_side_info()
    pr, po = planes_struct(rec), planes_struct(ref)
    if lib().orc_postfilter_frame(C.byref(frame.params), C.byref(s), C.byref(pr), C.byref(po), arr, threads) != 0:
        raise MemoryError("orc_postfilter_frame failed")
    return rec, ref, disp


class LrSearchFrame:
    """Host planes of one frame for the encoder restoration search:
    deblocked, CDEF output and source (uint16 arrays)."""

    def __init__(self, frame, deblocked, cdef_planes, source):
        self.frame = frame
        self.keep = (deblocked, cdef_planes, source)
        self.pd = planes_struct(deblocked)
        self.pc = planes_struct(cdef_planes)
        self.ps = planes_struct(source)

    def _args(self):
        return (C.byref(self.frame.params), C.byref(self.pd), C.byref(self.pc), C.byref(self.ps))

    def fit_wiener(self, plane, ur, uc, step=1):
        """(quantised taps [2, 3] int, unquantised fractions [2, 3])."""
        q = np.zeros((2, 3), dtype=np.int32)
        f = np.zeros((2, 3), dtype=np.float64)
        lib().orc_fit_wiener(*self._args(), plane, ur, uc, step, q.ctypes.data, f.ctypes.data)
        return q, f

    def fit_sgr(self, plane, ur, uc, set_, step=1):
        """(xqd [2] int, unquantised (w0, w2), exact squared error)."""
        q = np.zeros(2, dtype=np.int32)
        w = np.zeros(2, dtype=np.float64)
        e = C.c_uint64()
        lib().orc_fit_sgr(*self._args(), plane, ur, uc, set_, step, q.ctypes.data, w.ctypes.data, C.byref(e))
        return q, w, e.value

    def unit_sse(self, plane, ur, uc, unit_bytes=None):
        """Exact squared error of the unit with the rvpf_lr_unit bytes (None: CDEF output)."""
        if unit_bytes is None:
            return lib().orc_unit_sse(*self._args(), plane, ur, uc, None)
        u = np.ascontiguousarray(unit_bytes, dtype=np.uint8)
        return lib().orc_unit_sse(*self._args(), plane, ur, uc, u.ctypes.data)

    def search_unit(self, plane, ur, uc, lam, sets, step=1):
        """(chosen rvpf_lr_unit bytes [12], cost)."""
        out = np.zeros(12, dtype=np.uint8)
        cost = C.c_double()
        st = np.ascontiguousarray(sets, dtype=np.int32)
        lib().orc_lr_search_unit(*self._args(), plane, ur, uc, lam, st.ctypes.data, len(st), step, out.ctypes.data,
                                 C.byref(cost))
        return out, cost.value


def cdef_candidates(thorough):
    """([(pri, sec)] luma, [(pri, sec)] chroma) of the encoder's CDEF search."""
    a = [np.zeros(44, dtype=np.int32), np.zeros(44, dtype=np.int32), np.zeros(24, dtype=np.int32),
         np.zeros(24, dtype=np.int32)]
    ny, nuv = C.c_int(), C.c_int()
    lib().orc_cdef_candidates(int(thorough), *[x.ctypes.data for x in a], C.byref(ny), C.byref(nuv))
    return ([(int(a[0][i]), int(a[1][i])) for i in range(ny.value)],
            [(int(a[2][i]), int(a[3][i])) for i in range(nuv.value)])


def cdef_search(frame, deblocked, source, base_q_idx, lam, thorough, threads=1):
    """cdef::search: (RvpfCdefSearchResult, cdef_idx table, per-block luma
    errors [nb, ny], chroma errors [nb, nuv])."""
    ycand, uvcand = cdef_candidates(thorough)
    g = frame.geom
    maxb = ((g.mi_rows + 15) // 16) * ((g.mi_cols + 15) // 16)
    ys = np.zeros(maxb * len(ycand), dtype=np.uint64)
    uvs = np.zeros(maxb * len(uvcand), dtype=np.uint64)
    table = frame.cdef_idx.copy()  # rows past the 64x64 grid keep their value
    res = abi.RvpfCdefSearchResult()
    nb = C.c_int()
    s = frame.host_side_info()
    pd, ps = planes_struct(deblocked), planes_struct(source)
    lib().orc_cdef_search(C.byref(frame.params), C.byref(s), C.byref(pd), C.byref(ps), base_q_idx, lam,
                          int(thorough), threads, C.byref(res), table.ctypes.data, ys.ctypes.data, uvs.ctypes.data,
                          C.byref(nb))
    n = nb.value
    return res, table, ys[:n * len(ycand)].reshape(n, len(ycand)), uvs[:n * len(uvcand)].reshape(n, len(uvcand))


def lf_search(frame, recon, source, threads=1):
    """search_lf: (levels [4], squared errors [6 scales, 3 planes])."""
    out = np.zeros(4, dtype=np.int32)
    sse = np.zeros((6, 3), dtype=np.uint64)
    s = frame.host_side_info()
    pr, ps = planes_struct(recon), planes_struct(source)
    lib().orc_lf_search(C.byref(frame.params), C.byref(s), C.byref(pr), C.byref(ps), threads, out.ctypes.data,
                        sse.ctypes.data)
    return [int(v) for v in out], sse
