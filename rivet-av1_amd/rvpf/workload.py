"""Seeded synthetic frames for the post-filter chain (host numpy arrays).

Wraps librvpf_synth.so (csrc/synth.c). A ``Frame`` holds the frame
parameters, the side info in the boundary layout of include/rvpf.h, the
source picture and the reconstruction, all as numpy arrays.
"""
import ctypes as C
import os

import numpy as np

from . import abi

_synth = None


def synth_lib():
    global _synth
    if _synth is None:
        path = abi.lib_path("librvpf_synth.so")
        if not os.path.exists(path):
            raise RuntimeError(
                "librvpf_synth.so is not built (run __graft_entry__.build()): %s" % path)
        lib = C.CDLL(path)
        lib.synth_geometry.argtypes = [C.POINTER(abi.SynthCfg), C.POINTER(abi.SynthGeom)]
        lib.synth_geometry.restype = C.c_int
        lib.synth_frame.argtypes = [C.POINTER(abi.SynthCfg), C.POINTER(abi.RvpfFrameParams)] + [C.c_void_p] * 14
        lib.synth_frame.restype = C.c_int
        lib.synth_struct_sizes.argtypes = [C.POINTER(C.c_int32)]
        _synth = lib
    return _synth


def struct_sizes():
    out = (C.c_int32 * 11)()
    synth_lib().synth_struct_sizes(out)
    return list(out)


def _ptr(a):
    return None if a is None else a.ctypes.data


class Frame:
    """One synthetic frame: params, side info, source and reconstruction."""

    def __init__(self, width, height, bit_depth=8, ssx=1, ssy=1, mono=False, sb_size=64,
                 qindex=128, seed=1, frame_index=0, pan_px=0.0, lf=True, lf_extras=False,
                 cdef=True, cdef_bits=3, lr=2, lr_unit_size=64, grain=False, grain_lag=3,
                 grain_y_points=8, mc_identity=False, with_source=True, superres_denom=0):
        cfg = abi.SynthCfg()
        cfg.width, cfg.height, cfg.bit_depth = width, height, bit_depth
        cfg.ssx, cfg.ssy, cfg.mono = (1, 1, 1) if mono else (ssx, ssy, 0)
        cfg.sb_size, cfg.qindex, cfg.seed = sb_size, qindex, seed
        cfg.frame_index, cfg.pan_px = frame_index, pan_px
        cfg.lf, cfg.lf_extras, cfg.cdef, cfg.cdef_bits = int(lf), int(lf_extras), int(cdef), cdef_bits
        cfg.lr, cfg.lr_unit_size = lr, lr_unit_size
        cfg.grain, cfg.grain_lag, cfg.grain_y_points = int(grain), grain_lag, grain_y_points
        cfg.mc_identity = int(mc_identity)
        cfg.superres_denom = superres_denom
        self.cfg = cfg
        g = abi.SynthGeom()
        if synth_lib().synth_geometry(C.byref(cfg), C.byref(g)) != 0:
            raise ValueError("bad synthetic frame configuration")
        self.geom = g
        self.num_planes = g.num_planes
        n4 = g.mi_stride * g.mi_alloc_rows
        self.mi = np.zeros(n4 * 8, dtype=np.uint8)
        self.tx = [np.zeros(n4, dtype=np.uint8) for _ in range(3)]
        self.cdef_idx = np.zeros(g.cdef_stride * g.cdef_rows, dtype=np.int8)
        self.lr_units = [np.zeros(max(1, g.lr_unit_rows[p] * g.lr_unit_cols[p]) * 12, dtype=np.uint8)
                         for p in range(3)]
        self.src = [np.zeros((g.plane_h[p], g.plane_w[p]), dtype=np.uint16) if p < g.num_planes else None
                    for p in range(3)] if with_source else [None] * 3
        self.rec = [np.zeros((g.plane_h[p], g.plane_w[p]), dtype=np.uint16) if p < g.num_planes else None
                    for p in range(3)]
        self.params = abi.RvpfFrameParams()
        rc = synth_lib().synth_frame(
            C.byref(cfg), C.byref(self.params), _ptr(self.mi), _ptr(self.tx[0]), _ptr(self.tx[1]),
            _ptr(self.tx[2]), _ptr(self.cdef_idx), _ptr(self.lr_units[0]), _ptr(self.lr_units[1]),
            _ptr(self.lr_units[2]), _ptr(self.src[0]), _ptr(self.src[1]), _ptr(self.src[2]),
            _ptr(self.rec[0]), _ptr(self.rec[1]), _ptr(self.rec[2]))
        if rc != 0:
            raise ValueError("synth_frame failed")

    # ------------------------------------------------------------------
    @property
    def bit_depth(self):
        return self.params.bit_depth

    def plane_shape(self, p):
        return (self.geom.plane_h[p], self.geom.plane_w[p])

    def crop_shape(self, p):
        return (self.geom.crop_h[p], self.geom.crop_w[p])

    def ref_shape(self, p):
        """Allocated shape of a reference-frame plane (upscaled with
        super-resolution, else the reconstruction's)."""
        return (self.geom.plane_h[p], self.geom.ref_w[p])

    def ref_area(self, p):
        """(rows, cols) of a reference-frame plane that the chain defines:
        the mode-info area, or the displayed size with super-resolution."""
        return self.crop_shape(p) if self.params.use_superres else self.mi_area(p)

    def mi_area(self, p):
        """(rows, cols) of the mode-info-aligned area of plane p."""
        sx = self.params.ssx if p else 0
        sy = self.params.ssy if p else 0
        return ((self.params.mi_rows * 4) >> sy, (self.params.mi_cols * 4) >> sx)

    def host_side_info(self):
        """rvpf_side_info pointing at these host arrays (for the oracle)."""
        s = abi.RvpfSideInfo()
        s.mi = self.mi.ctypes.data
        s.mi_stride = self.geom.mi_stride
        for p in range(3):
            s.tx_size[p] = self.tx[p].ctypes.data
            s.lr_units[p] = self.lr_units[p].ctypes.data
        s.cdef_idx = self.cdef_idx.ctypes.data
        s.cdef_stride = self.geom.cdef_stride
        return s

    def luma_pixels(self):
        return self.params.upscaled_width * self.params.frame_height


# Resolutions of the mixed multi-stream workload.
BATCH_SIZES = [(640, 360), (1280, 720), (1920, 1080), (2560, 1440), (3840, 2160)]


def batch_stream_configs(n_streams, frames_per_stream, seed):
    """Frame configurations of the mixed multi-stream workload: per stream a
    size drawn uniformly from BATCH_SIZES, half 8-bit and half 10-bit, 4:2:0
    except every eighth stream (4:4:4), a random qindex in [48, 200], SB 128,
    every filter on (restoration mixing Wiener / self-guided per 64-sample
    unit) with film grain (lag 3, 8 luma points), frames of a GOP with a
    1.5 pixel per frame pan. Returns [(bit depth, [Frame kwargs])]."""
    import random
    rng = random.Random(seed)
    streams = []
    for s in range(n_streams):
        w, h = BATCH_SIZES[rng.randrange(len(BATCH_SIZES))]
        bd = 8 if s % 2 == 0 else 10
        ss = 0 if s % 8 == 3 else 1
        q = rng.randint(48, 200)
        streams.append((bd, [dict(width=w, height=h, bit_depth=bd, ssx=ss, ssy=ss, sb_size=128, qindex=q,
                                  seed=seed * 1000 + s, frame_index=i, pan_px=1.5, lr=2, lr_unit_size=64,
                                  grain=True, grain_lag=3, grain_y_points=8, with_source=False)
                             for i in range(frames_per_stream)]))
    return streams


def make_frames(configs, workers=16):
    """Frames of many configurations, generated on a pool of threads (the
    generator runs in C without the interpreter lock)."""
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=max(1, workers)) as ex:
        return list(ex.map(lambda kw: Frame(**kw), configs))
