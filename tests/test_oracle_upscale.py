"""The oracle's super-resolution upscaling (7.16, orc_upscale) pinned by
properties the upscaling process fixes independently of its formulas:
constant planes stay constant (every filter phase sums to 128), a linear
ramp is reproduced at the centre-aligned source position x' = (x + 1/2)
W/U - 1/2 (so a phase, offset or step error shows), samples past the
mode-info width never matter, and only the upscaled rows and columns are
written."""
import numpy as np
import pytest

import pyoracle
from rvpf import workload

CASES = [(226, 130, 10, 1, 1, 12), (200, 64, 8, 0, 0, 16), (333, 90, 12, 1, 0, 9), (96, 40, 10, 1, 1, 13)]


def frame(w, h, bd, ssx, ssy, denom):
    return workload.Frame(w, h, bit_depth=bd, ssx=ssx, ssy=ssy, qindex=100, lr=2, seed=w + denom,
                          superres_denom=denom, with_source=False)


@pytest.mark.parametrize("w,h,bd,ssx,ssy,denom", CASES)
def test_geometry(w, h, bd, ssx, ssy, denom):
    f = frame(w, h, bd, ssx, ssy, denom)
    p = f.params
    assert p.use_superres == 1 and p.upscaled_width == w
    assert p.frame_width == (w * 8 + denom // 2) // denom
    for pl in range(3):
        assert f.ref_shape(pl)[1] >= f.crop_shape(pl)[1] + 32


@pytest.mark.parametrize("w,h,bd,ssx,ssy,denom", CASES)
def test_constant_plane(w, h, bd, ssx, ssy, denom):
    f = frame(w, h, bd, ssx, ssy, denom)
    for v in (0, 1, (1 << bd) - 1, 3 << (bd - 4)):
        up = pyoracle.upscale(f, [np.full(f.plane_shape(p), v, dtype=np.uint16) for p in range(3)])
        for p in range(3):
            ch, cw = f.crop_shape(p)
            assert np.all(up[p][:ch, :cw] == v)


@pytest.mark.parametrize("w,h,bd,ssx,ssy,denom", CASES)
def test_ramp_is_reproduced_at_the_centre_aligned_position(w, h, bd, ssx, ssy, denom):
    f = frame(w, h, bd, ssx, ssy, denom)
    slope, base = (1 if bd == 8 else 4), 16
    maxv = (1 << bd) - 1
    planes = []
    for p in range(3):
        ph, pw = f.plane_shape(p)
        planes.append(np.tile(np.minimum(base + slope * np.arange(pw), maxv).astype(np.uint16), (ph, 1)))
    up = pyoracle.upscale(f, planes)
    for p in range(3):
        sx = ssx if p else 0
        dw = (f.params.frame_width + sx) >> sx
        uw = (w + sx) >> sx
        max_x = (f.params.mi_cols >> sx) * 4 - 1
        x = np.arange(uw)
        src = (x + 0.5) * dw / uw - 0.5
        # 8-tap support away from the clamped edges and below the clipped top of the ramp
        inside = (src >= 4) & (src <= max_x - 5) & (src + 5 < (maxv - base) / slope)
        want = base + slope * src
        got = up[p][0, :uw].astype(np.float64)
        # 1/64-sample phases and the 14-bit step: within 1/2 sample of position, plus rounding
        assert inside.sum() > 20
        assert np.max(np.abs(got - want)[inside]) <= slope * 0.5 + 1, (p, np.max(np.abs(got - want)[inside]))
        # a one-sample shift would be off by the slope everywhere
        assert np.mean(np.abs(got - want)[inside]) < max(slope * 0.25, 0.5)


@pytest.mark.parametrize("w,denom", [(3840, 9), (3840, 11), (2000, 13), (1000, 10), (777, 15), (3840, 16)])
def test_positions_are_unbiased_on_a_steep_ramp(w, denom):
    """On a 12-bit sawtooth of slope 16 the output's mean signed error
    against the ramp at the centre-aligned position measures the average
    position error in 1/16 sample: it stays below 1/32 sample (a dropped
    step-error correction or a 1/32-sample shift of initial_subpel_x moves
    it past that), and every sample is within a sample of its position."""
    slope, period, base = 16, 240, 16
    f = workload.Frame(w, 8, bit_depth=12, ssx=1, ssy=1, qindex=100, lr=0, seed=1, superres_denom=denom,
                       with_source=False)
    planes = []
    for p in range(3):
        ph, pw = f.plane_shape(p)
        planes.append(np.tile((base + slope * (np.arange(pw) % period)).astype(np.uint16), (ph, 1)))
    up = pyoracle.upscale(f, planes)
    for p in range(3):
        sx = 1 if p else 0
        dw, uw = (f.params.frame_width + sx) >> sx, (w + sx) >> sx
        max_x = (f.params.mi_cols >> sx) * 4 - 1
        pos = (np.arange(uw) + 0.5) * dw / uw - 0.5
        saw = np.mod(pos, period)
        inside = (saw >= 4) & (saw <= period - 5) & (pos >= 4) & (pos <= max_x - 5)
        err = up[p][0, :uw].astype(np.float64) - (base + slope * saw)
        assert inside.sum() > 300
        assert abs(err[inside].mean()) < slope / 32, (p, err[inside].mean())
        assert np.abs(err[inside]).max() < slope, (p, np.abs(err[inside]).max())


@pytest.mark.parametrize("w,h,bd,ssx,ssy,denom", CASES)
def test_reads_stop_at_the_mode_info_width_and_writes_at_the_upscaled_size(w, h, bd, ssx, ssy, denom):
    f = frame(w, h, bd, ssx, ssy, denom)
    rng = np.random.default_rng(w)
    planes = [rng.integers(0, 1 << bd, size=f.plane_shape(p), dtype=np.uint16) for p in range(3)]
    up1 = pyoracle.upscale(f, planes)
    for p in range(3):
        sx = ssx if p else 0
        max_x = (f.params.mi_cols >> sx) * 4 - 1
        planes[p][:, max_x + 1:] = rng.integers(0, 1 << bd, size=planes[p][:, max_x + 1:].shape, dtype=np.uint16)
    up2 = pyoracle.upscale(f, planes)
    for p in range(3):
        ch, cw = f.crop_shape(p)
        assert np.array_equal(up1[p], up2[p])
        assert not np.any(up1[p][ch:, :]) and not np.any(up1[p][:, cw:])


def test_chain_with_only_upscaling():
    """Deblocking, CDEF, restoration and grain off: the reference frame is
    the upscaled reconstruction."""
    f = workload.Frame(226, 130, bit_depth=10, qindex=100, lf=False, cdef=False, lr=0, seed=5, superres_denom=11,
                       with_source=False)
    deb, ref, disp = pyoracle.postfilter_frame(f, threads=2)
    up = pyoracle.upscale(f, f.rec)
    for p in range(3):
        ch, cw = f.crop_shape(p)
        assert np.array_equal(ref[p][:ch, :cw], up[p][:ch, :cw])
        assert np.array_equal(disp[p], up[p][:ch, :cw])
