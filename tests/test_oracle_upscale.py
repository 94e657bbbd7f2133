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


def _tdiv(a, b):
    """Integer division truncating toward zero (the spec's "/", section 4.6)."""
    q = abs(a) // abs(b)
    return q if (a >= 0) == (b >= 0) else -q


def _spec_positions(dw, uw):
    """(integer source position, 1/64 phase) of every output column, from
    7.16's position arithmetic (SUPERRES_SCALE_BITS 14, EXTRA_BITS 8)."""
    step = _tdiv((dw << 14) + _tdiv(uw, 2), uw)
    err = uw * step - (dw << 14)
    init = _tdiv(-((uw - dw) << 13) + _tdiv(uw, 2), uw) + (1 << 7) - _tdiv(err, 2)
    init &= (1 << 14) - 1
    pos = [-(1 << 14) + init + x * step for x in range(uw)]
    return [p >> 14 for p in pos], [(p & ((1 << 14) - 1)) >> 8 for p in pos]


def _upscale_filter():
    """UPSCALE_FILTER of the library's tables (its rows hash-checked against
    the reference's src/tables.rs in test_tables.py)."""
    import os
    import re
    path = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "rivet-av1_amd", "csrc",
                        "av1_tables.h")
    text = open(path).read()
    body = text[text.index("AV1_UPSCALE_FILTER[64][8]"):]
    body = body[body.index("{") + 1:body.index("};")]
    vals = [int(v) for v in re.findall(r"-?\d+", body)]
    return np.array(vals, dtype=np.int64).reshape(64, 8)


@pytest.mark.parametrize("w,denom,ssx", [(226, 12, 1), (333, 9, 1), (200, 16, 0), (97, 13, 1), (640, 11, 0),
                                          (150, 10, 1), (777, 15, 0), (514, 14, 1)])
def test_phase_and_offset_of_every_column_by_impulse_response(w, denom, ssx):
    """Each output column's integer source position and 1/64 phase, read
    back from the oracle's impulse responses, equal 7.16's position
    arithmetic exactly, and the response is the filter row of that phase
    with the taps clamped at column 0 and at the mode-info edge.

    12-bit planes: a mid-grey background of 2048 with impulses of +1024
    make the output 2048 + 8 * tap exactly (taps sum to 128), never clipped.
    Row r holds impulses at the source columns congruent to r modulo 16, so
    each output column sees at most one impulse per row and all 16 rows
    together give its response to every source column."""
    F = _upscale_filter()
    f = workload.Frame(w, 32, bit_depth=12, ssx=ssx, ssy=1, qindex=100, lr=0, seed=7, superres_denom=denom,
                       with_source=False)
    planes = []
    for p in range(3):
        ph, pw = f.plane_shape(p)
        a = np.full((ph, pw), 2048, dtype=np.uint16)
        for r in range(ph):
            a[r, (r % 16)::16] += 1024
        planes.append(a)
    up = pyoracle.upscale(f, planes)
    for p in range(3):
        sx = ssx if p else 0
        dw, uw = (f.params.frame_width + sx) >> sx, (w + sx) >> sx
        max_x = (f.params.mi_cols >> sx) * 4 - 1
        px_spec, sub_spec = _spec_positions(dw, uw)
        # measured response of column x to source column j (16 rows cover all j mod 16)
        resp = np.zeros((uw, max_x + 1), dtype=np.int64)
        for r in range(16):
            d = up[p][r, :uw].astype(np.int64) - 2048
            assert np.all(d % 8 == 0)
            for x in range(uw):
                # the impulses of row r that column x can reach, one at most
                js = [j for j in range(r % 16, max_x + 1, 16) if abs(j - px_spec[x]) <= 8 or j in (0, max_x)]
                hit = [j for j in js if max(0, min(max_x, j)) == j and
                       any(max(0, min(max_x, px_spec[x] + k - 3)) == j for k in range(8))]
                assert len(hit) <= 1
                if hit:
                    resp[x, hit[0]] = d[x] // 8
                else:
                    assert d[x] == 0, (p, x, r)
        readout_checked = 0
        for x in range(uw):
            # the phase and offset the response itself shows (interior columns)
            nz = np.nonzero(resp[x])[0]
            if px_spec[x] - 3 >= 0 and px_spec[x] + 4 <= max_x:
                matches = []
                for sub in range(64):
                    row = F[sub]
                    k0 = int(np.nonzero(row)[0][0])
                    px = int(nz[0]) + 3 - k0
                    if 0 <= px - 3 and px + 4 <= max_x and np.array_equal(resp[x, px - 3:px + 5], row) and \
                            np.count_nonzero(resp[x]) == np.count_nonzero(row):
                        matches.append((px, sub))
                assert matches == [(px_spec[x], sub_spec[x])], (p, x, matches, px_spec[x], sub_spec[x])
                readout_checked += 1
            # every column, edges included: the clamped filter row of its spec position
            want = np.zeros(max_x + 1, dtype=np.int64)
            for k in range(8):
                want[max(0, min(max_x, px_spec[x] + k - 3))] += F[sub_spec[x]][k]
            assert np.array_equal(resp[x], want), (p, x)
        assert readout_checked > uw // 2
        # the positions are the centre-aligned ones to within 1/32 sample
        pos = np.array(px_spec) + np.array(sub_spec) / 64.0
        ideal = (np.arange(uw) + 0.5) * dw / uw - 0.5
        assert np.max(np.abs(pos - ideal)) < 1 / 32 + 1e-9, (p, np.max(np.abs(pos - ideal)))
