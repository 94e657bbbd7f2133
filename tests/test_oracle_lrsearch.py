"""The oracle's per-unit restoration choice (orc_lr_search_unit, the unit
closure of search(), src/encoder/lr.rs:95-186) checked against its
definition by exhaustive evaluation.

For every unit the candidates are: no restoration (cost = SSE of the CDEF
output + lambda), the fitted Wiener filter (SSE + 40 lambda for luma, 28
lambda for chroma) and the fitted self-guided filter of each set tried
(SSE + 20 lambda), in that order; the choice is the first strict minimum
(lr.rs:154-186: `if cost < best.0`). The SSEs come from orc_unit_sse, the
exact judge, whose filters are the decoder's (pinned by the LR tests and by
the GPU judge's equality with the HIP integer filter). The Wiener taps and
self-guided weights are the oracle's own fits (orc_fit_wiener/orc_fit_sgr,
pinned by stationarity on the GPU); here the squared error fit_sgr reports
is also checked against the judge.

The lambdas tried include, per unit, the critical values where two
candidates cost the same, nudged to either side, so a wrong rate constant,
a swapped luma/chroma rate, a non-strict comparison or a changed candidate
order changes at least one choice. A unit whose source equals its CDEF
output makes every filter tie with "none" at lambda 0, which only the
strict comparison resolves to "none".
"""
import numpy as np
import pytest

import pyoracle
from rvpf import lrdiff, workload

CASES = [
    dict(width=200, height=136, bit_depth=10, qindex=160, lr=2, seed=5),
    dict(width=176, height=120, bit_depth=8, qindex=96, lr=2, seed=31),
    dict(width=130, height=90, bit_depth=12, qindex=200, lr=2, seed=32, ssx=0, ssy=0, lr_unit_size=64),
    dict(width=150, height=100, bit_depth=10, qindex=120, lr=2, seed=33, mono=True, lr_unit_size=128),
]

NONE, WIENER, SGR = 0, 1, 2


def unit_bytes(kind, wiener=None, set_=0, xqd=(0, 0)):
    """An rvpf_lr_unit (12 bytes: type, sgr_set, sgr_xqd[2], wiener[2][3], pad[2])."""
    b = np.zeros(12, dtype=np.uint8)
    b[0] = kind
    if kind == WIENER:
        b[4:10] = np.asarray(wiener, dtype=np.int8).reshape(6).view(np.uint8)
    elif kind == SGR:
        b[1] = set_
        b[2:4] = np.asarray(xqd, dtype=np.int8).view(np.uint8)
    return b


def candidates(h, plane, ur, uc, sets):
    """[(sse, rate, unit bytes)] in the order search() tries them."""
    out = [(h.unit_sse(plane, ur, uc, None), 1.0, unit_bytes(NONE))]
    q, _ = h.fit_wiener(plane, ur, uc)
    wb = unit_bytes(WIENER, q)
    out.append((h.unit_sse(plane, ur, uc, wb), 40.0 if plane == 0 else 28.0, wb))
    for s in sets:
        xqd, _, e = h.fit_sgr(plane, ur, uc, s)
        sb = unit_bytes(SGR, set_=s, xqd=xqd)
        assert e == h.unit_sse(plane, ur, uc, sb), ("fit_sgr's error is not the judge's", plane, ur, uc, s)
        out.append((e, 20.0, sb))
    return out


def first_strict_min(cands, lam):
    best = 0
    best_cost = float(cands[0][0]) + lam * cands[0][1]
    for i in range(1, len(cands)):
        c = float(cands[i][0]) + lam * cands[i][1]
        if c < best_cost:
            best, best_cost = i, c
    return best, best_cost


def critical_lambdas(cands):
    """lambdas at which two candidates cost the same, and either side of them."""
    lams = {0.0, 1e12}
    for i in range(len(cands)):
        for j in range(i + 1, len(cands)):
            dr = cands[j][1] - cands[i][1]
            if dr == 0:
                continue
            lam = (float(cands[i][0]) - float(cands[j][0])) / dr
            if lam > 0:
                lams.update((lam * (1 - 1e-9), lam, lam * (1 + 1e-9)))
    return sorted(lams)


def host_frame(kw, src_is_cdef=False):
    f = workload.Frame(**kw)
    deb = pyoracle.loop_filter(f, [None if a is None else a.copy() for a in f.rec], threads=4)
    cd = pyoracle.cdef(f, deb, threads=4)
    src = [None if a is None else a.copy() for a in cd] if src_is_cdef else f.src
    return f, pyoracle.LrSearchFrame(f, deb, cd, src)


@pytest.mark.parametrize("kw", CASES, ids=["c%d" % i for i in range(len(CASES))])
def test_search_unit_follows_its_definition(kw):
    f, h = host_frame(kw)
    sets = lrdiff.SGR_SETS_NORMAL
    units, _ = lrdiff.unit_layout(f.params)
    kinds_seen = set()
    for p, ur, uc in units:
        p, ur, uc = int(p), int(ur), int(uc)
        cands = candidates(h, p, ur, uc, sets)
        for lam in critical_lambdas(cands):
            got, cost = h.search_unit(p, ur, uc, lam, sets)
            k, want_cost = first_strict_min(cands, lam)
            assert np.array_equal(got, cands[k][2]), (p, ur, uc, lam, got, cands[k][2])
            assert cost == want_cost, (p, ur, uc, lam, cost, want_cost)
            kinds_seen.add(int(got[0]))
    # the lambdas reach "none" and Wiener in every case, and a self-guided
    # choice where one is ever the cheapest (the 8-bit case)
    want = {NONE, WIENER, SGR} if kw["bit_depth"] == 8 else {NONE, WIENER}
    assert want <= kinds_seen, kinds_seen


@pytest.mark.parametrize("kw", CASES[:2], ids=["c0", "c1"])
def test_ties_keep_the_earlier_candidate(kw):
    """Source = CDEF output: "none" and the Wiener fit (zero taps, the exact
    identity) both have error 0, so at lambda 0 every unit ties and must
    keep "none"; any self-guided candidate with error 0 ties too."""
    f, h = host_frame(kw, src_is_cdef=True)
    units, _ = lrdiff.unit_layout(f.params)
    for p, ur, uc in units:
        p, ur, uc = int(p), int(ur), int(uc)
        cands = candidates(h, p, ur, uc, lrdiff.SGR_SETS_NORMAL)
        assert cands[0][0] == 0 and cands[1][0] == 0, (p, ur, uc)
        got, cost = h.search_unit(p, ur, uc, 0.0, lrdiff.SGR_SETS_NORMAL)
        assert got[0] == NONE and cost == 0.0, (p, ur, uc, got)


def test_sets_and_rates_of_the_reference():
    """The sets tried per effort and the rates are lr.rs:40-41 and :154-176."""
    assert lrdiff.SGR_SETS_NORMAL == (0, 3, 6, 10, 13, 14)
    assert lrdiff.SGR_SETS_FAST == (3, 10)
    assert (lrdiff.RATE_NONE, lrdiff.RATE_WIENER_LUMA, lrdiff.RATE_WIENER_CHROMA, lrdiff.RATE_SGR) == \
        (1.0, 40.0, 28.0, 20.0)
