"""The 4:2:0 request derivation pinned by what the mathematics fixes.

dsmvc_hip.filter_requests follows the reference's prepare_filter_requests
(src/vs_plugin.cpp:656-705): horizontal chroma is shifted by
0.25 - 0.25 * dw / sw (+ src_left scaled), vertical chroma by src_top scaled.
These tests do not compare that formula with a copy of itself.  They check
the property the formula exists for, using the oracle's FP64 solver:

A linear ramp f(X) = X / sw over the source picture (X in source-luma pixel
units, sample i centred at i + 0.5) is sampled at each plane's own sample
positions -- luma at pixel centres, left-sited chroma at X = 2j + 0.5
horizontally, centre-sited chroma at Y = 2j + 1 vertically.  Bilinear
interpolation reproduces a ramp exactly, so the least-squares descale of
those samples must return the ramp at the destination's sample positions
(away from the borders, where the padding rule bends the ramp): destination
pixel k of the low-resolution image sits at X = (k + 0.5 - src_left) *
sw / src_width, and destination chroma j at X = (2j + 0.5 - src_left) *
sw / src_width.  A wrong sign, a dropped term or the wrong active length
moves every interior sample by a visible amount.
"""
import numpy as np
import pytest

import dsmvc_hip as dh
from support import BILINEAR, F64_ALWAYS, make_request

INNER = 32   # bilinear plans decay ~0.27 per sample: 32 samples from the edge is exact in FP64


def _oracle_plan(oracle, r: dh.AxisRequest):
    return oracle.build(make_request(r.source_size, r.destination_size, r.active_length,
                                     r.shift, BILINEAR, f64_mode=F64_ALWAYS))


CASES = [(1920, 1080, 1280, 720, 0.0, 0.0), (1920, 1080, 1440, 810, 0.0, 0.0),
         (3840, 2160, 2560, 1440, 0.0, 0.0), (1920, 1080, 1280, 720, 0.75, -0.5)]


@pytest.mark.parametrize("sw,sh,dw,dh_,left,top", CASES)
def test_ramp_survives_descale_at_sited_positions(oracle, sw, sh, dw, dh_, left, top):
    hreq, vreq = dh.filter_requests(sw, sh, dw, dh_, 3, 1, 1, left, top, kernel=dh.BILINEAR,
                                    f64_mode=2)
    assert len(hreq) == 2 and len(vreq) == 2
    # horizontal luma and left-sited chroma
    luma = oracle.axis_f64(_oracle_plan(oracle, hreq[0]), (np.arange(sw) + 0.5) / sw)
    want = (np.arange(dw) + 0.5 - left) * sw / dw / sw
    assert np.abs(luma - want)[INNER:-INNER].max() < 1e-9
    cw = sw // 2
    chroma = oracle.axis_f64(_oracle_plan(oracle, hreq[1]), (2 * np.arange(cw) + 0.5) / sw)
    want_c = (2 * np.arange(dw // 2) + 0.5 - left) * sw / dw / sw
    assert np.abs(chroma - want_c)[INNER:-INNER].max() < 1e-9
    # vertical centre-sited chroma
    ch = sh // 2
    vch = oracle.axis_f64(_oracle_plan(oracle, vreq[1]), (2 * np.arange(ch) + 1.0) / sh)
    want_v = (2 * np.arange(dh_ // 2) + 1.0 - top) * sh / dh_ / sh
    assert np.abs(vch - want_v)[INNER:-INNER].max() < 1e-9


def test_planted_siting_mistakes_fail(oracle):
    """The same check rejects the plausible mistakes: the horizontal chroma
    shift with its sign flipped, without its 0.25 - 0.25 * dw / sw term, and
    centre-sited (vertical-style) chroma horizontally."""
    sw, dw = 1920, 1280
    hreq, _ = dh.filter_requests(sw, 1080, dw, 720, 3, 1, 1, kernel=dh.BILINEAR, f64_mode=2)
    good = hreq[1]
    x = (2 * np.arange(sw // 2) + 0.5) / sw
    want = (2 * np.arange(dw // 2) + 0.5) * sw / dw / sw
    for shift in (-good.shift, 0.0, 0.25 * dw / sw):
        bad = dh.AxisRequest(good.source_size, good.destination_size, good.active_length, shift,
                             dh.BILINEAR, f64_mode=2)
        err = np.abs(oracle.axis_f64(_oracle_plan(oracle, bad), x) - want)[INNER:-INNER].max()
        assert err > 1e-4, shift


def test_server_and_test_restatements_agree_with_filter_requests():
    """The served requests' derivation (src_left = src_top = 0) and the
    test-side restatement in workloads.py both equal filter_requests."""
    from workloads import chroma420_requests
    for sw, sh, dw, dh_ in [(1920, 1080, 1280, 720), (3840, 2160, 2560, 1442)]:
        hreq, vreq = dh.filter_requests(sw, sh, dw, dh_, 3, 1, 1)
        ch, cv = chroma420_requests(sw, sh, dw, dh_, dh.BICUBIC)
        for mine, theirs in ((hreq[1], ch), (vreq[1], cv)):
            assert (mine.source_size, mine.destination_size) == (theirs.source_size,
                                                                 theirs.destination_size)
            assert mine.active_length == theirs.active_length and mine.shift == theirs.shift
