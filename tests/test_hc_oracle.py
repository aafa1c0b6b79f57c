"""Pins the dxt_hc colour selector restatement (oracle/hc_oracle.c, which
follows the reference's E2/E4/E8 lookup tables) against a direct numpy sum of
per-texel distances over the whole block x codebook error matrix."""
import numpy as np
import pytest

import oracle
from tests.hc_cases import brute_force, make_case


@pytest.mark.parametrize("subblocks", [False, True])
@pytest.mark.parametrize("perceptual", [True, False])
def test_oracle_matches_brute_force(subblocks, perceptual):
    blocks, ei, cols, sel = make_case(300, 40, 700, seed=11 + subblocks + 2 * perceptual, subblocks=subblocks)
    best, err, det, used = oracle.hc_color_selectors(blocks, ei, cols, sel, subblocks, perceptual)
    b2, e2, d2, _ = brute_force(blocks, ei, cols, sel, subblocks, perceptual)
    assert np.array_equal(best, b2)
    assert np.array_equal(err, e2)
    assert np.array_equal(det, d2)
    assert np.array_equal(np.flatnonzero(used), np.unique(b2))


def test_ties_pick_lowest_index():
    """A codebook of identical selectors: every block must choose entry 0;
    a block of one colour equal to a palette entry has error 0."""
    blocks, ei, cols, sel = make_case(50, 5, 16, seed=3, dup=0.0)
    sel[:] = sel[0]
    best, err, *_ = oracle.hc_color_selectors(blocks, ei, cols, sel)
    assert (best == 0).all()
    blocks[:] = cols[ei, 2][:, None, :]
    sel[5] = 0xAAAAAAAA      # code 2 for every texel
    sel[9] = 0xAAAAAAAA
    best, err, *_ = oracle.hc_color_selectors(blocks, ei, cols, sel)
    assert (err == 0).all() and (best == 5).all()


def test_extreme_distances():
    """White texels against black palettes: the largest E2 (34 * 255^2 per
    texel with perceptual weights) sums exactly."""
    blocks = np.full((4, 16, 4), 255, np.uint8)
    cols = np.zeros((1, 4, 4), np.uint8)
    ei = np.zeros(4, np.uint32)
    sel = np.arange(8, dtype=np.uint32)
    best, err, det, _ = oracle.hc_color_selectors(blocks, ei, cols, sel)
    assert (err == 16 * 34 * 255 * 255).all() and (best == 0).all()
    assert det[0].sum() == 4 * 64 * 34 * 255 * 255


@pytest.mark.parametrize("scatter", [False, True])
def test_alpha_oracle_matches_brute_force(scatter):
    """The E3/E6 lookup restatement of the alpha selector task equals direct
    sums over the full error matrix; totals are taken against the second
    palette when one is given."""
    from tests.hc_cases import alpha_brute_force, make_alpha_case
    vals, ei, pal, pal2, sel = make_alpha_case(300, 30, 500, seed=5 + scatter)
    sv = pal2 if scatter else None
    best, err, det, used = oracle.hc_alpha_selectors(vals, ei, pal, sel, sv)
    b2, e2, d2 = alpha_brute_force(vals, ei, pal, sel, sv)
    assert np.array_equal(best, b2)
    assert np.array_equal(err, e2)
    assert np.array_equal(det, d2)
    assert np.array_equal(np.flatnonzero(used), np.unique(b2))


def test_alpha_extremes_and_ties():
    """255 against a palette of zeros: 16 * 255^2 per block; identical
    codebook entries: the first wins."""
    vals = np.full((3, 16), 255, np.uint8)
    pal = np.zeros((1, 8), np.uint8)
    sel = np.full(6, 0o7777777777777777, np.uint64)
    best, err, det, _ = oracle.hc_alpha_selectors(vals, np.zeros(3, np.uint32), pal, sel)
    assert (err == 16 * 255 * 255).all() and (best == 0).all()
    assert det[0].sum() == 3 * 128 * 255 * 255


@pytest.mark.parametrize("bits", [2, 3])
def test_selector_refine_is_an_argmin(bits):
    """Every refined code has the smallest total of its texel."""
    rng = np.random.default_rng(bits)
    k = 1 << bits
    det = rng.integers(0, 50, (400, 16 * k)).astype(np.uint32)     # small range: many ties
    out = oracle.hc_selector_refine(det, bits)
    codes = (out[:, None].astype(np.uint64) >> (bits * np.arange(16, dtype=np.uint64))[None, :]) & (k - 1)
    e = det.reshape(400, 16, k)
    chosen = np.take_along_axis(e, codes.astype(np.int64)[:, :, None], 2)[:, :, 0]
    assert np.array_equal(chosen, e.min(2))


def test_selector_refine_tie_order():
    """Ties follow the reference's tournament (crn_dxt_hc.cpp:1496-1500,
    1709-1716), which is not lowest-code-first: 2-bit, codes 1 and 3 tied
    lowest -> 3; 3-bit, codes 7 and 1 tied lowest -> 7, codes 0 and 7 -> 0,
    codes 4 and 2 -> 2, codes 3 and 6 -> 3; all equal -> 0."""
    d2 = np.array([[5, 1, 5, 1] * 16, [2] * 64], np.uint32)
    out2 = oracle.hc_selector_refine(d2, 2)
    assert out2[0] == 0xFFFFFFFF and out2[1] == 0

    def row(low):
        e = np.full(8, 9, np.uint32)
        e[list(low)] = 1
        return np.tile(e, 16)
    d3 = np.stack([row((7, 1)), row((0, 7)), row((4, 2)), row((3, 6)), np.full(128, 4, np.uint32)])
    out3 = oracle.hc_selector_refine(d3, 3)
    expect = [7, 0, 2, 3, 0]
    for o, c in zip(out3, expect):
        assert all(((int(o) >> (3 * p)) & 7) == c for p in range(16))
