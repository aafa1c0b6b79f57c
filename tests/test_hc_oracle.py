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
