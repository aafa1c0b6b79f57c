"""Pins the dxt_hc selector codebook composition in the oracle (merge, tree
VQ, quantize, assignment, refinement; crn_dxt_hc.cpp:1379-1504 and
1588-1720): the merge against a dictionary, the code layout on hand-made
words, and a closed form -- with room for every distinct selector and
selectors that are already each texel's best code, every block ends on an
entry whose codes are its best ones."""
import numpy as np
import pytest

import oracle
from tests.hc_cases import alpha_codebook_case, color_codebook_case


@pytest.mark.parametrize("bits", [2, 3])
def test_merge_matches_dictionary(bits):
    rng = np.random.default_rng(bits)
    wb = 32 if bits == 2 else 16
    sel = rng.integers(0, 40, 3000).astype(np.uint64) * np.uint64(0x9E3779B1 if bits == 2 else 0x9E3779B97F)
    sel &= np.uint64((1 << (16 * bits)) - 1)
    w = rng.integers(0, 1 << wb, 3000, dtype=np.uint64)
    vec, uw = oracle.hc_merge_selectors((sel << np.uint64(wb)) | w, bits)
    d = {}
    for s_, w_ in zip(sel.tolist(), w.tolist()):
        d[s_] = min(d.get(s_, 0) + w_, 2**32 - 1)
    keys = sorted(d)
    assert np.array_equal(uw, np.array([d[k] for k in keys], np.uint32))
    for row, k in zip(vec, keys):
        codes = [(k >> (bits * (15 - t))) & ((1 << bits) - 1) for t in range(16)]
        assert np.array_equal(row, (np.array(codes, np.float32) + 0.5) / (1 << bits))
    assert np.array_equal(oracle.hc_quantize_selectors(vec, bits).astype(np.uint64),
                          np.array([sum(int(c) << (bits * t) for t, c in
                                        enumerate([(k >> (bits * (15 - t))) & ((1 << bits) - 1) for t in range(16)]))
                                    for k in keys], np.uint64))


def test_color_closed_form():
    blocks, ei, cols, words, e2 = color_codebook_case(400, 12, seed=3)
    sel, idx, used = oracle.hc_color_selector_codebook(words, 4096, blocks, ei, cols)
    codes = (sel[idx][:, None].astype(np.uint64) >> (2 * np.arange(16, dtype=np.uint64))[None, :]) & np.uint64(3)
    err = np.take_along_axis(e2, codes.astype(np.int64)[:, :, None], 2)[:, :, 0].sum(1)
    assert np.array_equal(err, e2.min(2).sum(1))
    assert used[idx].all()


def test_alpha_closed_form():
    vals, ei, pal, pal2, words, e3 = alpha_codebook_case(300, 9, seed=4, channels=2)
    sel, idx, used = oracle.hc_alpha_selector_codebook(words, 4096, vals, ei, pal, pal)
    for c in range(2):
        codes = (sel[idx[c]][:, None] >> (3 * np.arange(16, dtype=np.uint64))[None, :]) & np.uint64(7)
        err = np.take_along_axis(e3[c], codes.astype(np.int64)[:, :, None], 2)[:, :, 0].sum(1)
        assert np.array_equal(err, e3[c].min(2).sum(1))
        assert used[idx[c]].all()


def test_small_codebooks_stay_within_size():
    blocks, ei, cols, words, _ = color_codebook_case(600, 20, seed=5, optimal=False)
    for k in (1, 3, 50):
        sel, idx, used = oracle.hc_color_selector_codebook(words, k, blocks, ei, cols)
        assert 1 <= len(sel) <= k and idx.max() < len(sel)
