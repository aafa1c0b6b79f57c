"""The tree clusterizer oracle sums in the GPU's chunked order; the
reference sums in position order (crn_tree_clusterizer.h:265-570).  This
measures what the order changes:
  * where every float sum is exact (dyadic coordinates, small weights) the
    two orders give the same codebook and the same leaves -- so the order is
    the only difference between the restatement and the reference's loop;
  * on dxt_hc inputs the float sums round differently, the trees diverge
    (about half the blocks keep their endpoints and selectors at 2048^2 with
    4096 / 8192 entries) while the encoded quality stays the same: 35.9790
    vs 35.9784 dB there, and within 0.1 dB on the small case below (256^2,
    256 / 512 entries)."""
import numpy as np

import oracle
from crunch2_amd import dxt1, texture


def test_orders_agree_when_sums_are_exact():
    rng = np.random.default_rng(3)
    v = (rng.integers(0, 16, (6000, 6)) / 16.0).astype(np.float32)
    w = rng.integers(1, 4, 6000).astype(np.uint32)
    try:
        oracle.tree_position_order(False)
        a = oracle.tree_codebook(v, w, 64)
        oracle.tree_position_order(True)
        b = oracle.tree_codebook(v, w, 64)
    finally:
        oracle.tree_position_order(False)
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(a[1], b[1])


def psnr(blocks, r):
    ep, sel, ce, cs = r[0], r[1], r[2].astype(np.int64), r[3].astype(np.int64)
    e, s = ce[ep[:, 0]], cs[sel[:, 0]]

    def ex(c):
        r_, g, b_ = (c >> 11) & 31, (c >> 5) & 63, c & 31
        return np.stack([(r_ << 3) | (r_ >> 2), (g << 2) | (g >> 4), (b_ << 3) | (b_ >> 2)], -1)
    c0, c1 = ex(e & 0xFFFF), ex(e >> 16)
    pal = np.stack([c0, (2 * c0 + c1) // 3, (c0 + 2 * c1) // 3, c1], 1)
    idx = (s[:, None] >> (2 * np.arange(16))[None]) & 3
    d = np.take_along_axis(pal, idx[:, :, None].repeat(3, 2), 1) - blocks[:, :, :3].astype(np.int64)
    return 10 * np.log10(255.0 ** 2 / (d * d).mean())


def test_order_changes_trees_not_quality():
    blocks = dxt1.image_blocks(texture.synth(256, 256, 11, texture.RGB))
    levels = [(0, len(blocks), 64, 1.0)]
    try:
        oracle.tree_position_order(False)
        a = oracle.hc_compress(blocks, levels, codebook_sizes=(256, 512, 8, 8))
        oracle.tree_position_order(True)
        b = oracle.hc_compress(blocks, levels, codebook_sizes=(256, 512, 8, 8))
    finally:
        oracle.tree_position_order(False)
    same = (a[2][a[0][:, 0]] == b[2][b[0][:, 0]]).mean()
    print("blocks with equal endpoints %.3f, PSNR %.4f vs %.4f" % (same, psnr(blocks, a), psnr(blocks, b)))
    assert 0.2 < same < 1.0
    assert abs(psnr(blocks, a) - psnr(blocks, b)) < 0.1
