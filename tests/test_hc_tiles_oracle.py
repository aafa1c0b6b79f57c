"""Pins the adaptive tile restatement (oracle/hc_tiles.c; crn_dxt_hc.cpp:
318-561, crn_dxt_fast.cpp, crn_tree_clusterizer.h:573-753) with closed forms
and invariants: dxt_fast on solid and two-colour blocks, split_vectors on
two values and on separated groups, palettized endpoints of two-colour tiles,
and the tile bookkeeping (encodings cover each chunk, every block's tile
holds its texels, the backwards walk of odd row pairs, flat chunks pick the
whole 8x8 tile)."""
import numpy as np
import pytest

import oracle
from tests.hc_cases import blocks_from_texture

# tiles of each encoding and the texels (block index 0..3 = b, b+1, b+w, b+w+1) each tile covers
TILES = [[8], [6, 7], [4, 5], [6, 1, 3], [7, 0, 2], [4, 2, 3], [5, 0, 1], [0, 2, 1, 3]]
COVER = {0: {0}, 1: {2}, 2: {1}, 3: {3}, 4: {0, 2}, 5: {1, 3}, 6: {0, 1}, 7: {2, 3}, 8: {0, 1, 2, 3}}


def test_dxt_fast_solid_uses_single_colour_tables():
    om5, om6 = oracle.dxt1_omatch(0), oracle.dxt1_omatch(1)
    for r, g, b in ((0, 0, 0), (255, 255, 255), (13, 200, 77), (128, 64, 32)):
        px = np.tile(np.array([r, g, b, 255], np.uint8), (16, 1))
        lo, hi, sel = oracle.dxt_fast_color(px)
        a = int(om5[r, 0]) << 11 | int(om6[g, 0]) << 5 | int(om5[b, 0])
        c = int(om5[r, 1]) << 11 | int(om6[g, 1]) << 5 | int(om5[b, 1])
        assert (lo, hi) == (max(a, c), min(a, c))
        assert set(sel.tolist()) <= {2, 3}


def expand(c):
    r, g, b = c >> 11, (c >> 5) & 63, c & 31
    return np.array([(r << 3) | (r >> 2), (g << 2) | (g >> 4), (b << 3) | (b >> 2)])


@pytest.mark.parametrize("n", [16, 32, 64])
def test_dxt_fast_two_exact_colours_is_lossless(n):
    rng = np.random.default_rng(n)
    for _ in range(20):
        c0, c1 = rng.integers(0, 65536, 2)
        if c0 == c1:
            continue
        e0, e1 = expand(int(c0)), expand(int(c1))
        pick = rng.integers(0, 2, n)
        pick[0], pick[1] = 0, 1
        px = np.zeros((n, 4), np.uint8)
        px[:, :3] = np.where(pick[:, None] == 0, e0, e1)
        lo, hi, sel = oracle.dxt_fast_color(px)
        pal = [expand(lo), expand(hi), (2 * expand(lo) + expand(hi)) // 3, (expand(lo) + 2 * expand(hi)) // 3]
        rec = np.array([pal[s] for s in sel])
        if (e0 == e1).all():
            continue
        assert (rec == px[:, :3]).all()


def test_split_vectors_closed_forms():
    v = np.array([[0.2], [0.8]], np.float32)
    assert np.allclose(oracle.split_vectors(v, [3, 5]), [[0.2], [0.8]])
    one = oracle.split_vectors(np.array([[0.4, 0.1, 0.3]], np.float32), [7])
    assert np.allclose(one[0], one[1]) and np.allclose(one[0], [0.4, 0.1, 0.3])
    rng = np.random.default_rng(0)
    a = 0.1 + 0.01 * rng.random((10, 3))
    b = 0.7 + 0.01 * rng.random((12, 3))
    v = np.concatenate([a, b]).astype(np.float32)
    w = rng.integers(1, 9, 22)
    r = oracle.split_vectors(v, w)
    ma = (a * w[:10, None]).sum(0) / w[:10].sum()
    mb = (b * w[10:, None]).sum(0) / w[10:].sum()
    got = sorted([r[0], r[1]], key=lambda x: x[0])
    assert np.allclose(got[0], ma, atol=1e-6) and np.allclose(got[1], mb, atol=1e-6)


def level_list(w, h, faces=1):
    return [(0, w * h, w, 1.0)]


def test_flat_chunks_take_the_whole_tile():
    blocks = np.zeros((8 * 8, 16, 4), np.uint8)
    blocks[:, :, :3] = 90
    enc, ti, src, vec = oracle.hc_tiles(blocks, level_list(8, 8))
    assert (enc == 0).all()
    used = src != 0xFFFFFFFF
    assert (src[used] & 15 == 8).all() and used.sum() == 16
    assert np.allclose(vec[used][:, :3], vec[used][:, 3:6])      # one colour: both ends equal


@pytest.mark.parametrize("alpha", [(), (3,), (3, 0)])
def test_tile_bookkeeping(alpha):
    w, h = 16, 12
    blocks = blocks_from_texture(64, 5)[: w * h].copy()
    blocks[:, :, 3] = np.arange(16, dtype=np.uint8)[None, :] * 16
    has_color = len(alpha) != 2
    enc, ti, src, vec = oracle.hc_tiles(blocks, [(0, w * h, w, 1.5)], has_color=has_color, alpha_comp=alpha)
    for by in range(0, h, 2):
        for bx in range(0, w, 2):
            b = by * w + bx
            quad = [b, b + 1, b + w, b + w + 1]
            e = enc[b]
            assert (enc[quad] == e).all()
            slots = set(ti[quad].tolist())
            assert len(slots) == len(TILES[e])
            # every block's slot holds a tile of this chunk that covers the block
            for q, blk in enumerate(quad):
                s = src[ti[blk]]
                assert s != 0xFFFFFFFF and s >> 4 == b
                assert q in COVER[s & 15]
            # tile slots of a chunk: 4 consecutive, walked backwards on odd row pairs
            base = min(slots) & ~3
            pair = by // 2
            expect = by * w + (bx // 2) * 4 if pair % 2 == 0 else by * w + 2 * w - 4 - (bx // 2) * 4
            assert base == expect


def test_two_colour_tile_endpoints():
    """A chunk of two colours: every chosen tile's palettized endpoints are the
    two colours in the perceptual scaling, shorter vector first (or one colour
    twice when the tile holds only one)."""
    blocks = np.zeros((4, 16, 4), np.uint8)
    c0, c1 = np.array([40, 120, 200]), np.array([220, 30, 60])
    rng = np.random.default_rng(1)
    pick = rng.integers(0, 2, (4, 16))
    blocks[:, :, :3] = np.where(pick[:, :, None] == 0, c0, c1)
    enc, ti, src, vec = oracle.hc_tiles(blocks, [(0, 4, 2, 1.0)])
    scale = np.array([0.5, 1.0, 0.25], np.float32)
    v0 = (c0.astype(np.float32) / np.float32(255)) * scale
    v1 = (c1.astype(np.float32) / np.float32(255)) * scale
    lo, hi = (v0, v1) if np.linalg.norm(v0) <= np.linalg.norm(v1) else (v1, v0)
    for s in np.flatnonzero(src != 0xFFFFFFFF):
        assert np.allclose(vec[s, :3], lo, atol=1e-6) and np.allclose(vec[s, 3:6], hi, atol=1e-6)
