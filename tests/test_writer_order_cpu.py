"""Palette ordering in the CRN writer (crnlib's crn_comp::optimize_color /
optimize_alpha and the selector orders, crn_comp.cpp:767-1352): the reordered
file transcodes (oracle) to exactly the same DXTn bytes as the file written
in the given order -- the order is a bijection applied to palettes and
indices alike -- and it is smaller, for colour, DXT5 and two-channel alpha."""
import numpy as np
import pytest

import oracle
from crunch2_amd import texture
from tests.test_hc_roundtrip_oracle import FORMATS
from tests.hc_compress_cases import texture_blocks


@pytest.mark.parametrize("fmt_name", ["dxt1", "dxt5", "dxn"])
def test_ordered_palettes_decode_the_same_and_pack_smaller(fmt_name):
    fmt, kind, alpha, has_color = FORMATS[fmt_name]
    size, nlevels = 128, 4
    blocks, levels = texture_blocks(size, 5, kind, nlevels)
    ep, sel, ce, cs, ae, as_ = oracle.hc_compress(blocks, levels, has_color=has_color, alpha_comp=alpha,
                                                  codebook_sizes=(256, 256, 128, 128))
    kw = {}
    if has_color:
        kw.update(color_endpoints=ce, color_selectors=cs, ce_idx=ep[:, 0], cs_idx=sel[:, 0])
    if alpha:
        kw.update(alpha_endpoints=ae, alpha_selectors=as_, a0e_idx=ep[:, 1], a0s_idx=sel[:, 1])
    if len(alpha) == 2:
        kw.update(a1e_idx=ep[:, 2], a1s_idx=sel[:, 2])
    plain = texture.pack(size, size, nlevels, 1, fmt, **kw)
    ordered = texture.pack(size, size, nlevels, 1, fmt, order_palettes=True, **kw)
    a, b = oracle.unpack_all(plain), oracle.unpack_all(ordered)
    for l in range(nlevels):
        assert np.array_equal(a[l][0], b[l][0]), l
    print(fmt_name, len(plain), "->", len(ordered), flush=True)
    assert len(ordered) < len(plain)


def test_selector_chain_order_on_hand_made_words():
    # four colour selector words: the chain starts at the word nearest the
    # all-zero word and walks to the nearest remaining word each step
    ce = np.array([0xFFFF0000], np.uint32)
    cs = np.array([0x55555555, 0x00000000, 0xFFFFFFFF, 0x00000001], np.uint32)
    # one 8x8 level (2x2 blocks), every block on endpoint 0, selectors 0..3
    data = texture.pack(8, 8, 1, 1, 0, color_endpoints=ce, color_selectors=cs, ce_idx=np.zeros(4, np.uint32),
                        cs_idx=np.arange(4, dtype=np.uint32), order_palettes=True)
    pal = oracle.Context(data).palette(1)
    # the zero word first, then 0x00000001 (one texel differs by code 1),
    # then 0x55555555, then 0xFFFFFFFF; palettes decode to raw DXT codes, so
    # compare through a second file written in that order without reordering
    ref = texture.pack(8, 8, 1, 1, 0, color_endpoints=ce,
                       color_selectors=np.array([0, 1, 0x55555555, 0xFFFFFFFF], np.uint32),
                       ce_idx=np.zeros(4, np.uint32), cs_idx=np.array([2, 0, 3, 1], np.uint32))
    assert np.array_equal(pal, oracle.Context(ref).palette(1))
    assert np.array_equal(oracle.unpack_all(data)[0][0], oracle.unpack_all(ref)[0][0])
