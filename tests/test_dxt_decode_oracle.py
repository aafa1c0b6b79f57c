"""The oracle's DXTn decode (oracle/dxt_decode.py) against hand-computed
values: palettes of extreme endpoints in both colour modes, the DXT5 8- and
6-value ramps, DXT3 nibble order, selector bit order, channel placement per
format, and the reference's DXT5 fall-through restated on request."""
import numpy as np

from oracle import dxt_decode as dd


def color_block(c0, c1, sel):
    return bytes([c0 & 255, c0 >> 8, c1 & 255, c1 >> 8, sel & 255, (sel >> 8) & 255, (sel >> 16) & 255, sel >> 24])


def alpha_block(l, h, sel48):
    return bytes([l, h] + [(sel48 >> (8 * k)) & 255 for k in range(6)])


def test_four_colour_palette_of_white_and_black():
    # c0 = white > c1 = black: 255, 0, (2*255 + 0) // 3 = 170, (255 + 0) // 3 = 85
    sel = 0b11100100  # texels 0..3 take selectors 0, 1, 2, 3
    t = dd.decode_blocks(color_block(0xFFFF, 0x0000, sel), dd.FMT_DXT1)[0]
    assert t[:4].tolist() == [[255, 255, 255, 255], [0, 0, 0, 255], [170, 170, 170, 255], [85, 85, 85, 255]]
    assert (t[4:] == [255, 255, 255, 255]).all()


def test_three_colour_palette_has_transparent_black():
    # c0 = 0 <= c1 = white: (0 + 255) >> 1 = 127 and (0, 0, 0, 0) for selector 3
    t = dd.decode_blocks(color_block(0x0000, 0xFFFF, 0b11100100), dd.FMT_DXT1)[0]
    assert t[:4].tolist() == [[0, 0, 0, 255], [255, 255, 255, 255], [127, 127, 127, 255], [0, 0, 0, 0]]


def test_565_expansion_and_channel_order():
    # pure red 5-bit 16 -> (16 << 3) | (16 >> 2) = 132; green 6-bit 33 -> 132 + 4 = (33 << 2) | (33 >> 4) = 134
    c0 = (16 << 11) | (33 << 5) | 1
    t = dd.decode_blocks(color_block(c0, c0, 0), dd.FMT_DXT1)[0]
    assert t[0].tolist() == [132, 134, 8, 255]


def test_dxt5_value_ramps():
    v = dd.alpha_values(np.array([255, 0]), np.array([0, 255]))
    assert v[0].tolist() == [255, 0, 218, 182, 145, 109, 72, 36]
    assert v[1].tolist() == [0, 255, 51, 102, 153, 204, 0, 255]


def test_selector_bit_order():
    # 3-bit selectors: texel t at bits 3t of the 48-bit field; 2-bit at 2t
    for t in (0, 5, 15):
        a = dd.decode_blocks(alpha_block(255, 0, 1 << (3 * t)), dd.FMT_DXT5A)[0][:, 3]
        assert a[t] == 0 and (np.delete(a, t) == 255).all()
        c = dd.decode_blocks(color_block(0xFFFF, 0, 1 << (2 * t)), dd.FMT_DXT1)[0][:, 0]
        assert c[t] == 0 and (np.delete(c, t) == 255).all()


def test_dxt3_nibbles():
    alpha = bytes([0x10, 0x32, 0x54, 0x76, 0x98, 0xBA, 0xDC, 0xFE])   # texel t holds nibble t
    t = dd.decode_blocks(alpha + color_block(0xFFFF, 0, 0), dd.FMT_DXT3)[0]
    assert t[:, 3].tolist() == [(k << 4) | k for k in range(16)]


def test_channel_placement():
    x, y = alpha_block(200, 200, 0), alpha_block(30, 30, 0)
    assert dd.decode_blocks(x + y, dd.FMT_DXN_XY)[0, 0].tolist() == [200, 30, 255, 255]
    assert dd.decode_blocks(x + y, dd.FMT_DXN_YX)[0, 0].tolist() == [30, 200, 255, 255]
    assert dd.decode_blocks(x, dd.FMT_DXT5A)[0, 0].tolist() == [255, 255, 255, 200]
    d5 = dd.decode_blocks(x + color_block(0xFFFF, 0, 0), dd.FMT_DXT5)[0, 0].tolist()
    assert d5 == [255, 255, 255, 200]


def test_reference_dxt5_fall_through():
    # the colour block read as an alpha block: l = 0xFF, h = 0xFF, selectors from bytes 2..7 = 0
    blk = alpha_block(200, 200, 0) + color_block(0xFFFF, 0, 0)
    ref = dd.decode_blocks(blk, dd.FMT_DXT5, reference_dxt5=True)[0, 0].tolist()
    assert ref == [255, 200, 255, 255]


def test_reference_dxt5_equals_dxn_yx_of_the_same_bytes():
    rng = np.random.default_rng(3)
    b = rng.integers(0, 256, (256, 16), dtype=np.uint8).tobytes()
    theirs = dd.decode_blocks(b, dd.FMT_DXT5, reference_dxt5=True)
    assert np.array_equal(theirs, dd.decode_blocks(b, dd.FMT_DXN_YX))
    assert not np.array_equal(dd.decode_blocks(b, dd.FMT_DXT5), theirs)
