"""CPU restatement of crnlib's DXTn block decode -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module, and only as the checker.

Follows crn_decompress_block (crnlib/crnlib.cpp:420-543) and the block helpers
it calls, per block, in plain integer numpy:
  colour palette ........ dxt1_block::get_block_colors (crn_dxt.cpp:291-300):
                          c0 > c1 -> get_block_colors4 ((2a + b) / 3,
                          (a + 2b) / 3, :247-260); else get_block_colors3
                          ((a + b) >> 1 and transparent black, :234-245);
                          endpoints expanded by unpack_color (:167-181)
  DXT3 alpha ............ dxt3_block::get_alpha scaled (crn_dxt.cpp:350-366)
  DXT5 alpha values ..... dxt5_block::get_block_values (crn_dxt.cpp:395-455)
  selectors ............. crn_dxt.h:161-165 (2 bits, row bytes) and 318-330
                          (3 bits, 48-bit little-endian field)
  channel placement ..... crnlib.cpp:432-531 (DXN_XY: X red, Y green; DXN_YX
                          swapped; DXT5A: alpha over white; swizzled DXT5
                          decoded as DXT5)
`reference_dxt5=True` restates the reference's DXT5 case exactly, including
its missing `break` (crnlib.cpp:485-503): after writing colour and alpha it
falls into the DXN code and overwrites every texel with (block 1 read as an
alpha block, block 0's alpha, 255, 255).  The product decode ends the case
where the alpha is written; tests compare it with `reference_dxt5=False` and
show that the two differ only in that case.

Parity pins (tests/test_dxt_decode_oracle.py): hand-computed palettes of
extreme endpoints (0 / 255 / 565 extremes) for both colour modes, the DXT5
8- and 6-value ramps of l = 255, h = 0 and the reverse, DXT3 nibble order,
and selector bit order on single-bit patterns.
"""
import numpy as np

FMT_DXT1, FMT_DXT3, FMT_DXT5 = 0, 1, 2
FMT_DXT5_CCXY, FMT_DXT5_XGXR, FMT_DXT5_XGBR, FMT_DXT5_AGBR = 3, 4, 5, 6
FMT_DXN_XY, FMT_DXN_YX, FMT_DXT5A = 7, 8, 9


def bytes_per_block(fmt):
    return 8 if fmt in (FMT_DXT1, FMT_DXT5A) else 16


def _expand(c):
    c = c.astype(np.int64)
    r, g, b = (c >> 11) & 31, (c >> 5) & 63, c & 31
    return np.stack([(r << 3) | (r >> 2), (g << 2) | (g >> 4), (b << 3) | (b >> 2)], -1)


def color_palette(c0, c1):
    """[n, 4, 4] RGBA palettes of colour blocks with endpoint words c0, c1."""
    e0, e1 = _expand(c0), _expand(c1)
    four = (c0 > c1)[:, None]
    p2 = np.where(four, (2 * e0 + e1) // 3, (e0 + e1) >> 1)
    p3 = np.where(four, (e0 + 2 * e1) // 3, 0)
    n = len(c0)
    pal = np.zeros((n, 4, 4), np.int64)
    pal[:, 0, :3], pal[:, 1, :3], pal[:, 2, :3], pal[:, 3, :3] = e0, e1, p2, p3
    pal[:, :3, 3] = 255
    pal[:, 3, 3] = np.where(four[:, 0], 255, 0)
    return pal


def alpha_values(l, h):
    """[n, 8] DXT5 alpha values of endpoint bytes l, h."""
    l, h = l.astype(np.int64), h.astype(np.int64)
    eight = (l > h)[:, None]
    v8 = np.stack([l, h] + [(l * (7 - k) + h * k) // 7 for k in range(1, 7)], 1)
    v6 = np.stack([l, h] + [(l * (5 - k) + h * k) // 5 for k in range(1, 5)] +
                  [np.zeros_like(l), np.full_like(l, 255)], 1)
    return np.where(eight, v8, v6)


def _color_texels(blk8):
    """blk8: [n, 8] uint8 colour blocks -> [n, 16, 4]."""
    w = blk8.astype(np.int64)
    c0 = w[:, 0] | (w[:, 1] << 8)
    c1 = w[:, 2] | (w[:, 3] << 8)
    sel = w[:, 4] | (w[:, 5] << 8) | (w[:, 6] << 16) | (w[:, 7] << 24)
    idx = (sel[:, None] >> (2 * np.arange(16))[None, :]) & 3
    pal = color_palette(c0, c1)
    return np.take_along_axis(pal, idx[:, :, None].repeat(4, 2), 1)


def _alpha_texels(blk8):
    """blk8: [n, 8] uint8 DXT5 alpha blocks -> [n, 16] values."""
    w = blk8.astype(np.int64)
    s = np.zeros(len(w), np.int64)
    for k in range(6):
        s |= w[:, 2 + k] << (8 * k)
    idx = (s[:, None] >> (3 * np.arange(16))[None, :]) & 7
    return np.take_along_axis(alpha_values(w[:, 0], w[:, 1]), idx, 1)


def decode_blocks(blocks, fmt, reference_dxt5=False):
    """blocks: bytes or uint8 array of n whole blocks -> [n, 16, 4] uint8 texels."""
    b = np.frombuffer(bytes(blocks), np.uint8).reshape(-1, bytes_per_block(fmt))
    n = len(b)
    out = np.zeros((n, 16, 4), np.int64)
    if fmt == FMT_DXT1:
        out = _color_texels(b)
    elif fmt == FMT_DXT3:
        out = _color_texels(b[:, 8:])
        a = b[:, :8].astype(np.int64)
        nib = np.stack([(a[:, t >> 1] >> (4 * (t & 1))) & 15 for t in range(16)], 1)
        out[:, :, 3] = (nib << 4) | nib
    elif fmt in (FMT_DXT5, FMT_DXT5_CCXY, FMT_DXT5_XGXR, FMT_DXT5_XGBR, FMT_DXT5_AGBR):
        out = _color_texels(b[:, 8:])
        out[:, :, 3] = _alpha_texels(b[:, :8])
        if reference_dxt5:
            # falls through into the DXN code with crn_fmt != DXN_XY
            v0, v1 = _alpha_texels(b[:, :8]), _alpha_texels(b[:, 8:])
            out = np.stack([v1, v0, np.full_like(v0, 255), np.full_like(v0, 255)], -1)
    elif fmt in (FMT_DXN_XY, FMT_DXN_YX):
        v0, v1 = _alpha_texels(b[:, :8]), _alpha_texels(b[:, 8:])
        x, y = (v0, v1) if fmt == FMT_DXN_XY else (v1, v0)
        out = np.stack([x, y, np.full_like(x, 255), np.full_like(x, 255)], -1)
    elif fmt == FMT_DXT5A:
        v = _alpha_texels(b)
        out = np.stack([np.full_like(v, 255)] * 3 + [v], -1)
    else:
        raise ValueError("format %d" % fmt)
    return out.astype(np.uint8)


def decode_image(blocks, fmt, width, height):
    """A level's blocks (raster order, minimal pitch) -> [height, width, 4]."""
    bx, by = (width + 3) // 4, (height + 3) // 4
    t = decode_blocks(blocks, fmt).reshape(by, bx, 4, 4, 4).transpose(0, 2, 1, 3, 4).reshape(by * 4, bx * 4, 4)
    return np.ascontiguousarray(t[:height, :width])
