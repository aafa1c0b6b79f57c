"""Decoding dxt_hc's output (palettes + per-block indices) back to texels,
the way a CRN transcoder turns them into DXT blocks: colour selectors and
alpha selectors are in linear order (code 0 = first endpoint .. last =
second endpoint), mapped to DXT order before the standard palettes apply."""
import numpy as np

DXT1_FROM_LINEAR = np.array([0, 2, 3, 1])
DXT5_FROM_LINEAR = np.array([0, 2, 3, 4, 5, 6, 7, 1])


def rgb565(c):
    r, g, b = (c >> 11) & 31, (c >> 5) & 63, c & 31
    return np.stack([(r << 3) | (r >> 2), (g << 2) | (g >> 4), (b << 3) | (b >> 2)], -1).astype(np.int64)


def dxt1_palette(lo, hi):
    c0, c1 = rgb565(lo), rgb565(hi)
    four = np.stack([c0, c1, (2 * c0 + c1) // 3, (c0 + 2 * c1) // 3], 1)
    three = np.stack([c0, c1, (c0 + c1) // 2, np.zeros_like(c0)], 1)
    return np.where((lo > hi)[:, None, None], four, three)


def dxt5_palette(lo, hi):
    lo = lo.astype(np.int64)
    hi = hi.astype(np.int64)
    eight = np.stack([lo, hi] + [((7 - i) * lo + i * hi) // 7 for i in range(1, 7)], 1)
    six = np.stack([lo, hi] + [((5 - i) * lo + i * hi) // 5 for i in range(1, 5)] +
                   [np.zeros_like(lo), np.full_like(lo, 255)], 1)
    return np.where((lo > hi)[:, None], eight, six)


def decode_color(ep, sel, color_endpoints, color_selectors):
    e = color_endpoints[ep[:, 0].astype(np.int64)].astype(np.int64)
    s = color_selectors[sel[:, 0].astype(np.int64)].astype(np.int64)
    pal = dxt1_palette(e & 0xFFFF, e >> 16)                                   # n, 4, 3
    codes = (s[:, None] >> (2 * np.arange(16))[None, :]) & 3
    raw = DXT1_FROM_LINEAR[codes]
    return np.take_along_axis(pal, raw[:, :, None].repeat(3, 2), 1)           # n, 16, 3


def decode_alpha(ep, sel, c, alpha_endpoints, alpha_selectors):
    e = alpha_endpoints[ep[:, 1 + c].astype(np.int64)].astype(np.int64)
    s = alpha_selectors[sel[:, 1 + c].astype(np.int64)].astype(np.uint64)
    pal = dxt5_palette(e & 0xFF, (e >> 8) & 0xFF)                              # n, 8
    codes = ((s[:, None] >> (3 * np.arange(16, dtype=np.uint64))[None, :]) & np.uint64(7)).astype(np.int64)
    return np.take_along_axis(pal, DXT5_FROM_LINEAR[codes], 1)                # n, 16


def psnr(a, b):
    mse = np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2)
    return 99.0 if mse == 0 else 10 * np.log10(255.0 ** 2 / mse)


def dxt1_texels(raw):
    """Standard DXT1 decode of a byte array of 8-byte blocks -> (n, 16, 3)."""
    b = np.frombuffer(bytes(raw), np.uint8).reshape(-1, 8).astype(np.int64)
    lo = b[:, 0] | (b[:, 1] << 8)
    hi = b[:, 2] | (b[:, 3] << 8)
    sel = b[:, 4] | (b[:, 5] << 8) | (b[:, 6] << 16) | (b[:, 7] << 24)
    pal = dxt1_palette(lo, hi)
    codes = (sel[:, None] >> (2 * np.arange(16))[None, :]) & 3
    return np.take_along_axis(pal, codes[:, :, None].repeat(3, 2), 1)


def dxt5a_values(raw8):
    """Decode of the 8-byte alpha half of DXT5 / DXN blocks -> (n, 16)."""
    b = np.frombuffer(bytes(raw8), np.uint8).reshape(-1, 8).astype(np.int64)
    lo, hi = b[:, 0], b[:, 1]
    bits = np.zeros(len(b), np.uint64)
    for k in range(6):
        bits |= b[:, 2 + k].astype(np.uint64) << np.uint64(8 * k)
    codes = ((bits[:, None] >> (3 * np.arange(16, dtype=np.uint64))[None, :]) & np.uint64(7)).astype(np.int64)
    return np.take_along_axis(dxt5_palette(lo, hi), codes, 1)
