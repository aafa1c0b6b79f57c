"""Seeded synthetic textures and the CRN writer (host side of libcrn2).

The recipe and writer live in crunch2_amd/csrc/synth.cpp and crn_pack.cpp;
this module wraps them for the tests and the benchmark.
"""
import ctypes as C
import math

import numpy as np

from ._lib import lib, check, PackInput, FMT_DXT1, FMT_DXT5, FMT_DXN_XY, FMT_DXN_YX, FMT_DXT5A

RGB, RGBA, NORMAL = 0, 1, 2

# palette sizes crnlib derives for quality 128 on a texture of >= 8192 blocks
# (crnlib/crn_comp.cpp:532-577): colour endpoints/selectors, alpha endpoints/selectors
Q128_DXT1 = (2413, 2695, 0, 0)
Q128_DXT5 = (3265, 3549, 1944, 2662)
Q128_ALPHA = (0, 0, 1944, 2662)


def synth(w, h, seed, kind=RGB):
    out = np.zeros((h, w, 4), np.uint8)
    check(lib().crn2_synth_rgba(w, h, seed, kind, out.ctypes.data_as(C.c_void_p)), "crn2_synth_rgba")
    return out


def downsample(img):
    h, w = img.shape[:2]
    out = np.zeros((max(1, h >> 1), max(1, w >> 1), 4), np.uint8)
    src = np.ascontiguousarray(img)
    check(lib().crn2_downsample(src.ctypes.data_as(C.c_void_p), w, h, out.ctypes.data_as(C.c_void_p)))
    return out


def mip_chain(img, levels=None):
    h, w = img.shape[:2]
    n = levels or (int(math.floor(math.log2(max(w, h)))) + 1)
    chain = [np.ascontiguousarray(img)]
    for _ in range(n - 1):
        chain.append(downsample(chain[-1]))
    return chain


def compress(faces_levels, fmt, palette_sizes):
    """faces_levels[f][l] = RGBA8 array of level l of face f; returns .crn bytes."""
    faces = len(faces_levels)
    nl = len(faces_levels[0])
    h, w = faces_levels[0][0].shape[:2]
    arr = (C.c_void_p * (faces * nl))()
    keep = []
    for f in range(faces):
        for l in range(nl):
            a = np.ascontiguousarray(faces_levels[f][l], dtype=np.uint8)
            keep.append(a)
            arr[f * nl + l] = a.ctypes.data
    ps = (C.c_uint32 * 4)(*palette_sizes)
    out = C.c_void_p()
    size = C.c_uint32()
    check(lib().crn2_compress_rgba(arr, w, h, nl, faces, fmt, ps, C.byref(out), C.byref(size)), "crn2_compress_rgba")
    data = C.string_at(out, size.value)
    lib().crn2_free(out)
    return data


def pack(width, height, levels, faces, fmt, color_endpoints=None, color_selectors=None, alpha_endpoints=None,
         alpha_selectors=None, ce_idx=None, cs_idx=None, a0e_idx=None, a0s_idx=None, a1e_idx=None, a1s_idx=None,
         order_palettes=False):
    """Writes a .crn from explicit palettes and per-block indices (blocks of the
    even-padded grid of every level, faces stacked, rows in raster order).
    Colour selectors are linear 2-bit words, alpha selectors linear 3-bit words.
    With order_palettes the palettes are reordered as crnlib's compressor
    orders them before packing (the decoded texture is the same)."""
    keep = []

    def ptr(a, dt):
        if a is None:
            return None, 0
        a = np.ascontiguousarray(a, dtype=dt)
        keep.append(a)
        return a.ctypes.data, a.size

    p = PackInput()
    p.width, p.height, p.levels, p.faces, p.format = width, height, levels, faces, fmt
    p.color_endpoints, p.n_color_endpoints = ptr(color_endpoints, np.uint32)
    p.color_selectors_linear, p.n_color_selectors = ptr(color_selectors, np.uint32)
    p.alpha_endpoints, p.n_alpha_endpoints = ptr(alpha_endpoints, np.uint32)
    p.alpha_selectors_linear, p.n_alpha_selectors = ptr(alpha_selectors, np.uint64)
    p.color_endpoint_idx, _ = ptr(ce_idx, np.uint32)
    p.color_selector_idx, _ = ptr(cs_idx, np.uint32)
    p.alpha0_endpoint_idx, _ = ptr(a0e_idx, np.uint32)
    p.alpha0_selector_idx, _ = ptr(a0s_idx, np.uint32)
    p.alpha1_endpoint_idx, _ = ptr(a1e_idx, np.uint32)
    p.alpha1_selector_idx, _ = ptr(a1s_idx, np.uint32)
    p.order_palettes = int(bool(order_palettes))
    out = C.c_void_p()
    size = C.c_uint32()
    rc = lib().crn2_pack(C.byref(p), C.byref(out), C.byref(size))
    if rc != 0:
        raise RuntimeError("crn2_pack failed: %d" % rc)
    data = C.string_at(out, size.value)
    lib().crn2_free(out)
    return data


def padded_grid(width, height, levels):
    """(gw, gh) of the even-padded block grid of each level."""
    out = []
    for l in range(levels):
        w, h = max(1, width >> l), max(1, height >> l)
        bx, by = (w + 3) // 4, (h + 3) // 4
        out.append(((bx + 1) & ~1, (by + 1) & ~1))
    return out


def compress_hc(faces_levels, fmt, palette_sizes, perceptual=True, tile_params=(2.0, 2.0, 3.0)):
    """faces_levels[f][l] = RGBA8 array of level l of face f; returns .crn
    bytes from the GPU dxt_hc (crn2_compress_crn: the clustered path of
    crn_compress) and the project's CRN writer.  Needs a GPU."""
    faces = len(faces_levels)
    nl = len(faces_levels[0])
    h, w = faces_levels[0][0].shape[:2]
    arr = (C.c_void_p * (faces * nl))()
    keep = []
    for f in range(faces):
        for l in range(nl):
            a = np.ascontiguousarray(faces_levels[f][l], dtype=np.uint8)
            keep.append(a)
            arr[f * nl + l] = a.ctypes.data
    cb = (C.c_uint32 * 4)(*[max(8, int(v)) for v in palette_sizes])
    tp = (C.c_float * 3)(*tile_params)
    out = C.c_void_p()
    size = C.c_uint32()
    check(lib().crn2_compress_crn(arr, w, h, nl, faces, fmt, int(bool(perceptual)), cb, tp, 0, 0, C.byref(out),
                                  C.byref(size)), "crn2_compress_crn")
    data = C.string_at(out, size.value)
    lib().crn2_free(out)
    return data


def q128_palette_sizes(width, height, fmt):
    """The palette sizes crnlib's compressor derives for quality 128
    (crn_comp::quantize_images, crnlib/crn_comp.cpp:541-577)."""
    entries = min(max(((width + 3) // 4) * ((height + 3) // 4), 8), 8192)
    q = np.float32(128.0 / 255.0)
    cmul = np.float32(0.75) if fmt == FMT_DXT5 else np.float32(1.0)

    def size(lo, e):
        qq = np.float32(q ** np.float32(e))
        v = np.float32(0.5) + (np.float32(lo) + (np.float32(entries) - np.float32(lo)) * qq)
        return int(min(max(int(v), 8), 8192))
    return (size(64, np.float32(1.8) * cmul), size(96, np.float32(1.65) * cmul), size(24, np.float32(2.1)),
            size(48, np.float32(1.65)))


def synth_crn(size, seed, fmt, palette_sizes=None, faces=1, kind=None, writer="simple"):
    """Seeded texture (full mip chain) compressed to .crn in `fmt`: by the
    simple block quantizer (writer "simple", CPU) or by the GPU dxt_hc
    (writer "hc", palette sizes of crnlib's quality 128 unless given)."""
    if kind is None:
        kind = {FMT_DXT1: RGB, FMT_DXT5: RGBA, FMT_DXT5A: RGBA, FMT_DXN_XY: NORMAL, FMT_DXN_YX: NORMAL}[fmt]
    w, h = size if isinstance(size, tuple) else (size, size)
    fl = [mip_chain(synth(w, h, seed + 7919 * f, kind)) for f in range(faces)]
    if writer == "hc":
        return compress_hc(fl, fmt, palette_sizes or q128_palette_sizes(w, h, fmt))
    if palette_sizes is None:
        palette_sizes = {FMT_DXT1: Q128_DXT1, FMT_DXT5: Q128_DXT5}.get(fmt, Q128_ALPHA)
    return compress(fl, fmt, palette_sizes)


def mixed_set(n=512, seed=2024, workers=None, writer="simple"):
    """The serving-style texture set: n seeded textures with sides
    log-uniform in [64, 2048] and full mip chains, 60 % RGB -> DXT1, 25 % RGBA
    -> DXT5, 15 % normal maps -> DXN_XY, each packed to .crn by the project's
    writer.  Returns [(format, side, crn bytes)] in a fixed order; the files
    are built on `workers` threads (the work is in libcrn2, outside the GIL)."""
    import os
    from concurrent.futures import ThreadPoolExecutor
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        side = int(round(2 ** rng.uniform(6, 11)))
        u = rng.random()
        fmt = FMT_DXT1 if u < 0.60 else (FMT_DXT5 if u < 0.85 else FMT_DXN_XY)
        jobs.append((fmt, side, 5000 + i))
    if writer == "hc":
        # mip chains on CPU threads, the GPU dxt_hc and writer on the calling thread in order
        kinds = {FMT_DXT1: RGB, FMT_DXT5: RGBA, FMT_DXN_XY: NORMAL}
        workers = workers or min(16, os.cpu_count() or 1)
        with ThreadPoolExecutor(max_workers=workers) as pool:
            chains = pool.map(lambda j: mip_chain(synth(j[1], j[1], j[2], kinds[j[0]])), jobs)
            data = [compress_hc([ch], j[0], q128_palette_sizes(j[1], j[1], j[0])) for j, ch in zip(jobs, chains)]
    else:
        workers = workers or min(16, os.cpu_count() or 1)
        with ThreadPoolExecutor(max_workers=workers) as pool:
            data = list(pool.map(lambda j: synth_crn((j[1], j[1]), j[2], j[0]), jobs))
    return [(j[0], j[1], d) for j, d in zip(jobs, data)]
