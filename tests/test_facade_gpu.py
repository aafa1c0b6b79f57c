"""The reference-compatible C++ boundary on the GPU, called through the
reference's mangled names: the crnd:: transcoder (crnd_unpack_begin /
crnd_unpack_level / crnd_unpack_end) and crn_decompress_crn_to_dds against
the oracle's transcode; crn_compress to DDS against the per-level GPU
encoders and the oracle's optimizer; crn_compress to CRN through the
transcoder; the block compressor API and the two optimizer classes against
the oracle."""
import ctypes as C
import struct

import numpy as np
import pytest

import oracle
from crunch2_amd import dxt1, texture
from crunch2_amd._lib import lib
from tests import crn_cases

pytestmark = pytest.mark.gpu


def fn(name, res, args):
    f = getattr(lib(), name)
    f.restype, f.argtypes = res, args
    return f


@pytest.mark.parametrize("case", [2, 3, 5])
def test_crnd_unpack_through_mangled_names(gpu, case):
    data, _ = crn_cases.case(case)
    print("crnd facade case", case, flush=True)
    begin = fn("_ZN4crnd17crnd_unpack_beginEPKvj", C.c_void_p, [C.c_void_p, C.c_uint32])
    level = fn("_ZN4crnd17crnd_unpack_levelEPvPS0_jjj", C.c_bool,
               [C.c_void_p, C.POINTER(C.c_void_p), C.c_uint32, C.c_uint32, C.c_uint32])
    end = fn("_ZN4crnd15crnd_unpack_endEPv", C.c_bool, [C.c_void_p])
    buf = C.create_string_buffer(data, len(data))
    ctx = begin(buf, len(data))
    assert ctx
    ref = oracle.unpack_all(data)
    info = oracle.info(data)
    for l in sorted(ref):
        faces = ref[l]
        n = faces[0].nbytes
        outs = [C.create_string_buffer(n) for _ in faces]
        ptrs = (C.c_void_p * len(outs))(*[C.cast(o, C.c_void_p) for o in outs])
        assert level(ctx, ptrs, n, 0, l)
        for f, o in enumerate(outs):
            assert o.raw == faces[f].tobytes(), (l, f)
    assert not level(ctx, ptrs, n, 0, len(ref))            # level out of range
    assert end(ctx)
    del info


@pytest.mark.parametrize("case", [2, 3, 5])
def test_crn_to_dds(gpu, case):
    data, _ = crn_cases.case(case)
    f = fn("_Z25crn_decompress_crn_to_ddsPKvRj", C.c_void_p, [C.c_void_p, C.POINTER(C.c_uint32)])
    size = C.c_uint32(len(data))
    p = f(C.create_string_buffer(data, len(data)), C.byref(size))
    assert p and size.value > 128
    raw = C.string_at(p, size.value)
    fn("_Z14crn_free_blockPv", None, [C.c_void_p])(p)
    ref = oracle.unpack_all(data)
    hdr = struct.unpack("<31I", raw[4:128])
    info = oracle.info(data)
    assert hdr[3] == info["width"] and hdr[2] == info["height"]
    faces = len(ref[0])
    body = b"".join(ref[l][fc].tobytes() for fc in range(faces) for l in sorted(ref))
    assert raw[128:] == body
    bad = C.c_uint32(10)
    assert not f(C.create_string_buffer(data[:10], 10), C.byref(bad)) and bad.value == 0


class CompParams(C.Structure):
    _fields_ = [("size_of_obj", C.c_uint32), ("file_type", C.c_uint32), ("faces", C.c_uint32),
                ("width", C.c_uint32), ("height", C.c_uint32), ("levels", C.c_uint32),
                ("format", C.c_int64), ("flags", C.c_uint32), ("pad1", C.c_uint32),
                ("images", (C.c_void_p * 16) * 6), ("target_bitrate", C.c_float), ("quality_level", C.c_uint32),
                ("dxt1a_alpha_threshold", C.c_uint32), ("dxt_quality", C.c_uint32),
                ("dxt_compressor_type", C.c_uint32), ("alpha_component", C.c_uint32),
                ("color_derating", C.c_float), ("alpha_derating", C.c_float), ("pal", C.c_uint32 * 4),
                ("num_helper_threads", C.c_uint32), ("userdata0", C.c_uint32), ("userdata1", C.c_uint32),
                ("pad2", C.c_uint32), ("progress_func", C.c_void_p), ("progress_data", C.c_void_p)]


def comp_params(w, h, levels, fmt, file_type):
    p = CompParams()
    p.size_of_obj = C.sizeof(CompParams)
    p.file_type, p.faces, p.width, p.height, p.levels, p.format = file_type, 1, w, h, levels, fmt
    p.flags = 1 | 2 | 8                      # perceptual, hierarchical, both block types: the reference's defaults
    p.quality_level, p.dxt1a_alpha_threshold, p.dxt_quality, p.alpha_component = 255, 128, 4, 3
    p.color_derating = p.alpha_derating = 2.0
    return p


def test_comp_params_layout_matches_the_header(gpu):
    # offsets g++ gives the members of include/cxx/crnlib.h's crn_comp_params
    # (the reference's member order; crn_format is 8 bytes wide)
    assert CompParams.format.offset == 24 and CompParams.flags.offset == 32 and CompParams.images.offset == 40
    assert CompParams.target_bitrate.offset == 808 and CompParams.progress_func.offset == 872
    assert C.sizeof(CompParams) == 888 and C.sizeof(D1Params) == 40 and C.sizeof(D1Results) == 32


def compress(p, levels_imgs):
    keep = [np.ascontiguousarray(im) for im in levels_imgs]
    for l, im in enumerate(keep):
        p.images[0][l] = im.ctypes.data
    f = fn("_Z12crn_compressRK15crn_comp_paramsRjPjPf", C.c_void_p,
           [C.POINTER(CompParams), C.POINTER(C.c_uint32), C.c_void_p, C.c_void_p])
    size = C.c_uint32(0)
    out = f(C.byref(p), C.byref(size), None, None)
    if not out:
        return None
    raw = C.string_at(out, size.value)
    fn("_Z14crn_free_blockPv", None, [C.c_void_p])(out)
    return raw


def test_crn_compress_dds_dxt1_matches_the_oracle_search(gpu):
    img = texture.synth(64, 64, 5, texture.RGB)
    chain = texture.mip_chain(img, 3)
    raw = compress(comp_params(64, 64, 3, 0, 1), chain)
    assert raw is not None and raw[:4] == b"DDS "
    off = 128
    for l, im in enumerate(chain):
        blocks = dxt1.image_blocks(im)
        n = len(blocks)
        got = np.frombuffer(raw[off:off + 8 * n], np.uint64)
        off += 8 * n
        sel, err, lo, hi = oracle.dxt1_blocks(blocks, 4, oracle.DXT1_PERCEPTUAL | oracle.DXT1_ALPHA_BLOCKS)[:4]
        sel_w = (sel.astype(np.uint64) << (2 * np.arange(16, dtype=np.uint64))).sum(1)
        exp = lo.astype(np.uint64) | (hi.astype(np.uint64) << 16) | (sel_w << 32)
        assert np.array_equal(got, exp), l
    assert off == len(raw)


@pytest.mark.parametrize("fmt", [0, 2, 7])
def test_crn_compress_crn_round_trip(gpu, fmt):
    kind = {0: texture.RGB, 2: texture.RGBA, 7: texture.NORMAL}[fmt]
    img = texture.synth(128, 128, 9, kind)
    chain = texture.mip_chain(img, 4)
    p = comp_params(128, 128, 4, fmt, 0)
    p.quality_level = 128
    raw = compress(p, chain)
    assert raw is not None
    assert oracle.validate_file(raw)
    ref = oracle.unpack_all(raw)
    assert sorted(ref) == [0, 1, 2, 3]
    # the CRN decodes to a close copy of each level
    from oracle import dxt_decode as dd
    for l, im in enumerate(chain):
        h, w = im.shape[:2]
        dec = dd.decode_image(ref[l][0].tobytes(), fmt, w, h).astype(np.float64)
        ch = [0, 1, 2] if fmt == 0 else ([0, 1] if fmt == 7 else [0, 1, 2, 3])
        mse = ((dec[..., ch] - im[..., ch].astype(np.float64)) ** 2).mean()
        psnr = 10 * np.log10(255.0 ** 2 / max(mse, 1e-9))
        print("format", fmt, "level", l, "psnr %.2f" % psnr, flush=True)
        assert psnr > (30.0 if l == 0 else 22.0)     # small mips share few palette entries


def test_block_compressor_api(gpu):
    p = comp_params(4, 4, 1, 0, 1)
    create = fn("_Z27crn_create_block_compressorRK15crn_comp_params", C.c_void_p, [C.POINTER(CompParams)])
    block = fn("_Z18crn_compress_blockPvPKjS_", None, [C.c_void_p, C.c_void_p, C.c_void_p])
    free = fn("_Z25crn_free_block_compressorPv", None, [C.c_void_p])
    blocks = dxt1.image_blocks(texture.synth(32, 32, 11, texture.RGBA))
    for fmt in (0, 2, 9, 7):
        p.format = fmt
        ctx = create(C.byref(p))
        assert ctx
        bpb = 8 if fmt in (0, 9) else 16
        batch = np.zeros((len(blocks), bpb), np.uint8)
        assert lib().crn2_compress_blocks(blocks.ctypes.data_as(C.c_void_p), len(blocks), fmt, 4, 1 | 2,
                                          batch.ctypes.data_as(C.c_void_p)) == 0
        for i in range(0, len(blocks), 7):
            out = np.zeros(bpb, np.uint8)
            block(ctx, blocks[i].ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
            assert np.array_equal(out, batch[i]), (fmt, i)
        free(ctx)
    # colour part of DXT1 against the oracle
    sel, err, lo, hi = oracle.dxt1_blocks(blocks, 4, oracle.DXT1_PERCEPTUAL | oracle.DXT1_ALPHA_BLOCKS)[:4]
    p.format = 0
    ctx = create(C.byref(p))
    out = np.zeros(8, np.uint8)
    block(ctx, blocks[3].ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p))
    w = out.view(np.uint16)
    assert (w[0], w[1]) == (lo[3], hi[3])
    free(ctx)


class D1Params(C.Structure):
    _fields_ = [("block_index", C.c_uint32), ("pixels", C.c_void_p), ("num_pixels", C.c_uint32),
                ("threshold", C.c_uint32), ("quality", C.c_uint32), ("have_alpha", C.c_bool),
                ("use_alpha_blocks", C.c_bool), ("perceptual", C.c_bool), ("grayscale", C.c_bool),
                ("caching", C.c_bool), ("transparent_black", C.c_bool), ("force_alpha", C.c_bool)]


class D1Results(C.Structure):
    _fields_ = [("error", C.c_uint64), ("low", C.c_uint16), ("high", C.c_uint16), ("selectors", C.c_void_p),
                ("alpha_block", C.c_bool), ("reordered", C.c_bool), ("alt", C.c_bool), ("enforce", C.c_bool),
                ("enforced", C.c_uint8)]


def test_dxt1_optimizer_class(gpu):
    ctor = fn("_ZN6crnlib23dxt1_endpoint_optimizerC1Ev", None, [C.c_void_p])
    compute = fn("_ZN6crnlib23dxt1_endpoint_optimizer7computeERKNS0_6paramsERNS0_7resultsE", C.c_bool,
                 [C.c_void_p, C.POINTER(D1Params), C.POINTER(D1Results)])
    obj = C.create_string_buffer(64)
    ctor(obj)
    blocks = dxt1.image_blocks(texture.synth(32, 32, 12, texture.RGB))
    ref = oracle.dxt1_blocks(blocks, 4, oracle.DXT1_PERCEPTUAL | oracle.DXT1_ALPHA_BLOCKS)
    for i in range(0, len(blocks), 5):
        p = D1Params(0, blocks[i].ctypes.data, 16, 128, 4, False, True, True, False, True, False, False)
        sel = np.zeros(16, np.uint8)
        r = D1Results()
        r.selectors = sel.ctypes.data
        assert compute(obj, C.byref(p), C.byref(r))
        assert (r.error, r.low, r.high, r.alpha_block) == (ref[1][i], ref[2][i], ref[3][i], ref[4][i]), i
        assert np.array_equal(sel, ref[0][i])
    # any number of texels on the dxt_hc path (uber, four colours only)
    px = texture.synth(64, 64, 13, texture.RGB).reshape(-1, 4)[:1000].copy()
    p = D1Params(0, px.ctypes.data, len(px), 128, 4, False, False, True, False, False, False, False)
    sel = np.zeros(len(px), np.uint8)
    r = D1Results()
    r.selectors = sel.ctypes.data
    assert compute(obj, C.byref(p), C.byref(r))
    osel, ores = oracle.dxt1_cluster(px, 4, oracle.DXT1_PERCEPTUAL)
    assert (r.error, r.low, r.high) == (ores.error, ores.low, ores.high)
    assert np.array_equal(sel, osel)
    # DXT1A transparency at normal quality, and the transparent-black trial
    ab = blocks.copy()
    ab[::3, ::4, 3] = 40
    ref_a = oracle.dxt1_blocks(ab, 2, oracle.DXT1_PERCEPTUAL | oracle.DXT1_ALPHA_BLOCKS | oracle.DXT1_HAVE_ALPHA |
                               oracle.dxt1_threshold(128))
    for i in range(0, len(ab), 7):
        p = D1Params(0, ab[i].ctypes.data, 16, 128, 2, True, True, True, False, True, False, False)
        sel = np.zeros(16, np.uint8)
        r = D1Results()
        r.selectors = sel.ctypes.data
        assert compute(obj, C.byref(p), C.byref(r))
        assert (r.error, r.low, r.high, r.alpha_block) == (ref_a[1][i], ref_a[2][i], ref_a[3][i], ref_a[4][i]), i
        assert np.array_equal(sel, ref_a[0][i])
    # 1-D lattice of dark texels: the transparent-black trial
    bb = blocks.copy()
    bb[:, ::2, :3] = 2
    ref_b = oracle.dxt1_blocks(bb, 4, oracle.DXT1_PERCEPTUAL | oracle.DXT1_ALPHA_BLOCKS |
                               oracle.DXT1_BLACK_AS_TRANSPARENT | oracle.dxt1_threshold(128))
    for i in range(0, len(bb), 9):
        p = D1Params(0, bb[i].ctypes.data, 16, 128, 4, False, True, True, False, True, True, False)
        sel = np.zeros(16, np.uint8)
        r = D1Results()
        r.selectors = sel.ctypes.data
        assert compute(obj, C.byref(p), C.byref(r))
        assert (r.error, r.low, r.high, r.alpha_block) == (ref_b[1][i], ref_b[2][i], ref_b[3][i], ref_b[4][i]), i
        assert np.array_equal(sel, ref_b[0][i])
    # a cluster of more than 16 texels with alpha: refused, not approximated
    p = D1Params(0, px.ctypes.data, len(px), 128, 4, True, True, True, False, True, False, False)
    assert not compute(obj, C.byref(p), C.byref(D1Results()))


class D5Params(C.Structure):
    _fields_ = [("block_index", C.c_uint32), ("pixels", C.c_void_p), ("num_pixels", C.c_uint32),
                ("comp_index", C.c_uint32), ("quality", C.c_uint32), ("use_both", C.c_bool)]


class D5Results(C.Structure):
    _fields_ = [("selectors", C.c_void_p), ("error", C.c_uint64), ("first", C.c_uint8), ("second", C.c_uint8),
                ("block_type", C.c_uint8), ("reordered", C.c_bool)]


def test_dxt5_optimizer_class(gpu):
    ctor = fn("_ZN6crnlib23dxt5_endpoint_optimizerC1Ev", None, [C.c_void_p])
    compute = fn("_ZN6crnlib23dxt5_endpoint_optimizer7computeERKNS0_6paramsERNS0_7resultsE", C.c_bool,
                 [C.c_void_p, C.POINTER(D5Params), C.POINTER(D5Results)])
    obj = C.create_string_buffer(64)
    ctor(obj)
    blocks = dxt1.image_blocks(texture.synth(32, 32, 14, texture.RGBA))
    vals = np.ascontiguousarray(blocks[..., 3])
    sel_ref, err_ref, first, second = oracle.dxt5a_blocks(vals, 4, True)[:4]
    for i in range(0, len(blocks), 3):
        p = D5Params(0, blocks[i].ctypes.data, 16, 3, 4, True)
        sel = np.zeros(16, np.uint8)
        r = D5Results()
        r.selectors = sel.ctypes.data
        assert compute(obj, C.byref(p), C.byref(r))
        assert r.error == err_ref[i], i
        assert (r.first, r.second) == (first[i], second[i])
        assert np.array_equal(sel, sel_ref[i])


class MipmapParams(C.Structure):
    # include/cxx/crnlib.h's crn_mipmap_params: 21 four-byte members
    _fields_ = [("size_of_obj", C.c_uint32), ("mode", C.c_uint32), ("filter", C.c_uint32),
                ("gamma_filtering", C.c_uint32), ("gamma", C.c_float), ("blurriness", C.c_float),
                ("max_levels", C.c_uint32), ("min_mip_size", C.c_uint32), ("renormalize", C.c_uint32),
                ("rtopmip", C.c_uint32), ("tiled", C.c_uint32), ("scale_mode", C.c_uint32), ("scale_x", C.c_float),
                ("scale_y", C.c_float), ("window", C.c_uint32 * 4), ("clamp_scale", C.c_uint32),
                ("clamp_width", C.c_uint32), ("clamp_height", C.c_uint32)]


def mip_params(**kw):
    m = MipmapParams()
    m.size_of_obj = C.sizeof(MipmapParams)
    m.mode, m.filter, m.gamma_filtering, m.gamma, m.blurriness = 0, 4, 1, 2.2, 0.9
    m.max_levels, m.min_mip_size, m.scale_x, m.scale_y = 16, 1, 1.0, 1.0
    for k, v in kw.items():
        setattr(m, k, v)
    return m


@pytest.mark.parametrize("case", [
    dict(mode=2),                                              # generate, crnlib defaults (Kaiser, sRGB)
    dict(mode=0, filter=0, blurriness=1.0, tiled=1),           # source-or-generate, one level given
    dict(mode=2, filter=2, gamma_filtering=0, max_levels=4, min_mip_size=2),
    dict(mode=2, filter=3, helper_threads=2, gamma=1.75),      # threaded order; gamma is not used
])
def test_crn_compress_generates_mips_like_generate_mipmaps(gpu, case):
    case = dict(case)
    threads = case.pop("helper_threads", 0)
    img = texture.synth(96, 64, 7, texture.RGBA)
    m = mip_params(**case)
    p = comp_params(96, 64, 1, 0, 1)
    p.num_helper_threads = threads
    keep = np.ascontiguousarray(img)
    p.images[0][0] = keep.ctypes.data
    f = fn("_Z12crn_compressRK15crn_comp_paramsRK17crn_mipmap_paramsRjPjPf", C.c_void_p,
           [C.POINTER(CompParams), C.POINTER(MipmapParams), C.POINTER(C.c_uint32), C.c_void_p, C.c_void_p])
    size = C.c_uint32(0)
    out = f(C.byref(p), C.byref(m), C.byref(size), None, None)
    assert out
    raw = C.string_at(out, size.value)
    fn("_Z14crn_free_blockPv", None, [C.c_void_p])(out)
    flags = (oracle.MIP_SRGB if m.gamma_filtering else 0) | (oracle.MIP_WRAP if m.tiled else 0) | \
        oracle.MIP_ALPHA | (oracle.MIP_THREADED if threads else 0)
    levels, _ = oracle.mip_generate(img, m.filter, m.blurriness, 2.2, flags, m.max_levels, m.min_mip_size)
    p2 = comp_params(96, 64, len(levels), 0, 1)
    exp = compress(p2, levels)
    assert raw == exp


@pytest.mark.parametrize("fmt,kind", [(0, texture.RGB), (0, texture.RGBA), (2, texture.RGBA), (7, texture.NORMAL)])
def test_crn_compress_clustered_dds_matches_the_oracle(gpu, fmt, kind):
    # DDS below quality 255: qdxt1 / qdxt5 (crn_dds_comp.cpp:158-205) with the
    # elements interleaved as qdxt_pack does (crn_mipmapped_texture.cpp:2536-2553);
    # DXT1 of an RGBA source: transparent texels below the threshold of 128
    img = texture.synth(96, 64, 13 + fmt, kind)
    chain = texture.mip_chain(img, 4)
    p = comp_params(96, 64, 4, fmt, 1)
    p.quality_level = 100
    raw = compress(p, chain)
    assert raw is not None and raw[:4] == b"DDS "
    from crunch2_amd.qdxt import level_blocks
    blocks, desc = level_blocks(chain)
    P, A, H = oracle.QDXT_PERCEPTUAL, oracle.QDXT_ALPHA_BLOCKS, oracle.QDXT_HIERARCHICAL
    elems = []
    if fmt == 0:
        elems.append(oracle.qdxt1(blocks, desc, 100, 4, P | A | H, 128, 1.0)[0])
    elif fmt == 2:
        elems.append(oracle.qdxt5(blocks, desc, 100, 4, oracle.QDXT5_BOTH | H, 3)[0])
        elems.append(oracle.qdxt1(blocks, desc, 100, 4, P | H, 128, 0.75)[0])
    else:
        elems.append(oracle.qdxt5(blocks, desc, 100, 4, oracle.QDXT5_BOTH | H, 0)[0])
        elems.append(oracle.qdxt5(blocks, desc, 100, 4, oracle.QDXT5_BOTH | H, 1)[0])
    exp = np.stack(elems, 1).reshape(-1)
    got = np.frombuffer(raw[128:], np.uint64)
    assert got.shape == exp.shape and np.array_equal(got, exp)
