"""GPU transcoder parity: the HIP path (through the C ABI) against the known
answers of tests/crn_cases.py and the CPU oracle, bit-exact."""
import ctypes as C
import sys

import numpy as np
import pytest

import oracle
from crunch2_amd import crnd, texture
from crunch2_amd._lib import FMT_DXT1, FMT_DXT5, FMT_DXN_XY, FMT_DXT5A, lib
from tests import crn_cases

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def gpu_all_levels(data):
    ctx = crnd.unpack_begin(data)
    assert ctx is not None
    inf = crnd.get_texture_info(data)
    out = {}
    for l in range(inf["levels"]):
        r = crnd.unpack_level(ctx, l)
        assert r is not False, l
        out[l] = r
    assert crnd.unpack_end(ctx)
    return out


@pytest.mark.parametrize("i", range(len(crn_cases.CASES)))
def test_known_answer(gpu, i):
    say("known-answer case", crn_cases.CASES[i][:6])
    data, expected = crn_cases.case(i)
    got = gpu_all_levels(data)
    for l, faces in expected.items():
        for f, exp in enumerate(faces):
            assert np.array_equal(got[l][f], exp), (i, l, f)


SYNTH = [
    ((256, 256), 11, FMT_DXT1, 1),
    ((256, 128), 12, FMT_DXT5, 1),
    ((128, 128), 13, FMT_DXN_XY, 1),
    ((96, 200), 14, FMT_DXT5A, 1),
    ((64, 64), 15, FMT_DXT1, 6),
    ((64, 64), 16, FMT_DXT5, 6),
    ((1024, 1024), 17, FMT_DXT1, 1),
]


@pytest.mark.parametrize("size,seed,fmt,faces", SYNTH)
def test_synthetic_matches_oracle(gpu, size, seed, fmt, faces):
    say("synthetic", size, seed, fmt, faces)
    data = texture.synth_crn(size, seed, fmt, faces=faces)
    ref = oracle.unpack_all(data)
    got = gpu_all_levels(data)
    for l in ref:
        for f in range(len(ref[l])):
            assert np.array_equal(got[l][f], ref[l][f]), (l, f)


def test_device_levels_api_and_pitch(gpu):
    import torch
    say("device api")
    data = texture.synth_crn(200, 21, FMT_DXT5)
    ref = oracle.unpack_all(data)
    ctx = crnd.unpack_begin(data)
    n = crnd.levels_size(ctx, 0, len(ref))
    buf = torch.zeros(n, dtype=torch.uint8, device="cuda")
    crnd.unpack_levels_device(ctx, buf.data_ptr(), n, 0, len(ref))
    torch.cuda.synchronize()
    got = buf.cpu().numpy()
    exp = np.concatenate([ref[l][0] for l in sorted(ref)])
    assert np.array_equal(got, exp)
    # padded pitch through the host entry point keeps the caller's gap bytes
    pitch = 64 * 16 + 12
    by = 50
    dst = [np.full(pitch * by, 0xA5, np.uint8)]
    r = crnd.unpack_level(ctx, 0, row_pitch=pitch, dst=dst)
    rows = r[0].reshape(by, pitch)
    assert np.array_equal(rows[:, :800].reshape(-1), ref[0][0])
    assert (rows[:, 800:] == 0xA5).all()
    # argument checks (crn_decomp.h:3569-3575, 4446)
    assert crnd.unpack_level(ctx, 0, row_pitch=802) is False
    assert crnd.unpack_level(ctx, 0, row_pitch=16 * 49) is False
    assert crnd.unpack_level(ctx, 0, dst=[np.zeros(100, np.uint8)]) is False
    assert crnd.unpack_level(ctx, len(ref)) is False
    assert crnd.unpack_end(ctx)
    assert not crnd.unpack_end(ctx)


def test_segmented_level(gpu):
    say("segmented")
    data = texture.synth_crn(128, 22, FMT_DXT1)
    ref = oracle.unpack_all(data)
    ctx = crnd.unpack_begin(data)
    b = np.frombuffer(data, np.uint8).copy()
    size = C.c_uint32()
    p = lib().crn2_get_level_data(b.ctypes.data_as(C.c_void_p), b.nbytes, 2, C.byref(size))
    off = p - b.ctypes.data
    seg = data[off:off + size.value]
    got = crnd.unpack_level_segmented(ctx, seg, 2)
    assert np.array_equal(got[0], ref[2][0])
    crnd.unpack_end(ctx)


def test_emscripten_crn_decompress(gpu):
    say("crn_decompress")
    data = texture.synth_crn((64, 32), 23, FMT_DXT1)
    ref = oracle.unpack_all(data)
    got = crnd.crn_decompress(data, 1, 4)
    exp = np.concatenate([ref[l][0] for l in range(1, 5)])
    assert np.array_equal(got, exp)


def test_begin_rejects_bad_files(gpu):
    say("bad files")
    data = texture.synth_crn(32, 24, FMT_DXT1)
    assert crnd.unpack_begin(data[:61]) is None
    bad = bytearray(data)
    bad[1] = 0
    assert crnd.unpack_begin(bytes(bad)) is None


@pytest.mark.parametrize("fmt", [FMT_DXT1, FMT_DXT5])
def test_full_size_4096_matches_oracle(gpu, fmt):
    say("4096^2 full mip chain", fmt)
    data = texture.synth_crn(4096, 1, fmt)
    ref = oracle.unpack_all(data)
    got = gpu_all_levels(data)
    for l in ref:
        assert np.array_equal(got[l][0], ref[l][0]), l


def test_batch_plan_mixed_textures_and_graph_replay(gpu):
    """One plan over textures of every format family, odd sizes and a cube
    map; eager run against the oracle, then the captured graph replayed three
    times (outputs cleared by eager work in between), each replay bit-exact."""
    import torch
    say("batch plan + graph replay")
    specs = [((256, 256), 31, FMT_DXT1, 1), ((96, 200), 32, FMT_DXT5A, 1), ((64, 64), 33, FMT_DXT5, 6),
             ((128, 72), 34, FMT_DXN_XY, 1), ((4, 4), 35, FMT_DXT1, 1)]
    files = [texture.synth_crn(sz, seed, fmt, faces=faces) for sz, seed, fmt, faces in specs]
    ctxs, outs, exps = [], [], []
    for d in files:
        c = crnd.unpack_begin(d)
        assert c is not None
        inf = crnd.get_texture_info(d)
        outs.append(torch.zeros(crnd.levels_size(c, 0, inf["levels"]), dtype=torch.uint8, device="cuda"))
        ctxs.append(c)
        ref = oracle.unpack_all(d)
        exps.append(np.concatenate([np.concatenate([f.reshape(-1) for f in ref[l]]) for l in sorted(ref)]))
    n = len(ctxs)
    plan = C.c_void_p()
    assert lib().crn2_batch_create((C.c_void_p * n)(*[c.handle for c in ctxs]), n,
                                   (C.c_void_p * n)(*[o.data_ptr() for o in outs]),
                                   (C.c_uint64 * n)(*[o.numel() for o in outs]), C.byref(plan)) == 0
    stream = torch.cuda.current_stream()
    assert lib().crn2_batch_run(plan, C.c_void_p(stream.cuda_stream)) == 0
    torch.cuda.synchronize()
    for o, e in zip(outs, exps):
        assert np.array_equal(o.cpu().numpy(), e)
    for rep in range(3):
        for o in outs:
            o.zero_()
        junk = torch.randn(1 << 20, device="cuda").sum()
        torch.cuda.synchronize()
        assert lib().crn2_batch_run_graph(plan, C.c_void_p(stream.cuda_stream)) == 0
        torch.cuda.synchronize()
        for o, e in zip(outs, exps):
            assert np.array_equal(o.cpu().numpy(), e), rep
        del junk
    assert lib().crn2_batch_status(plan, C.c_void_p(stream.cuda_stream)) == 0
    assert lib().crn2_batch_destroy(plan) == 0
    for c in ctxs:
        assert crnd.unpack_end(c)


@pytest.mark.parametrize("fmt,pal,factor", crn_cases.SHRUNK)
def test_out_of_range_palette_indices_fail_like_oracle(gpu, fmt, pal, factor):
    """A palette cut short makes some levels reach past it: each level must
    fail or succeed as the oracle's does, and match it byte for byte when it
    succeeds."""
    say("shrunk palette", fmt, pal, factor)
    data = crn_cases.shrink_palette(texture.synth_crn((256, 256), 31, fmt), pal, factor, oracle.crc16)
    levels = crnd.get_texture_info(data)["levels"]
    ctx = crnd.unpack_begin(data)
    assert ctx is not None
    with oracle.Context(data) as oc:
        assert oc.h
        fails = 0
        for l in range(levels):
            want = oc.unpack_level(l)
            got = crnd.unpack_level(ctx, l)
            assert (got is False) == (want is False), l
            if want is False:
                fails += 1
            else:
                assert np.array_equal(got[0], want[0]), l
    assert crnd.unpack_end(ctx)
    assert 0 < fails < levels
