"""GPU mip generation (crn2_mip_plan_*, crn2_generate_mipmaps_device) against
the oracle's restatement of crnlib's generate_mipmaps, bit for bit: every
filter, clamp and wrap edges, sRGB and linear, 3 and 4 components, the
single-thread Resampler's per-level pass order (Y then X on tall and thin
images) and the multithreaded resampler's, renorm_normal_map, level limits,
odd and one-texel sides, and a whole 4096 x 4096 chain.  At the bench size
(8192 x 8192) the two deepest levels are compared with the oracle's
resample() of the same image, and wrap-mode shift equivariance and constant
images are checked on the whole chain."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle
from crunch2_amd import mips, texture
from crunch2_amd._lib import CRN2_EINVAL, lib

pytestmark = pytest.mark.gpu

BOX, TENT, LANCZOS4, MITCHELL, KAISER = range(5)
SRGB, WRAP, ALPHA, RENORM, THREADED = 1, 2, 4, 8, 16


def gpu_levels(img, filt=KAISER, flags=SRGB | ALPHA, fs=0.9, max_levels=0, min_size=1):
    p = mips.params(filt, flags, fs, 2.2, max_levels, min_size)
    return mips.generate(img, p)


def oracle_levels(img, filt=KAISER, flags=SRGB | ALPHA, fs=0.9, max_levels=0, min_size=1):
    return oracle.mip_generate(img, filt, fs, 2.2, flags, max_levels, min_size)


def rand_img(h, w, seed):
    return np.random.default_rng(seed).integers(0, 256, (h, w, 4), dtype=np.uint8)


def assert_same(got, exp, what):
    assert len(got) == len(exp), what
    for l, (a, b) in enumerate(zip(got, exp)):
        assert a.shape == b.shape, (what, l)
        if not np.array_equal(a, b):
            bad = np.argwhere((a != b).any(-1))
            raise AssertionError("%s level %d: %d texels differ, first %s: %s vs %s" %
                                 (what, l, len(bad), bad[0], a[tuple(bad[0])], b[tuple(bad[0])]))


CASES = [
    (64, 64, KAISER, SRGB | ALPHA, 0.9),
    (100, 37, KAISER, SRGB | ALPHA, 0.9),
    (37, 100, LANCZOS4, ALPHA, 1.0),
    (1, 64, MITCHELL, SRGB, 0.9),
    (64, 1, TENT, SRGB | ALPHA | WRAP, 0.9),
    (8, 256, BOX, ALPHA, 1.0),
    (300, 2, KAISER, SRGB | ALPHA, 0.9),
    (1024, 3, KAISER, ALPHA, 0.9),
    (131, 257, KAISER, SRGB | ALPHA | WRAP, 0.9),
    (256, 256, KAISER, SRGB | ALPHA | THREADED, 0.9),
    (2, 300, LANCZOS4, SRGB | ALPHA | THREADED, 1.3),
    (200, 3, MITCHELL, SRGB, 0.5),
]


@pytest.mark.parametrize("case", CASES, ids=["%dx%d-f%d-fl%d-s%g" % c for c in CASES])
def test_chain_matches_oracle(gpu, case):
    h, w, filt, flags, fs = case
    img = rand_img(h, w, h * 131 + w)
    exp, orders = oracle_levels(img, filt, flags, fs)
    got = gpu_levels(img, filt, flags, fs)
    assert_same(got, exp, case)
    plan = mips.Plan(w, h, mips.params(filt, flags, fs))
    assert plan.y_first.tolist() == orders.tolist()
    plan.close()


@pytest.mark.parametrize("filt", range(5))
@pytest.mark.parametrize("flags", [SRGB | ALPHA, ALPHA | WRAP, SRGB, SRGB | ALPHA | WRAP | THREADED])
def test_every_filter_and_mode(gpu, filt, flags):
    img = texture.synth(96, 128, filt + 3, texture.RGBA)
    assert_same(gpu_levels(img, filt, flags), oracle_levels(img, filt, flags)[0], (filt, flags))


def test_yfirst_levels_are_exercised(gpu):
    for h, w in [(2048, 2), (1024, 3), (3, 1024)]:
        img = rand_img(h, w, 5)
        exp, orders = oracle_levels(img)
        assert_same(gpu_levels(img), exp, (h, w))
        if h > w:
            assert orders[2:].any()


def test_renormalize_normal_map(gpu):
    img = texture.synth(128, 96, 4, texture.NORMAL)
    img[:16, :16, :3] = 128
    for filt in (KAISER, BOX):
        exp, _ = oracle_levels(img, filt, RENORM)
        assert_same(gpu_levels(img, filt, RENORM), exp, filt)
        plain, _ = oracle_levels(img, filt, 0)
        assert any(not np.array_equal(a, b) for a, b in zip(exp[1:], plain[1:]))


def test_level_limits(gpu):
    img = rand_img(100, 64, 2)
    for mx, ms in [(1, 1), (3, 1), (0, 4), (2, 16), (0, 100)]:
        assert_same(gpu_levels(img, max_levels=mx, min_size=ms), oracle_levels(img, max_levels=mx, min_size=ms)[0],
                    (mx, ms))


def test_plan_reuse_and_one_shot(gpu):
    a, b = rand_img(300, 200, 1), rand_img(300, 200, 2)
    p = mips.params()
    plan = mips.Plan(200, 300, p)
    ta, tb = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    out = torch.zeros(plan.chain_bytes + 64, dtype=torch.uint8, device="cuda")
    for img, t in ((a, ta), (b, tb), (a, ta)):
        plan.run(t, out)
        torch.cuda.synchronize()
        assert_same(plan.split(out.cpu().numpy()), oracle_levels(img)[0], "reuse")
    plan.close()
    one = torch.zeros(plan.chain_bytes, dtype=torch.uint8, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    assert lib().crn2_generate_mipmaps_device(C.c_void_p(tb.data_ptr()), 200, 300, C.byref(p),
                                              C.c_void_p(one.data_ptr()), one.numel(), C.c_void_p(s)) == 0
    assert torch.equal(one, mips.Plan(200, 300, p).run(tb))
    assert lib().crn2_generate_mipmaps_device(C.c_void_p(tb.data_ptr()), 200, 300, C.byref(p),
                                              C.c_void_p(one.data_ptr()), one.numel() - 1, C.c_void_p(s)) == \
        CRN2_EINVAL


def test_4096_chain_matches_oracle(gpu):
    img = texture.synth(4096, 4096, 17, texture.RGBA)
    exp, _ = oracle_levels(img)
    assert_same(gpu_levels(img), exp, "4096 kaiser srgb")


@pytest.fixture(scope="module")
def big(gpu):
    img = texture.synth(8192, 8192, 23, texture.RGBA)
    plan = mips.Plan(8192, 8192, mips.params())
    chain = plan.run(torch.from_numpy(img).cuda())
    torch.cuda.synchronize()
    lv = plan.split(chain.cpu().numpy())
    plan.close()
    return img, lv


def test_8192_deepest_levels_match_oracle(big):
    img, lv = big
    assert len(lv) == 14
    for l in (13, 12):
        w = h = 8192 >> l
        exp, order = oracle.mip_resample(img, w, h)
        assert order == 0
        assert np.array_equal(lv[l], exp), l


def test_8192_wrap_shift_and_constant(gpu):
    img = texture.synth(8192, 8192, 29, texture.RGBA)
    p = mips.params(flags=SRGB | ALPHA | WRAP)
    plan = mips.Plan(8192, 8192, p)
    t = torch.from_numpy(img).cuda()
    base = plan.split(plan.run(t).cpu().numpy())
    shifted = plan.split(plan.run(torch.roll(t, (64, 128), (0, 1))).cpu().numpy())
    for l in range(1, 7):
        assert np.array_equal(shifted[l], np.roll(base[l], (64 >> l, 128 >> l), (0, 1))), l
    const = plan.split(plan.run(torch.full_like(t, 200)).cpu().numpy())
    s2l, l2s = oracle.mip_tables(2.2)
    c = l2s[min(8191, int(np.float32(8192) * s2l[200] + np.float32(0.5)))]
    for l in const[1:]:
        assert (l[..., :3] == c).all() and (l[..., 3] == 200).all()
    plan.close()
