"""GPU DXT1 search at the qualities below Better (probe tables and the
projection evaluator, crn_dxt1.cpp:716-746, 1594-1757), with DXT1A
transparency (:2116, :300-343) and the transparent-black trial
(:2001-2079), against the oracle on the same blocks: every block bit-exact
(endpoints, selectors, error, flags)."""
import numpy as np
import pytest
import torch

import oracle
from crunch2_amd import dxt1, texture
from crunch2_amd._lib import lib

pytestmark = pytest.mark.gpu


def sample(seed, n_random=256):
    rng = np.random.default_rng(seed)
    b = np.concatenate([dxt1.image_blocks(texture.synth(64, 64, seed, texture.RGB)),
                        dxt1.image_blocks(texture.synth(64, 64, seed + 1, texture.RGBA)),
                        rng.integers(0, 256, (n_random, 16, 4)).astype(np.uint8)])
    b[..., 3] = 255
    return b


def same(got, ref):
    ok = np.ones(len(ref[0]), bool)
    for k in range(7):
        ok &= (np.asarray(got[k]).reshape(len(ok), -1) == np.asarray(ref[k]).reshape(len(ok), -1)).all(1)
    return ok


@pytest.mark.parametrize("quality", [0, 1, 2])
@pytest.mark.parametrize("flags", [dxt1.PERCEPTUAL | dxt1.ALPHA_BLOCKS, dxt1.PERCEPTUAL, dxt1.ALPHA_BLOCKS,
                                   dxt1.GRAYSCALE | dxt1.ALPHA_BLOCKS])
def test_low_qualities_match_oracle(gpu, quality, flags):
    print("dxt1 quality", quality, "flags", flags, flush=True)
    b = sample(3 + quality)
    ok = same(dxt1.encode_blocks(b, quality, flags), oracle.dxt1_blocks(b, quality, flags))
    assert ok.all(), np.flatnonzero(~ok)[:10]


def dxt1a_blocks(seed):
    rng = np.random.default_rng(seed)
    b = sample(seed, 128)
    b[..., 3] = np.where(rng.random(b.shape[:2]) < 0.3, rng.integers(0, 128, b.shape[:2]), 255).astype(np.uint8)
    b[-4:, :, 3] = 3
    b[-8:-4, :8, :3] = b[-8:-4, :1, :3]
    b[-8:-4, 8:, 3] = 0
    return b


@pytest.mark.parametrize("quality", [0, 2, 3, 4])
def test_dxt1a_matches_oracle(gpu, quality):
    print("dxt1a quality", quality, flush=True)
    b = dxt1a_blocks(20 + quality)
    flags = dxt1.PERCEPTUAL | dxt1.ALPHA_BLOCKS | dxt1.HAVE_ALPHA | dxt1.threshold(128)
    ok = same(dxt1.encode_blocks(b, quality, flags), oracle.dxt1_blocks(b, quality, flags))
    assert ok.all(), np.flatnonzero(~ok)[:10]


@pytest.mark.parametrize("quality", [1, 3, 4])
def test_black_as_transparent_matches_oracle(gpu, quality):
    print("black-as-transparent quality", quality, flush=True)
    rng = np.random.default_rng(40 + quality)
    b = sample(30 + quality, 128)
    dark = rng.random(b.shape[:2]) < 0.25
    b[dark, :3] = rng.integers(0, 5, (int(dark.sum()), 3))
    flags = dxt1.PERCEPTUAL | dxt1.ALPHA_BLOCKS | dxt1.BLACK_AS_TRANSPARENT | dxt1.threshold(128)
    ref = oracle.dxt1_blocks(b, quality, flags)
    ok = same(dxt1.encode_blocks(b, quality, flags), ref)
    assert ok.all(), np.flatnonzero(~ok)[:10]
    assert (np.asarray(ref[4]) == 1).sum() > 0


def test_dxt1a_image_entry(gpu):
    # the image entry with DXT1A for transparency: blocks with a texel of
    # alpha below 128 are encoded with that texel transparent
    img = texture.synth(256, 200, 55, texture.RGBA)
    blocks = dxt1.image_blocks(img)
    has = (blocks[..., 3] < 128).any(1)
    base = dxt1.PERCEPTUAL | dxt1.ALPHA_BLOCKS
    ref_plain = oracle.dxt1_blocks(blocks[~has], 2, base)
    ref_a = oracle.dxt1_blocks(blocks[has], 2, base | dxt1.HAVE_ALPHA | dxt1.threshold(128))
    t = torch.from_numpy(img).cuda()
    out = torch.empty(len(blocks), dtype=torch.int64, device="cuda")
    assert lib().crn2_encode_image_device(t.data_ptr(), 256, 200, 0, 2, 1 | 2 | 8 | (128 << 8), out.data_ptr(),
                                          None) == 0
    torch.cuda.synchronize()
    w = out.cpu().numpy().view(np.uint64)
    lohi = (w & 0xFFFFFFFF).astype(np.uint32)
    sel = (w >> 32).astype(np.uint32)
    assert has.sum() > 50 and (~has).sum() > 50
    for mask, ref in ((~has, ref_plain), (has, ref_a)):
        rsel = (ref[0].astype(np.uint32) << (2 * np.arange(16, dtype=np.uint32))).sum(1).astype(np.uint32)
        assert np.array_equal(lohi[mask] & 0xFFFF, ref[2])
        assert np.array_equal(lohi[mask] >> 16, ref[3])
        assert np.array_equal(sel[mask], rsel)


@pytest.mark.parametrize("quality", [3, 4])
def test_grayscale_sampling_matches_oracle(gpu, quality):
    # under grayscale sampling the error planes' bound prunes candidates the
    # luma evaluation would keep, so it is part of the search
    print("grayscale quality", quality, flush=True)
    b = sample(60 + quality, 128)
    for flags in (dxt1.GRAYSCALE | dxt1.ALPHA_BLOCKS, dxt1.GRAYSCALE | dxt1.PERCEPTUAL):
        ok = same(dxt1.encode_blocks(b, quality, flags), oracle.dxt1_blocks(b, quality, flags))
        assert ok.all(), (flags, np.flatnonzero(~ok)[:10])
