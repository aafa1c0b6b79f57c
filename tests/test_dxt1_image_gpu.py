"""Image-level plain DXT1 encode (crn2_encode_dxt1_image_device, dxt_image's
block gather with edge clamp, the search, dxt1_block packing) against the
oracle's optimizer on the same blocks.  Includes the 256x256 texture of the
plain-DXT1 workload (4,096 blocks, uber, perceptual, both block types): every
8-byte block and every per-block error bit-exact."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle
from crunch2_amd import dxt1, texture
from crunch2_amd._lib import lib

pytestmark = pytest.mark.gpu


def gpu_image(img, quality, flags):
    h, w = img.shape[:2]
    n = ((w + 3) // 4) * ((h + 3) // 4)
    d = torch.from_numpy(np.ascontiguousarray(img).reshape(-1)).cuda()
    blocks = torch.zeros(n, dtype=torch.int64, device="cuda")
    err = torch.zeros(n, dtype=torch.int64, device="cuda")
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib().crn2_encode_dxt1_image_device(C.c_void_p(d.data_ptr()), w, h, quality, flags,
                                               C.c_void_p(blocks.data_ptr()), 1, C.c_void_p(err.data_ptr()), s) == 0
    torch.cuda.synchronize()
    return blocks.cpu().numpy().view(np.uint64), err.cpu().numpy().view(np.uint64)


def oracle_image(img, quality, flags):
    b = dxt1.image_blocks(img)
    sel, err, lo, hi = oracle.dxt1_blocks(b, quality, flags)[:4]
    sel_w = (sel.astype(np.uint64) << (2 * np.arange(16, dtype=np.uint64))).sum(1)
    return lo.astype(np.uint64) | (hi.astype(np.uint64) << np.uint64(16)) | (sel_w << np.uint64(32)), err


@pytest.mark.parametrize("size,seed,flags,quality", [
    ((256, 256), 1, dxt1.PERCEPTUAL | dxt1.ALPHA_BLOCKS, 4),     # the plain DXT1 workload, 256x256 uber
    ((61, 37), 2, dxt1.PERCEPTUAL | dxt1.ALPHA_BLOCKS, 4),       # edge clamp on both axes
    ((64, 64), 3, dxt1.PERCEPTUAL, 3),
    ((5, 3), 4, dxt1.ALPHA_BLOCKS, 4)])
def test_image_encode_matches_oracle(gpu, size, seed, flags, quality):
    w, h = size
    img = texture.synth(w, h, seed, texture.RGB)
    print("dxt1 image", w, h, "flags", flags, "quality", quality, flush=True)
    got_b, got_e = gpu_image(img, quality, flags)
    exp_b, exp_e = oracle_image(img, quality, flags)
    assert len(got_b) == ((w + 3) // 4) * ((h + 3) // 4)
    assert np.array_equal(got_e, exp_e)
    assert np.array_equal(got_b, exp_b)


def test_image_encode_is_the_block_kernel(gpu):
    # the image entry gathers the same 16 texels the block entry is given
    img = texture.synth(64, 48, 5, texture.RGBA)
    got_b, got_e = gpu_image(img, 4, dxt1.PERCEPTUAL | dxt1.ALPHA_BLOCKS)
    sel, err, lo, hi = dxt1.encode_blocks(dxt1.image_blocks(img), 4, dxt1.PERCEPTUAL | dxt1.ALPHA_BLOCKS)[:4]
    assert np.array_equal(got_e, err)
    assert np.array_equal(got_b & np.uint64(0xFFFFFFFF), lo.astype(np.uint64) | (hi.astype(np.uint64) << np.uint64(16)))
