"""GPU DXT1 endpoint search vs the oracle restatement on the same blocks:
every block bit-exact (endpoints, selectors, error and flags)."""
import numpy as np
import pytest

import oracle
from crunch2_amd import dxt1, texture

pytestmark = pytest.mark.gpu


def blocks_of(seed, size=64, kind=texture.RGB):
    return dxt1.image_blocks(texture.synth(size, size, seed, kind))


@pytest.mark.parametrize("flags,quality", [(dxt1.PERCEPTUAL | dxt1.ALPHA_BLOCKS, 4), (dxt1.PERCEPTUAL, 4),
                                           (dxt1.ALPHA_BLOCKS, 4), (dxt1.PERCEPTUAL | dxt1.ALPHA_BLOCKS, 3)])
def test_matches_oracle(gpu, flags, quality):
    print("dxt1 blocks flags", flags, "quality", quality, flush=True)
    rng = np.random.default_rng(7)
    b = np.concatenate([blocks_of(1), blocks_of(2, kind=texture.RGBA),
                        rng.integers(0, 256, (256, 16, 4)).astype(np.uint8)])
    b[..., 3] = 255
    ref = oracle.dxt1_blocks(b, quality, flags)
    got = dxt1.encode_blocks(b, quality, flags)
    same_block = np.ones(len(b), bool)
    for k in range(7):
        eq = np.asarray(got[k]).reshape(len(b), -1) == np.asarray(ref[k]).reshape(len(b), -1)
        same_block &= eq.all(1)
    rate = same_block.mean()
    err_eq = (np.asarray(got[1]) == np.asarray(ref[1])).mean()
    print("bit-exact rate %.4f, error-equal rate %.4f" % (rate, err_eq), flush=True)
    assert err_eq == 1.0
    assert rate == 1.0
