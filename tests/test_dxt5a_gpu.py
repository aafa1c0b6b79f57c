"""GPU DXT5 alpha search vs the oracle restatement: bit-exact endpoints,
selectors, block type, reorder flag and error."""
import numpy as np
import pytest
import torch

import oracle
from crunch2_amd import dxt5a, texture
from tests import dxt5a_cases

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("quality,use_both", [(4, True), (3, True), (2, True), (4, False)])
def test_blocks_match_oracle(gpu, quality, use_both):
    print("dxt5a blocks q", quality, use_both, flush=True)
    b = dxt5a_cases.blocks(10 + quality)
    ref = oracle.dxt5a_blocks(b, quality, use_both)
    got = dxt5a.compute_blocks(b, quality, use_both)
    for k, name in enumerate(("selectors", "error", "first", "second", "type", "reordered")):
        assert np.array_equal(np.asarray(got[k]), np.asarray(ref[k])), name


@pytest.mark.parametrize("w,h,comp", [(64, 64, 3), (37, 21, 0), (256, 128, 1)])
def test_image_encode_matches_oracle(gpu, w, h, comp):
    print("dxt5a image", w, h, comp, flush=True)
    img = texture.synth(w, h, 31 + w, texture.RGBA if comp == 3 else texture.NORMAL)
    t = torch.from_numpy(img).cuda()
    bx, by = (w + 3) // 4, (h + 3) // 4
    err = torch.zeros(bx * by, dtype=torch.int64, device="cuda")
    blocks = dxt5a.encode_image(t, comp, errors=err).cpu().numpy().view(np.uint64)
    ys = np.minimum(np.arange(by * 4), h - 1)
    xs = np.minimum(np.arange(bx * 4), w - 1)
    ch = img[ys][:, xs][..., comp].reshape(by, 4, bx, 4).transpose(0, 2, 1, 3).reshape(-1, 16)
    sel, e, f, s, _, _ = oracle.dxt5a_blocks(ch)
    shifts = (3 * np.arange(16)).astype(np.uint64)
    exp = f.astype(np.uint64) | (s.astype(np.uint64) << np.uint64(8)) | (
        (sel.astype(np.uint64) << shifts[None, :]).sum(1).astype(np.uint64) << np.uint64(16))
    assert np.array_equal(blocks, exp)
    assert np.array_equal(err.cpu().numpy().view(np.uint64), e)
