"""GPU dxt_hc colour selector assignment (INT8 MFMA error matrix) against the
oracle: exact best index (lowest on ties), exact best error, exact details
totals and used flags."""
import numpy as np
import pytest
import torch

import oracle
from tests.hc_cases import make_case

pytestmark = pytest.mark.gpu


def run_gpu(blocks, ei, cols, sel, subblocks, perceptual):
    from crunch2_amd import hc
    d = torch.device("cuda:0")
    tb = torch.from_numpy(blocks).to(d)
    te = torch.from_numpy(ei.view(np.int32)).to(d)
    tc = torch.from_numpy(cols).to(d)
    ts = torch.from_numpy(sel.view(np.int32)).to(d)
    det = torch.zeros((sel.shape[0], 64), dtype=torch.int32, device=d)
    used = torch.zeros(sel.shape[0], dtype=torch.uint8, device=d)
    best, err = hc.color_selectors(tb, te, tc, ts, subblocks, perceptual, det, used)
    torch.cuda.synchronize()
    return (best.cpu().numpy().view(np.uint32), err.cpu().numpy().view(np.uint32),
            det.cpu().numpy().view(np.uint32), used.cpu().numpy())


@pytest.mark.parametrize("n,k,m", [(1, 1, 1), (37, 3, 33), (300, 40, 700), (2000, 64, 9000)])
@pytest.mark.parametrize("subblocks", [False, True])
@pytest.mark.parametrize("perceptual", [True, False])
def test_matches_oracle(gpu, n, k, m, subblocks, perceptual):
    print("hc case", n, k, m, subblocks, perceptual, flush=True)
    blocks, ei, cols, sel = make_case(n, k, m, seed=n + m, subblocks=subblocks)
    b, e, d, u = run_gpu(blocks, ei, cols, sel, subblocks, perceptual)
    rb, re_, rd, ru = oracle.hc_color_selectors(blocks, ei, cols, sel, subblocks, perceptual)
    assert np.array_equal(b, rb)
    assert np.array_equal(e, re_)
    assert np.array_equal(d, rd)
    assert np.array_equal(u, ru)


def test_extreme_and_ties(gpu):
    blocks = np.full((64, 16, 4), 255, np.uint8)
    blocks[32:, :, :3] = 0
    cols = np.zeros((2, 4, 4), np.uint8)
    cols[1] = 255
    ei = np.repeat(np.array([0, 1], np.uint32), 32)
    sel = np.array([7, 7, 0xFFFFFFFF, 0, 0], np.uint32)
    b, e, d, u = run_gpu(blocks, ei, cols, sel, False, True)
    rb, re_, rd, ru = oracle.hc_color_selectors(blocks, ei, cols, sel, False, True)
    assert np.array_equal(b, rb) and np.array_equal(e, re_) and np.array_equal(d, rd)
    assert e.max() == 16 * 34 * 255 * 255
