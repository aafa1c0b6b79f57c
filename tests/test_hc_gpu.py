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


def run_alpha_gpu(vals, ei, pal, sel, scatter):
    from crunch2_amd import hc
    d = torch.device("cuda:0")
    det = torch.zeros((sel.shape[0], 128), dtype=torch.int32, device=d)
    used = torch.zeros(sel.shape[0], dtype=torch.uint8, device=d)
    sv = torch.from_numpy(scatter).to(d) if scatter is not None else None
    best, err = hc.alpha_selectors(torch.from_numpy(vals).to(d), torch.from_numpy(ei.view(np.int32)).to(d),
                                   torch.from_numpy(pal).to(d), torch.from_numpy(sel.view(np.int64)).to(d), sv,
                                   det, used)
    torch.cuda.synchronize()
    return (best.cpu().numpy().view(np.uint32), err.cpu().numpy().view(np.uint32),
            det.cpu().numpy().view(np.uint32), used.cpu().numpy())


@pytest.mark.parametrize("n,k,m", [(1, 1, 1), (45, 4, 31), (700, 50, 1500), (3000, 200, 9000)])
@pytest.mark.parametrize("scatter", [False, True])
def test_alpha_matches_oracle(gpu, n, k, m, scatter):
    from tests.hc_cases import make_alpha_case
    print("hc alpha case", n, k, m, scatter, flush=True)
    vals, ei, pal, pal2, sel = make_alpha_case(n, k, m, seed=n + m + scatter)
    sv = pal2 if scatter else None
    b, e, d, u = run_alpha_gpu(vals, ei, pal, sel, sv)
    rb, re_, rd, ru = oracle.hc_alpha_selectors(vals, ei, pal, sel, sv)
    assert np.array_equal(b, rb)
    assert np.array_equal(e, re_)
    assert np.array_equal(d, rd)
    assert np.array_equal(u, ru)


def test_alpha_extremes(gpu):
    vals = np.full((40, 16), 255, np.uint8)
    vals[20:] = 0
    pal = np.zeros((2, 8), np.uint8)
    pal[1] = 255
    ei = np.zeros(40, np.uint32)
    sel = np.array([0o7777777777777777, 0, 0], np.uint64)
    b, e, d, u = run_alpha_gpu(vals, ei, pal, sel, None)
    rb, re_, rd, ru = oracle.hc_alpha_selectors(vals, ei, pal, sel)
    assert np.array_equal(b, rb) and np.array_equal(e, re_) and np.array_equal(d, rd)
    assert e.max() == 16 * 255 * 255


@pytest.mark.parametrize("bits", [2, 3])
def test_selector_refine_matches_oracle(gpu, bits):
    from crunch2_amd import hc
    rng = np.random.default_rng(10 + bits)
    m = 5000
    det = rng.integers(0, 40, (m, 16 << bits)).astype(np.uint32)
    det[::3] = rng.integers(0, 2**32, (len(det[::3]), 16 << bits), dtype=np.uint64).astype(np.uint32)
    out = hc.selector_refine(torch.from_numpy(det.view(np.int32)).cuda(), bits)
    torch.cuda.synchronize()
    got = out.cpu().numpy().view(np.uint32 if bits == 2 else np.uint64)
    assert np.array_equal(got, oracle.hc_selector_refine(det, bits))
