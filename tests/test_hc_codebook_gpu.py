"""GPU dxt_hc selector codebooks (merge, tree VQ, quantize, INT8 MFMA
assignment, refinement) against the oracle's composition: codebook, block
indices and used flags bit for bit."""
import numpy as np
import pytest
import torch

import oracle
from tests.hc_cases import alpha_codebook_case, color_codebook_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,k,size,optimal", [(300, 8, 4096, True), (3000, 40, 256, False), (9000, 64, 1024, True)])
def test_color_codebook_matches_oracle(gpu, n, k, size, optimal):
    from crunch2_amd import hc
    print("hc colour codebook", n, k, size, optimal, flush=True)
    blocks, ei, cols, words, _ = color_codebook_case(n, k, seed=n + size, optimal=optimal)
    d = torch.device("cuda:0")
    sel, idx, used = hc.color_selector_codebook(torch.from_numpy(words.view(np.int64)).to(d), size,
                                                torch.from_numpy(blocks).to(d),
                                                torch.from_numpy(ei.view(np.int32)).to(d), torch.from_numpy(cols).to(d))
    torch.cuda.synchronize()
    rs, ri, ru = oracle.hc_color_selector_codebook(words, size, blocks, ei, cols)
    assert np.array_equal(sel.cpu().numpy().view(np.uint32), rs)
    assert np.array_equal(idx.cpu().numpy().view(np.uint32), ri)
    assert np.array_equal(used.cpu().numpy(), ru)


@pytest.mark.parametrize("n,k,size,ch,optimal", [(200, 5, 4096, 1, True), (2500, 30, 300, 2, False),
                                                 (6000, 50, 2048, 1, True)])
def test_alpha_codebook_matches_oracle(gpu, n, k, size, ch, optimal):
    from crunch2_amd import hc
    print("hc alpha codebook", n, k, size, ch, optimal, flush=True)
    vals, ei, pal, pal2, words, _ = alpha_codebook_case(n, k, seed=n + ch, channels=ch, optimal=optimal)
    d = torch.device("cuda:0")
    sel, idx, used = hc.alpha_selector_codebook(torch.from_numpy(words.view(np.int64)).to(d), size,
                                                torch.from_numpy(vals).to(d), torch.from_numpy(ei.view(np.int32)).to(d),
                                                torch.from_numpy(pal).to(d), torch.from_numpy(pal2).to(d))
    torch.cuda.synchronize()
    rs, ri, ru = oracle.hc_alpha_selector_codebook(words, size, vals, ei, pal, pal2)
    assert np.array_equal(sel.cpu().numpy().view(np.uint64), rs)
    assert np.array_equal(idx.cpu().numpy().view(np.uint32), ri)
    assert np.array_equal(used.cpu().numpy(), ru)
