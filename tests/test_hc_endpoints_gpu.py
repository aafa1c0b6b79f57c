"""dxt_hc colour endpoint codebook on the GPU (crn2_hc_color_endpoints_device)
against the oracle: cluster endpoints after the refiner, cluster colours and
every block's selector word, bit for bit."""
import numpy as np
import pytest
import torch

import oracle
from crunch2_amd import hc
from tests.hc_endpoint_cases import clusters

pytestmark = pytest.mark.gpu


def run_gpu(px, ofs, bc, blocks, w, enc, perceptual):
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    ends, cv, bsel, ne = hc.color_endpoints(d(px), d(ofs.astype(np.int32)), d(bc.view(np.int32)), d(blocks), d(w),
                                            d(enc), perceptual)
    torch.cuda.synchronize()
    return ends.cpu().numpy().view(np.uint32), cv.cpu().numpy(), bsel.cpu().numpy().view(np.uint64), ne


@pytest.mark.parametrize("seed,perceptual,side,k", [(0, True, 128, 48), (1, False, 128, 48), (2, True, 256, 96)])
def test_endpoint_codebook_matches_oracle(gpu, seed, perceptual, side, k):
    """Every cluster, including those of more than 4,096 distinct colours
    (searched with their colour set in global memory), bit for bit."""
    print("hc endpoints seed", seed, "perceptual", perceptual, "side", side, "clusters", k, flush=True)
    px, ofs, bc, blocks, w, enc = clusters(side=side, k=k, seed=seed)
    ends, cv, bsel, ne = run_gpu(px, ofs, bc, blocks, w, enc, perceptual)
    rends, rcv, rbsel = oracle.hc_color_endpoints(px, ofs, bc, blocks, w, enc, perceptual)
    big = np.array([len(np.unique(px[ofs[c]:ofs[c + 1]].view(np.uint32))) > 4096 for c in range(k)])
    print("clusters beyond 4096 colours:", int(big.sum()), flush=True)
    assert ne == 0
    assert np.array_equal(cv, rcv)
    assert np.array_equal(ends, rends), np.nonzero(ends != rends)[0][:10]
    inb = bc != 0xFFFFFFFF
    assert np.array_equal(bsel[inb], rbsel[inb])
    assert (bsel[~inb] == 0).all()


@pytest.mark.parametrize("seed,side,k", [(0, 128, 40), (1, 256, 64)])
def test_alpha_endpoint_codebook_matches_oracle(gpu, seed, side, k):
    from tests.hc_endpoint_cases import alpha_clusters
    print("hc alpha endpoints seed", seed, "side", side, "clusters", k, flush=True)
    vals, ofs, bc, blocks, enc = alpha_clusters(side=side, k=k, seed=seed)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    ends, av, rv, bsel = hc.alpha_endpoints(d(vals), d(ofs.astype(np.int32)), d(bc.view(np.int32)), d(blocks), d(enc))
    torch.cuda.synchronize()
    ends, av, rv = ends.cpu().numpy().view(np.uint32), av.cpu().numpy(), rv.cpu().numpy()
    bsel = bsel.cpu().numpy().view(np.uint64)
    rends, rav, rrv, rbsel = oracle.hc_alpha_endpoints(vals, ofs, bc, blocks, enc)
    assert np.array_equal(av, rav)
    assert np.array_equal(ends, rends), np.nonzero(ends != rends)[0][:10]
    assert np.array_equal(rv, rrv)
    assert np.array_equal(bsel, rbsel)
