"""GPU endpoint clustering (crn2_hc_endpoint_clusters_device) against the
oracle: the codebook bit for bit and every vector's cluster."""
import numpy as np
import pytest
import torch

import oracle
from tests.tree_cases import tile_endpoints

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("dim,n,k", [(6, 1, 8), (6, 700, 64), (6, 20000, 4096), (2, 3000, 256), (2, 40000, 2048)])
def test_clusters_match_oracle(gpu, dim, n, k):
    from crunch2_amd import hc
    print("hc endpoint clusters", dim, n, k, flush=True)
    v, w = tile_endpoints(n, dim, seed=n + dim)
    d = torch.device("cuda:0")
    cl, cb = hc.endpoint_clusters(torch.from_numpy(v).to(d), torch.from_numpy(w.view(np.int32)).to(d), k)
    torch.cuda.synchronize()
    rcl, rcb = oracle.hc_endpoint_clusters(v, w, k)
    assert np.array_equal(cb.cpu().numpy().view(np.uint32), rcb.view(np.uint32))
    assert np.array_equal(cl.cpu().numpy().view(np.uint32), rcl)
