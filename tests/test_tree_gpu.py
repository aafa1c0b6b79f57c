"""GPU tree clusterizer (crn2_tree_codebook_device) against the oracle: the
codebook bit for bit and the leaf of every vector, on continuous, discrete
selector and separated data, across chunk boundaries, and on tiny and
degenerate inputs."""
import numpy as np
import pytest
import torch

import oracle
from tests.tree_cases import continuous, selector_vectors, separated

pytestmark = pytest.mark.gpu


def run_gpu(v, w, k):
    from crunch2_amd import hc
    d = torch.device("cuda:0")
    cb, leaf = hc.tree_codebook(torch.from_numpy(v).to(d), torch.from_numpy(w.view(np.int32)).to(d), k)
    torch.cuda.synchronize()
    return cb.cpu().numpy(), leaf.cpu().numpy().view(np.uint32)


def check(v, w, k):
    cb, leaf = run_gpu(v, w, k)
    cb2, leaf2 = oracle.tree_codebook(v, w, k)
    assert cb.shape == cb2.shape
    assert np.array_equal(cb.view(np.uint32), cb2.view(np.uint32))
    assert np.array_equal(leaf, leaf2)


@pytest.mark.parametrize("dim,n,k", [(6, 300, 16), (16, 2000, 64), (6, 5000, 300), (16, 4097, 128)])
def test_continuous(gpu, dim, n, k):
    print("tree continuous", dim, n, k, flush=True)
    v, w = continuous(n, dim, seed=7 * n + dim)
    check(v, w, k)


@pytest.mark.parametrize("bits,n,k", [(2, 3000, 256), (2, 12000, 1024), (3, 9000, 512)])
def test_selector_vectors(gpu, bits, n, k):
    print("tree selectors", bits, n, k, flush=True)
    v, w = selector_vectors(n, seed=n + bits, bits=bits)
    check(v, w, k)


def test_separated_large(gpu):
    print("tree separated 60000 x 6, 2048", flush=True)
    v, w = separated(60000, 6, 3, 700)
    check(v, w, 2048)


def test_tiny_and_degenerate(gpu):
    print("tree tiny", flush=True)
    check(np.array([[0.1] * 6], np.float32), np.array([3], np.uint32), 8)
    check(np.array([[0.0] * 6, [1.0] * 6], np.float32), np.array([1, 2], np.uint32), 8)
    check(np.full((40, 16), 0.375, np.float32), np.arange(1, 41, dtype=np.uint32), 64)
    v, w = continuous(500, 16, seed=1)
    check(v, w, 1)
    check(v[:3], w[:3], 5)
