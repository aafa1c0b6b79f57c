"""Pins the tree clusterizer restatement (oracle/tree_clusterizer.c):
known answers (separated blobs, one and two vectors, identical vectors),
invariants (leaves partition the vectors, every entry is its leaf's weighted
mean, at most max_splits entries, more splits never raise the total error),
and leaf-for-leaf agreement with an independent float64 model on continuous
data."""
import numpy as np
import pytest

import oracle
from tests.tree_cases import continuous, selector_vectors, separated
from tests.tree_ref import tree_codebook_f64


def weighted_sse(v, w, cb, leaf):
    d = v.astype(np.float64) - cb[leaf].astype(np.float64)
    return float((w.astype(np.float64) * (d * d).sum(1)).sum())


@pytest.mark.parametrize("dim,k", [(6, 4), (16, 5)])
def test_separated_blobs_are_found(dim, k):
    rng = np.random.default_rng(dim)
    centres = np.eye(dim)[:k] * 4.0
    lab = rng.integers(0, k, 600)
    lab[:k] = np.arange(k)
    v = (centres[lab] + rng.uniform(-0.01, 0.01, (600, dim))).astype(np.float32)
    w = rng.integers(1, 50, 600).astype(np.uint32)
    cb, leaf = oracle.tree_codebook(v, w, k)
    assert cb.shape == (k, dim)
    # each leaf is exactly one blob, its entry the blob's weighted mean
    for j in range(k):
        members = np.flatnonzero(leaf == j)
        assert len(np.unique(lab[members])) == 1
        mean = (v[members].astype(np.float64) * w[members, None]).sum(0) / w[members].sum()
        assert np.allclose(cb[j], mean, atol=1e-5)
    assert len(np.unique(leaf)) == k


def test_tiny_inputs():
    v = np.array([[0.1] * 6], np.float32)
    cb, leaf = oracle.tree_codebook(v, np.array([3], np.uint32), 8)
    assert cb.shape == (1, 6) and np.allclose(cb[0], 0.1)
    v2 = np.array([[0.0] * 6, [1.0] * 6], np.float32)
    cb, leaf = oracle.tree_codebook(v2, np.array([1, 2], np.uint32), 8)
    assert cb.shape == (2, 6)
    assert np.array_equal(cb[leaf], v2)
    same = np.full((40, 16), 0.375, np.float32)
    cb, leaf = oracle.tree_codebook(same, np.arange(1, 41, dtype=np.uint32), 64)
    assert cb.shape == (1, 16) and (leaf == 0).all()
    cb, _ = oracle.tree_codebook(v2, np.array([1, 2], np.uint32), 1)
    assert cb.shape == (1, 6) and np.allclose(cb[0], 2 / 3, atol=1e-6)


@pytest.mark.parametrize("dim", [6, 16])
def test_invariants(dim):
    v, w = continuous(3000, dim, seed=dim)
    prev = None
    for k in (1, 2, 7, 40, 160):
        cb, leaf = oracle.tree_codebook(v, w, k)
        assert 1 <= len(cb) <= k
        assert leaf.max() == len(cb) - 1 and len(np.unique(leaf)) == len(cb)
        for j in range(len(cb)):
            m = leaf == j
            mean = (v[m].astype(np.float64) * w[m, None]).sum(0) / w[m].sum()
            assert np.allclose(cb[j], mean, rtol=1e-4, atol=1e-5)
        e = weighted_sse(v, w, cb, leaf)
        if prev is not None:
            assert e <= prev * (1 + 1e-6)
        prev = e


@pytest.mark.parametrize("dim,n,k", [(6, 2000, 64), (16, 1500, 48), (6, 4097, 256)])
def test_matches_float64_model(dim, n, k):
    """Splitting one Gaussian blob is ill-conditioned (the Lloyd stopping test
    sits near its threshold), so float32 and float64 can part ways on large
    blobs; these seeds and sizes stay clear of that."""
    v, w = continuous(n, dim, seed=n + dim)
    cb, leaf = oracle.tree_codebook(v, w, k)
    cb2, leaf2 = tree_codebook_f64(v, w, k)
    assert cb.shape == cb2.shape
    assert np.array_equal(leaf, leaf2)
    assert np.allclose(cb, cb2, rtol=1e-4, atol=1e-5)


def test_selector_vectors_cover_chunks():
    """More than one 4096-position chunk, huge merged weights, discrete
    values: the invariants still hold."""
    v, w = selector_vectors(9000, seed=4)
    assert len(v) > 4096
    cb, leaf = oracle.tree_codebook(v, w, 300)
    assert len(cb) <= 300 and len(np.unique(leaf)) == len(cb)
    for j in range(0, len(cb), 7):
        m = leaf == j
        mean = (v[m].astype(np.float64) * w[m, None]).sum(0) / w[m].sum()
        assert np.allclose(cb[j], mean, rtol=1e-3, atol=1e-4)


@pytest.mark.parametrize("dim,n,blobs", [(6, 9000, 40), (16, 12000, 64), (6, 20000, 100)])
def test_matches_float64_model_past_one_chunk(dim, n, blobs):
    """Several 4096-position chunks: leaves, their order and the entries agree
    with the float64 model when every split has a wide margin."""
    v, w = separated(n, dim, n, blobs)
    cb, leaf = oracle.tree_codebook(v, w, blobs)
    cb2, leaf2 = tree_codebook_f64(v, w, blobs)
    assert cb.shape == cb2.shape == (blobs, dim)
    assert np.array_equal(leaf, leaf2)
    assert np.allclose(cb, cb2, rtol=1e-5, atol=1e-5)
