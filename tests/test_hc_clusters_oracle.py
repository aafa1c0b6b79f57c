"""Pins the endpoint clustering restatement (merge, tree VQ, nearest entry;
crn_dxt_hc.cpp:836-972, 1130-1239): the merge against sorted Python tuples,
the nearest-entry rule against a literal restatement of the reference's
pruned loop, and the composed result's invariants."""
import numpy as np
import pytest

import oracle
from tests.tree_cases import tile_endpoints


@pytest.mark.parametrize("dim", [2, 6])
def test_merge_matches_sorted_tuples(dim):
    v, w = tile_endpoints(3000, dim, seed=dim)
    w[::5] = 2**31
    uv, uw = oracle.hc_merge_vectors(v, w)
    d = {}
    for row, wt in zip(map(tuple, v.tolist()), w.tolist()):
        d[row] = min(d.get(row, 0) + wt, 2**32 - 1)
    keys = sorted(d)
    assert np.array_equal(uv, np.array(keys, np.float32))
    assert np.array_equal(uw, np.array([d[k] for k in keys], np.uint32))
    assert len(uv) < len(v)          # repeats were merged


def reference_nearest(v, cb, own):
    """The reference's loop (crn_dxt_hc.cpp:843-883): partial distances
    pruned against the distance to the vector's own leaf, strict '<', stop at
    zero -- in float32."""
    f = np.float32
    node = f(0)
    for c in range(len(v)):
        t = f(cb[own][c] - v[c])
        node = f(node + f(t * t))
    best, bi = f(1e37), 0
    for i in range(len(cb)):
        dist = f(0)
        pruned = False
        for c in range(len(v)):
            t = f(cb[i][c] - v[c])
            dist = f(dist + f(t * t))
            if len(v) == 6 and c in (1, 3) and dist > node:
                pruned = True
                break
        if pruned:
            continue
        if dist < best:
            best, bi = dist, i
            if best == 0:
                break
    return bi


@pytest.mark.parametrize("dim", [2, 6])
def test_nearest_matches_pruned_loop(dim):
    v, w = tile_endpoints(400, dim, seed=10 + dim)
    uv, uw = oracle.hc_merge_vectors(v, w)
    cb, leaf = oracle.tree_codebook(uv, uw, 40)
    got = oracle.hc_nearest(v, cb)
    row = {tuple(x): l for x, l in zip(map(tuple, uv.tolist()), leaf.tolist())}
    for i in range(len(v)):
        assert got[i] == reference_nearest(v[i], cb, row[tuple(v[i].tolist())])


@pytest.mark.parametrize("dim,k", [(6, 64), (2, 32), (6, 5000)])
def test_clusters_invariants(dim, k):
    v, w = tile_endpoints(5000, dim, seed=k + dim)
    cl, cb = oracle.hc_endpoint_clusters(v, w, k)
    uv, _ = oracle.hc_merge_vectors(v, w)
    assert len(cb) <= min(k, len(uv))
    d = ((v[:, None, :].astype(np.float64) - cb[None, :, :]) ** 2).sum(-1)
    assert np.allclose(d[np.arange(len(v)), cl], d.min(1), rtol=1e-5, atol=1e-9)
    if k >= len(uv):
        # nearly every distinct vector gets its own entry; splitting stops
        # early only where float rounding leaves a node's variance at <= 0
        assert len(cb) >= 0.99 * len(uv)
