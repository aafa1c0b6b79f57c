"""Seeded inputs for the tree clusterizer tests."""
import numpy as np


def continuous(n, dim, seed, blobs=12):
    """Continuous vectors around a few random centres with random integer
    weights: no exact or near ties in any distance."""
    rng = np.random.default_rng(seed)
    centres = rng.random((blobs, dim))
    lab = rng.integers(0, blobs, n)
    v = (centres[lab] + rng.normal(0, 0.08, (n, dim))).astype(np.float32)
    w = rng.integers(1, 1000, n).astype(np.uint32)
    return v, w


def selector_vectors(n, seed, bits=2):
    """Distinct selector vectors as dxt_hc builds them ((s + 0.5) / 2^bits per
    texel) from a few prototypes with random texel flips, with merged
    weights in a wide range (large sums included)."""
    rng = np.random.default_rng(seed)
    k = 1 << bits
    protos = rng.integers(0, k, (24, 16))
    sel = protos[rng.integers(0, 24, 3 * n)]
    flip = rng.random(sel.shape) < 0.15
    sel = np.where(flip, rng.integers(0, k, sel.shape), sel)
    sel = np.unique(sel, axis=0)[:n]
    v = ((sel + 0.5) / k).astype(np.float32)
    w = rng.integers(1, 5000, len(v)).astype(np.uint32)
    w[::17] = rng.integers(1 << 20, 1 << 31, len(w[::17]))
    return v, w


def separated(n, dim, seed, blobs):
    """Tight blobs far apart: every split is decided by a wide margin (the
    model comparison at sizes past one 4096-position chunk)."""
    rng = np.random.default_rng(seed)
    centres = rng.random((blobs, dim)) * 4
    lab = rng.integers(0, blobs, n)
    v = (centres[lab] + rng.uniform(-0.02, 0.02, (n, dim))).astype(np.float32)
    w = rng.integers(1, 1000, n).astype(np.uint32)
    return v, w


def tile_endpoints(n, dim, seed):
    """Endpoint vectors as dxt_hc's tiles produce them: quantized colour
    coordinates (multiples of 1/255, halved / quartered for red / blue in the
    colour case) with many exact repeats, and tile weights."""
    rng = np.random.default_rng(seed)
    base = rng.integers(0, 256, (max(4, n // 6), dim))
    q = base[rng.integers(0, len(base), n)] + rng.integers(-3, 4, (n, dim)) * (rng.random((n, 1)) < 0.5)
    v = (np.clip(q, 0, 255) * np.float32(1.0 / 255.0)).astype(np.float32)
    if dim == 6:
        v[:, 0] *= np.float32(0.5)
        v[:, 3] *= np.float32(0.5)
        v[:, 2] *= np.float32(0.25)
        v[:, 5] *= np.float32(0.25)
    w = rng.integers(16, 257, n).astype(np.uint32)
    return v, w
