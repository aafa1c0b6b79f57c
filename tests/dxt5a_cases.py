"""Block sets for the DXT5 alpha endpoint search: random, few-valued,
gradients, constant, extremes and channels of synthetic textures."""
import numpy as np

from crunch2_amd import texture


def values8(l, h):
    return [l, h] + [(l * (7 - i) + h * i) // 7 for i in range(1, 7)]


def values6(l, h):
    return [l, h] + [(l * (5 - i) + h * i) // 5 for i in range(1, 5)] + [0, 255]


def palette(first, second):
    """DXT5 decode palette by code (crn_dxt.cpp:382-431 formulas): 8 values
    when first > second, else 6 values plus 0 and 255."""
    return values8(first, second) if first > second else values6(first, second)


def blocks(seed=0, n_random=3000):
    rng = np.random.default_rng(seed)
    out = [rng.integers(0, 256, (n_random, 16))]
    few = rng.integers(0, 256, (n_random, 3))
    out.append(few[np.arange(n_random)[:, None], rng.integers(0, 3, (n_random, 16))])
    base = rng.integers(0, 200, (500, 1))
    out.append(np.clip(base + np.arange(16)[None, :] * rng.integers(0, 4, (500, 1)), 0, 255))
    out.append(np.full((20, 16), 77))
    out.append(np.array([[0] * 8 + [255] * 8, [255] * 16, [0] * 15 + [1], [254, 255] * 8]))
    img = texture.synth(64, 64, seed + 5, texture.RGBA)
    for c in (0, 3):
        ch = img[..., c].reshape(16, 4, 16, 4).transpose(0, 2, 1, 3).reshape(-1, 16)
        out.append(ch)
    nm = texture.synth(64, 64, seed + 6, texture.NORMAL)
    out.append(nm[..., 0].reshape(16, 4, 16, 4).transpose(0, 2, 1, 3).reshape(-1, 16))
    return np.ascontiguousarray(np.concatenate(out).astype(np.uint8))
