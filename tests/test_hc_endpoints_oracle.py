"""Pins the dxt_hc endpoint codebook / refiner oracle (oracle/hc_endpoints.c)
with checks that do not come from it: exact recovery of palettes that are
DXT1 already, the refiner's error recomputed from its endpoints with fixed
selectors, nearest-colour selectors and weights recomputed in numpy."""
import numpy as np
import pytest

import oracle
from tests.hc_endpoint_cases import clusters

W = np.array([8, 25, 1], np.int64)
TO_LIN = [0, 3, 1, 2]
FROM_LIN = [0, 2, 3, 1]


def unpack(c):
    r, g, b = (c >> 11) & 31, (c >> 5) & 63, c & 31
    return np.array([(r << 3) | (r >> 2), (g << 2) | (g >> 4), (b << 3) | (b >> 2)], np.int64)


def colors4(c0, c1):
    a, b = unpack(c0), unpack(c1)
    return np.stack([a, b, (2 * a + b) // 3, (2 * b + a) // 3])


def fixed_error(px, sel, lo, hi, perceptual=True):
    pal = colors4(lo, hi)
    d = (px[:, :3].astype(np.int64) - pal[sel]) ** 2
    return int((d * (W if perceptual else 1)).sum())


@pytest.mark.parametrize("seed", range(6))
def test_refiner_recovers_dxt1_palettes(seed):
    rng = np.random.default_rng(seed)
    hi = int(rng.integers(0, 0xFFFF))
    lo = int(rng.integers(hi + 1, 0x10000))
    n = int(rng.integers(16, 300))
    sel = rng.integers(0, 4, n).astype(np.uint8)
    sel[:4] = [0, 1, 2, 3]
    px = np.zeros((n, 4), np.uint8)
    px[:, :3] = colors4(lo, hi)[sel]
    ok, l, h, err = oracle.hc_refine_dxt1(px, sel)
    assert ok and err == 0
    assert np.array_equal(colors4(l, h), colors4(lo, hi))


@pytest.mark.parametrize("perceptual", [True, False])
def test_refiner_error_is_fixed_selector_error(perceptual):
    px, ofs, *_ = clusters(seed=1)
    rng = np.random.default_rng(2)
    for c in range(0, len(ofs) - 1, 3):
        p = px[ofs[c]:ofs[c + 1]]
        if not len(p):
            continue
        sel, r = oracle.dxt1_cluster(p, 4, oracle.DXT1_PERCEPTUAL if perceptual else 0)
        ok, l, h, err = oracle.hc_refine_dxt1(p, sel, perceptual, r.error)
        assert err == fixed_error(p, sel, l, h, perceptual)
        assert ok == (err < r.error)
        # random selectors: still the fixed-selector error of what it returns
        s2 = rng.integers(0, 4, len(p)).astype(np.uint8)
        _, l2, h2, e2 = oracle.hc_refine_dxt1(p, s2, perceptual)
        assert e2 == fixed_error(p, s2, l2, h2, perceptual)


def test_codebook_selectors_and_weights():
    px, ofs, bc, blocks, w, enc = clusters(seed=3)
    ends, cv, bsel = oracle.hc_color_endpoints(px, ofs, bc, blocks, w, enc)
    k = len(ofs) - 1
    f32 = np.float32
    ew = f32(1.15) + (f32(1.0) - f32(1.15)) * (np.arange(8, dtype=f32) / f32(7.0))   # lerp in float
    checked = 0
    for c in range(k):
        p = px[ofs[c]:ofs[c + 1]]
        mine = np.nonzero(bc == c)[0]
        if not len(p):
            assert (bsel[mine] == 0).all()
            continue
        sel, r = oracle.dxt1_cluster(p, 4, oracle.DXT1_PERCEPTUAL)
        pal = colors4(r.low, r.high)
        lin = pal[FROM_LIN]
        assert np.array_equal(cv[c][:, :3], lin)
        use = lin.copy()
        if r.alternate_rounding:
            use[1] = ((use[0] << 1) + use[3] + 1) // 3
            use[2] = ((use[3] << 1) + use[0] + 1) // 3
        ewt = int((((use[0] - use[3]) ** 2) * W).sum()) // 2000
        order = [(3 - TO_LIN[t]) if r.reordered else TO_LIN[t] for t in range(4)]
        for b in mine:
            d = (((blocks[b][:, None, :3].astype(np.int64) - use[order][None]) ** 2) * W).sum(2)
            s = np.array(order)[np.argmin(d, 1)]
            word = 0
            for v in s:
                word = (word << 2) | int(v)
            w0 = int(np.float32(ewt) * w[b])
            wt = int(np.float32(min(max(w0, 1), 2048)) * np.float32(ew[enc[b]]))
            assert int(bsel[b]) == (word << 32) | wt, (c, b)
            checked += 1
        # the refiner's endpoints replace the search's only when they are better
        ok, l, h, e = oracle.hc_refine_dxt1(p, sel, True, r.error)
        assert ends[c] == ((l | h << 16) if ok else (r.low | r.high << 16))
    assert checked > 900


TO_LIN5 = [0, 7, 1, 2, 3, 4, 5, 6]
FROM_LIN5 = [0, 2, 3, 4, 5, 6, 7, 1]


def values8(l, h):
    return np.array([l, h] + [(l * (8 - j) + h * (j - 1)) // 7 for j in range(2, 8)], np.int64)


@pytest.mark.parametrize("seed", range(6))
def test_dxt5_refiner_recovers_palettes(seed):
    rng = np.random.default_rng(10 + seed)
    h = int(rng.integers(0, 250))
    l = int(rng.integers(h + 2, 256))
    n = int(rng.integers(16, 400))
    sel = rng.integers(0, 8, n).astype(np.uint8)
    sel[:8] = np.arange(8)
    vals = values8(l, h)[sel].astype(np.uint8)
    ok, lo, hi, err = oracle.hc_refine_dxt5(vals, sel)
    assert ok and err == 0
    assert np.array_equal(values8(lo, hi), values8(l, h))


def test_dxt5_refiner_error_is_fixed_selector_error():
    from tests.hc_endpoint_cases import alpha_clusters
    vals, ofs, *_ = alpha_clusters(seed=1)
    rng = np.random.default_rng(4)
    for c in range(len(ofs) - 1):
        v = vals[ofs[c]:ofs[c + 1]]
        if not len(v):
            continue
        s = rng.integers(0, 8, len(v)).astype(np.uint8)
        _, lo, hi, err = oracle.hc_refine_dxt5(v, s)
        assert err == int(((v.astype(np.int64) - values8(lo, hi)[s]) ** 2).sum())


def test_alpha_codebook_selectors_and_weights():
    from tests.hc_endpoint_cases import alpha_clusters
    vals, ofs, bc, blocks, enc = alpha_clusters(seed=2)
    ends, av, rv, bsel = oracle.hc_alpha_endpoints(vals, ofs, bc, blocks, enc)
    f32 = np.float32
    checked = 0
    for c in range(len(ofs) - 1):
        v = vals[ofs[c]:ofs[c + 1]]
        mine = np.nonzero(bc == c)[0]
        if not len(v):
            assert (bsel[mine] == 0).all() and ends[c] == 0
            continue
        sel, err, first, second, _, reord = oracle.dxt5a_values(v)
        assert np.array_equal(av[c].astype(np.int64), values8(first, second)[FROM_LIN5] if first > second else
                              np.array([first, second] + [(first * (6 - j) + second * (j - 1)) // 5 for j in range(2, 6)]
                                       + [0, 255], np.int64)[FROM_LIN5])
        ok, lo, hi, rerr = oracle.hc_refine_dxt5(v, sel, err)
        f, s2 = (lo, hi) if ok else (first, second)
        assert int(ends[c]) == f | s2 << 8 | int(ok) << 16
        fv = values8(f, s2) if f > s2 else np.array([f, s2] + [(f * (6 - j) + s2 * (j - 1)) // 5 for j in range(2, 6)]
                                                    + [0, 255], np.int64)
        assert np.array_equal(rv[c].astype(np.int64), fv[FROM_LIN5])
        a = av[c].astype(np.int64)
        d = first - second
        ew = min(max((d * d) >> 3, 1), 2048)
        order = [(7 - TO_LIN5[t]) if reord else TO_LIN5[t] for t in range(8)]
        for b in mine:
            dist = np.abs(blocks[b][:, None].astype(np.int64) - a[np.array(order)][None, :])
            s = np.array(order)[np.argmin(dist, 1)]
            word = 0
            for x in s:
                word = (word << 3) | int(x)
            wt = int(f32(ew) * (f32(1.15) + (f32(1.0) - f32(1.15)) * (f32(enc[b]) / f32(7.0))))
            assert int(bsel[b]) == (word << 16) | wt, (c, b)
            checked += 1
    assert checked > 900
