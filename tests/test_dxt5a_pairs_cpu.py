"""Pins the DXT5 alpha search's pair stage (crn_dxt5a.cpp:93-104 with the
evaluator :198-262) independently of the oracle's code: below Better the
result is exactly the first minimum, in the order (unique pair i < j in
first-appearance order, 8-value then 6-value palette), of the error summed
over unique values of the nearest palette entry, then put in canonical
order (crn_dxt5a.cpp:151-184); at Better and Uber the box search can only
improve on it, so the reported error never exceeds it.  A reversed pair
order, a '<=' improvement or a dropped 6-value palette fails the first
check; a box that replaced the pair stage fails the second."""
import numpy as np
import pytest

import oracle
from tests import dxt5a_cases


def palette(l, h, six):
    if not six:
        return [l, h] + [(l * (7 - i) + h * i) // 7 for i in range(1, 7)]
    return [l, h] + [(l * (5 - i) + h * i) // 5 for i in range(1, 5)] + [0, 255]


def pair_stage(v, use_both):
    vals, first = np.unique(np.asarray(v, np.int64), return_index=True)
    order = np.argsort(first)
    uv = vals[order]
    w = np.array([(np.asarray(v) == x).sum() for x in uv], np.int64)
    if len(uv) == 1:
        return 0, int(uv[0]), int(uv[0]), 0
    best = None
    for i in range(len(uv) - 1):
        for j in range(i + 1, len(uv)):
            for six in ((0, 1) if use_both else (0,)):
                p = np.array(palette(int(uv[i]), int(uv[j]), six), np.int64)
                e = int(((uv[:, None] - p[None, :]) ** 2 * w[:, None]).min(1).sum())
                if best is None or e < best[0]:
                    best = (e, int(uv[i]), int(uv[j]), six)
    e, f, s, t = best
    # canonical order (crn_dxt5a.cpp:151-184)
    if f != s and ((t == 1 and f > s) or (t == 0 and f <= s)):
        f, s = s, f
    return e, f, s, t


@pytest.mark.parametrize("use_both", [True, False])
def test_pair_stage_below_better_is_exact(use_both):
    b = dxt5a_cases.blocks(7)
    b = np.concatenate([b[np.random.default_rng(0).choice(len(b) - 30, 300, replace=False)], b[-30:]])
    for quality in (0, 2):
        sel, err, f, s, t, r = oracle.dxt5a_blocks(b, quality, use_both)
        for i in range(len(b)):
            e, ef, es, et = pair_stage(b[i], use_both)
            assert int(err[i]) == e, (quality, i)
            if ef != es:
                assert (int(f[i]), int(s[i]), int(t[i])) == (ef, es, et), (quality, i)


@pytest.mark.parametrize("quality", [3, 4])
def test_box_search_never_worse_than_pairs(quality):
    b = dxt5a_cases.blocks(8)
    b = b[np.random.default_rng(1).choice(len(b), 300, replace=False)]
    sel, err, f, s, t, r = oracle.dxt5a_blocks(b, quality, True)
    better = 0
    for i in range(len(b)):
        e = pair_stage(b[i], True)[0]
        assert int(err[i]) <= e, i
        better += int(err[i]) < e
    assert better > 0
