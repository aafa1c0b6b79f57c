"""The host half of GPU mip generation (libcrn2's contributor lists, gamma
tables and level counts, built on the CPU before any launch) equals the
oracle's restatement of crnlib's Resampler bit for bit, for every filter, edge
mode, filter scale and the minification and magnification branches.  No GPU
call is made."""
import numpy as np
import pytest

import oracle
from crunch2_amd import _lib, mips


@pytest.mark.parametrize("filt", range(5))
@pytest.mark.parametrize("wrap", [False, True])
def test_contributor_lists_match_oracle(filt, wrap):
    checked = 0
    for src, dst in [(64, 32), (64, 16), (37, 18), (37, 1), (1, 1), (5, 2), (8192, 1), (300, 150), (4, 1),
                     (2, 1), (4096, 2048), (16384, 1), (3, 7), (2, 9)]:
        for fs in (0.9, 1.0, 0.6, 1.7):
            try:
                a = oracle.mip_contributors(src, dst, filt, fs, wrap)
            except AssertionError:
                # the reference's make_clist fails here too (zero filter sum)
                with pytest.raises(RuntimeError):
                    mips.contributors(src, dst, filt, fs, wrap)
                continue
            b = mips.contributors(src, dst, filt, fs, wrap)
            for x, y in zip(a, b):
                assert x.shape == y.shape and np.array_equal(x.view(np.uint32), y.view(np.uint32)), (src, dst, fs)
            checked += 1
    assert checked > 40


@pytest.mark.parametrize("gamma", [2.2, 1.75, 1.0])
def test_gamma_tables_match_oracle(gamma):
    s1, l1 = oracle.mip_tables(gamma)
    s2, l2 = mips.gamma_tables(gamma)
    assert np.array_equal(s1.view(np.uint32), s2.view(np.uint32)) and np.array_equal(l1, l2)


def test_level_counts_match_oracle():
    L = _lib.lib()
    for w, h in [(1, 1), (2, 1), (256, 256), (300, 20), (4096, 1), (8192, 8192), (16384, 3), (5, 7)]:
        for ms in (1, 2, 4, 16):
            for mx in (0, 1, 3, 16):
                assert L.crn2_mip_level_count(w, h, ms, mx) == oracle.mip_level_count(w, h, ms, mx)
    assert L.crn2_mip_level_count(0, 4, 1, 0) == 0


def test_plan_rejects_bad_parameters_before_any_device_work():
    import ctypes as C
    L = _lib.lib()
    h = C.c_void_p()
    for w, hh, p in [(0, 4, mips.params()), (16385, 4, mips.params()), (4, 4, mips.params(filter=5)),
                     (4, 4, mips.params(filter_scale=0.0)), (4, 4, mips.params(source_gamma=-1.0)),
                     (4, 4, mips.params(flags=64))]:
        assert L.crn2_mip_plan_create(w, hh, C.byref(p), C.byref(h)) == _lib.CRN2_EINVAL
    assert L.crn2_mip_plan_run(None, None, None, 0, None) == _lib.CRN2_EINVAL
