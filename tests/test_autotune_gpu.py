"""The launch-shape autotuner: with tuning on, every kernel family gives the
same results as its default shape, and a choice is recorded per size class."""
import ctypes as C

import numpy as np
import pytest
import torch

from crunch2_amd import dxt1, dxt5a, hc, texture
from crunch2_amd._lib import lib
from tests.hc_cases import make_case

pytestmark = pytest.mark.gpu


def choice(kernel, work):
    sh, ms = C.c_int(), C.c_float()
    assert lib().crn2_autotune_choice(kernel, work, C.byref(sh), C.byref(ms)) == 0
    return sh.value, ms.value


def test_tuned_shapes_give_identical_results(gpu):
    blocks = dxt1.image_blocks(texture.synth(256, 256, 8, texture.RGB))
    tb = torch.from_numpy(blocks).cuda()
    b, ei, cols, sel = make_case(3000, 50, 5000, seed=9)
    args = [torch.from_numpy(x).cuda() for x in (b, ei.view(np.int32), cols, sel.view(np.int32))]
    vals = torch.from_numpy(np.ascontiguousarray(blocks[:, :, 3])).cuda()
    try:
        assert lib().crn2_autotune_mode(0) == 0
        d0 = dxt1.encode_blocks_device(tb).cpu().numpy()
        h0 = [t.cpu().numpy() for t in hc.color_selectors(*args)]
        a0 = dxt5a.compute_blocks(vals.cpu().numpy())
        assert lib().crn2_autotune_mode(1) == 0
        d1 = dxt1.encode_blocks_device(tb).cpu().numpy()
        h1 = [t.cpu().numpy() for t in hc.color_selectors(*args)]
        a1 = dxt5a.compute_blocks(vals.cpu().numpy())
        print("dxt1 shape", choice(1, len(blocks)), "hc shape", choice(0, 3000 * 5000),
              "dxt5a shape", choice(2, len(blocks)), flush=True)
    finally:
        lib().crn2_autotune_mode(0)
    assert np.array_equal(d0, d1)
    for x, y in zip(h0, h1):
        assert np.array_equal(x, y)
    for x, y in zip(a0, a1):
        assert np.array_equal(x, y)
