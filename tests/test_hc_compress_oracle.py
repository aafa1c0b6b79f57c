"""The oracle's dxt_hc composition (crn_dxt_hc.cpp:97-316) end to end:
palettes without repeats, indices in range, reference codes that say what
they mean, and decoded texels close to the source."""
import numpy as np
import pytest

import oracle
from crunch2_amd import texture
from tests.hc_compress_cases import texture_blocks
from tests.hc_decode import decode_alpha, decode_color, psnr


@pytest.mark.parametrize("fmt", ["dxt1", "dxt5", "dxn"])
def test_compress_end_to_end(fmt):
    kind = {"dxt1": texture.RGB, "dxt5": texture.RGBA, "dxn": texture.NORMAL}[fmt]
    alpha = {"dxt1": (), "dxt5": (3,), "dxn": (0, 1)}[fmt]
    has_color = fmt != "dxn"
    blocks, levels = texture_blocks(64, 7, kind, 3)
    ep, sel, ce, cs, ae, as_ = oracle.hc_compress(blocks, levels, has_color=has_color, alpha_comp=alpha,
                                                  codebook_sizes=(128, 128, 64, 64))
    for pal in (ce, cs, ae, as_):
        if pal is not None:
            assert len(np.unique(pal)) == len(pal)
    if has_color:
        assert ep[:, 0].max() < len(ce) and sel[:, 0].max() < len(cs) and len(ce) <= 128 and len(cs) <= 128
        rec = decode_color(ep, sel, ce, cs)
        assert psnr(rec, blocks[:, :, :3]) > 28.0
    for a, c in enumerate(alpha):
        assert ep[:, 1 + a].max() < len(ae) and sel[:, 1 + a].max() < len(as_)
        rec = decode_alpha(ep, sel, a, ae, as_)
        assert psnr(rec, blocks[:, :, c]) > 28.0
    # reference codes: 1 = same endpoints as the previous block, 2 = as the block above
    comps = ([0] if has_color else []) + [1 + a for a in range(len(alpha))]
    for first, nb, w, _ in levels:
        for r in range(nb):
            b = first + r
            code = ep[b, 3]
            same_left = (r > 0 or False) and all(ep[b, c] == ep[b - 1, c] for c in comps) and (r // w > 0 or r % w > 0)
            same_top = r // w > 0 and all(ep[b, c] == ep[b - w, c] for c in comps)
            assert code == (1 if same_left else 2 if same_top else 0)
