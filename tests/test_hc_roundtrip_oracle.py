"""dxt_hc output through the CRN format: the oracle's dxt_hc palettes and
indices written by the project's CRN writer (texture.pack) and read back by
the oracle's transcoder give exactly the texels that decoding the palettes
directly gives -- the selector order (linear, texel p at bits bits*p), the
endpoint packing and the per-level block grids agree end to end."""
import numpy as np
import pytest

import oracle
from crunch2_amd import texture
from crunch2_amd._lib import FMT_DXN_XY, FMT_DXT1, FMT_DXT5
from tests.hc_compress_cases import texture_blocks
from tests.hc_decode import decode_alpha, decode_color, dxt1_texels, dxt5a_values

FORMATS = {"dxt1": (FMT_DXT1, texture.RGB, (), True), "dxt5": (FMT_DXT5, texture.RGBA, (3,), True),
           "dxn": (FMT_DXN_XY, texture.NORMAL, (0, 1), False)}


def roundtrip(fmt_name, size, nlevels, encode):
    fmt, kind, alpha, has_color = FORMATS[fmt_name]
    blocks, levels = texture_blocks(size, 21, kind, nlevels)
    ep, sel, ce, cs, ae, as_ = encode(blocks, levels, has_color, alpha)
    kw = {}
    if has_color:
        kw.update(color_endpoints=ce, color_selectors=cs, ce_idx=ep[:, 0], cs_idx=sel[:, 0])
    if alpha:
        kw.update(alpha_endpoints=ae, alpha_selectors=as_, a0e_idx=ep[:, 1], a0s_idx=sel[:, 1])
    if len(alpha) == 2:
        kw.update(a1e_idx=ep[:, 2], a1s_idx=sel[:, 2])
    data = texture.pack(size, size, nlevels, 1, fmt, **kw)
    return data, blocks, levels, ep, sel, ce, cs, ae, as_, alpha, has_color


def check(data, levels, ep, sel, ce, cs, ae, as_, alpha, has_color, unpacked):
    for l, (first, nb, w, _) in enumerate(levels):
        raw = np.frombuffer(bytes(unpacked[l][0]), np.uint8)
        bpb = 16 if alpha and has_color or len(alpha) == 2 else 8
        blk = raw.reshape(-1, bpb)
        assert blk.shape[0] == nb
        s = slice(first, first + nb)
        if has_color:
            col = blk[:, bpb - 8:]
            assert np.array_equal(dxt1_texels(col.tobytes()), decode_color(ep[s], sel[s], ce, cs))
        for a in range(len(alpha)):
            part = blk[:, 8 * a:8 * a + 8]
            assert np.array_equal(dxt5a_values(part.tobytes()), decode_alpha(ep[s], sel[s], a, ae, as_))


@pytest.mark.parametrize("fmt_name", ["dxt1", "dxt5", "dxn"])
def test_oracle_dxt_hc_through_crn(fmt_name):
    def enc(blocks, levels, has_color, alpha):
        return oracle.hc_compress(blocks, levels, has_color=has_color, alpha_comp=alpha,
                                  codebook_sizes=(256, 256, 128, 128))
    data, blocks, levels, *rest = roundtrip(fmt_name, 64, 4, enc)
    check(data, levels, *rest, oracle.unpack_all(data))
