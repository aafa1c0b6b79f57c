"""GPU adaptive tile choice (crn2_hc_tiles_device) against the oracle:
encodings, tile indices, tile sources and palettized endpoints bit for bit,
for colour only, colour + alpha, two alpha channels, several levels and a
cubemap's stacked faces."""
import numpy as np
import pytest
import torch

import oracle
from crunch2_amd import texture

pytestmark = pytest.mark.gpu


def image_blocks(img):
    h, w = img.shape[:2]
    return np.ascontiguousarray(img.reshape(h // 4, 4, w // 4, 4, 4).transpose(0, 2, 1, 3, 4).reshape(-1, 16, 4))


def mip_blocks(size, seed, kind, levels):
    img = texture.synth(size, size, seed, kind)
    out, desc, first = [], [], 0
    for _ in range(levels):
        b = image_blocks(img)
        out.append(b)
        w = img.shape[1] // 4
        desc.append((first, len(b), w))
        first += len(b)
        img = img.reshape(img.shape[0] // 2, 2, img.shape[1] // 2, 2, 4).astype(np.uint16).mean((1, 3)).astype(np.uint8)
    return np.concatenate(out), desc


CASES = {
    "rgb": dict(kind=0, has_color=True, alpha=()),
    "rgba": dict(kind=1, has_color=True, alpha=(3,)),
    "dxn": dict(kind=2, has_color=False, alpha=(0, 1)),
}


@pytest.mark.parametrize("case", ["rgb", "rgba", "dxn"])
@pytest.mark.parametrize("size,levels", [(32, 1), (256, 4)])
def test_tiles_match_oracle(gpu, case, size, levels):
    from crunch2_amd import hc
    c = CASES[case]
    print("hc tiles", case, size, levels, flush=True)
    kind = {0: texture.RGB, 1: texture.RGBA, 2: texture.NORMAL}[c["kind"]]
    blocks, desc = mip_blocks(size, size + levels, kind, levels)
    d = torch.device("cuda:0")
    enc, ti, src, vec = hc.tiles(torch.from_numpy(blocks).to(d), desc, has_color=c["has_color"],
                                 alpha_comp=c["alpha"])
    torch.cuda.synchronize()
    renc, rti, rsrc, rvec = oracle.hc_tiles(blocks, desc, has_color=c["has_color"], alpha_comp=c["alpha"])
    assert np.array_equal(enc.cpu().numpy(), renc)
    assert np.array_equal(ti.cpu().numpy().view(np.uint32), rti)
    assert np.array_equal(src.cpu().numpy().view(np.uint32), rsrc)
    assert np.array_equal(vec.cpu().numpy().view(np.uint32), rvec.view(np.uint32))


def test_tiles_cubemap_faces(gpu):
    from crunch2_amd import hc
    print("hc tiles cubemap", flush=True)
    faces = [image_blocks(texture.synth(64, 64, 50 + f, texture.RGB)) for f in range(6)]
    blocks = np.concatenate(faces)       # faces stacked: 16 x 96 blocks
    desc = [(0, len(blocks), 16)]
    d = torch.device("cuda:0")
    enc, ti, src, vec = hc.tiles(torch.from_numpy(blocks).to(d), desc, num_faces=6)
    torch.cuda.synchronize()
    renc, rti, rsrc, rvec = oracle.hc_tiles(blocks, desc, num_faces=6)
    assert np.array_equal(enc.cpu().numpy(), renc)
    assert np.array_equal(ti.cpu().numpy().view(np.uint32), rti)
    assert np.array_equal(vec.cpu().numpy().view(np.uint32), rvec.view(np.uint32))
