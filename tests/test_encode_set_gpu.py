"""Texture-set encode (crn2_encode_set_*): every item's DXTn blocks equal the
per-image encoder's (crn2_encode_image_device, itself pinned to the oracle)
for a mixed set of DXT1 / DXT5 / DXT5A / DXN images of uneven sizes; the same
after graph capture on several replays with other work between them; the
autotuned launch shapes give the same bytes; DXT1 items against the oracle's
optimizer."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle
from crunch2_amd import dxt1, texture
from crunch2_amd._lib import FMT_DXN_XY, FMT_DXT1, FMT_DXT5, FMT_DXT5A, FMT_DXN_YX, EncodeItem, lib

pytestmark = pytest.mark.gpu
KIND = {FMT_DXT1: texture.RGB, FMT_DXT5: texture.RGBA, FMT_DXT5A: texture.RGBA, FMT_DXN_XY: texture.NORMAL,
        FMT_DXN_YX: texture.NORMAL}


def make_set(seed, n):
    rng = np.random.default_rng(seed)
    imgs = []
    fmts = [FMT_DXT1, FMT_DXT5, FMT_DXT5A, FMT_DXN_XY, FMT_DXN_YX]
    for i in range(n):
        side = int(round(2 ** rng.uniform(2, 7)))
        w, h = side, max(1, side + int(rng.integers(-3, 4)))
        fmt = fmts[i % len(fmts)]
        for lv in texture.mip_chain(texture.synth(w, h, 100 + i, KIND[fmt]), 3):
            imgs.append((fmt, lv))
    return imgs


def per_image(fmt, img, quality, flags):
    h, w = img.shape[:2]
    bpb = 8 if fmt in (FMT_DXT1, FMT_DXT5A) else 16
    d = torch.from_numpy(np.ascontiguousarray(img).reshape(-1)).cuda()
    out = torch.zeros(((w + 3) // 4) * ((h + 3) // 4) * bpb, dtype=torch.uint8, device="cuda")
    assert lib().crn2_encode_image_device(C.c_void_p(d.data_ptr()), w, h, fmt, quality, flags,
                                          C.c_void_p(out.data_ptr()), None) == 0
    torch.cuda.synchronize()
    return out.cpu().numpy()


@pytest.mark.parametrize("tune", [0, 1])
def test_set_matches_per_image_encoder(gpu, tune):
    imgs = make_set(3 + tune, 14)
    quality, flags = 4, 1 | 2
    print("encode set of", len(imgs), "images, autotune", tune, flush=True)
    srcs = [torch.from_numpy(np.ascontiguousarray(im).reshape(-1)).cuda() for _, im in imgs]
    outs = []
    items = (EncodeItem * len(imgs))()
    for k, ((fmt, im), src) in enumerate(zip(imgs, srcs)):
        h, w = im.shape[:2]
        bpb = 8 if fmt in (FMT_DXT1, FMT_DXT5A) else 16
        outs.append(torch.full((((w + 3) // 4) * ((h + 3) // 4) * bpb,), 0xEE, dtype=torch.uint8, device="cuda"))
        items[k] = EncodeItem(src.data_ptr(), w, h, fmt, outs[-1].data_ptr())
    lib().crn2_autotune_mode(tune)
    try:
        s = C.c_void_p()
        assert lib().crn2_encode_set_create(items, len(imgs), quality, flags, C.byref(s)) == 0
        st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
        assert lib().crn2_encode_set_run(s, st) == 0
        torch.cuda.synchronize()
        exp = [per_image(fmt, im, quality, flags) for fmt, im in imgs]
        for k in range(len(imgs)):
            assert np.array_equal(outs[k].cpu().numpy(), exp[k]), k
        for rep in range(3):
            for o in outs:
                o.fill_(0x11)                 # eager work between the replays
            assert lib().crn2_encode_set_run_graph(s, st) == 0
            torch.cuda.synchronize()
            for k in range(len(imgs)):
                assert np.array_equal(outs[k].cpu().numpy(), exp[k]), (rep, k)
        ce, ae, cs, as_ = C.c_uint64(), C.c_uint64(), C.c_int(), C.c_int()
        assert lib().crn2_encode_set_info(s, C.byref(ce), C.byref(ae), C.byref(cs), C.byref(as_)) == 0
        print("elements", ce.value, ae.value, "shapes", cs.value, as_.value, flush=True)
        assert 0 <= cs.value < 6 and 0 <= as_.value < 6
        assert lib().crn2_encode_set_destroy(s) == 0
    finally:
        lib().crn2_autotune_mode(0)
    # DXT1 items against the oracle's optimizer
    for k, (fmt, im) in enumerate(imgs):
        if fmt != FMT_DXT1 or k % 3:
            continue
        sel, err, lo, hi = oracle.dxt1_blocks(dxt1.image_blocks(im), quality,
                                              oracle.DXT1_PERCEPTUAL | oracle.DXT1_ALPHA_BLOCKS)[:4]
        w = exp[k].view(np.uint16).reshape(-1, 4)
        assert np.array_equal(w[:, 0], lo) and np.array_equal(w[:, 1], hi), k
