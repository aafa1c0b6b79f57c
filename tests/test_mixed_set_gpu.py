"""The serving-style set at its full size: 512 textures, sides log-uniform in
[64, 2048], full mip chains, 60 % DXT1 / 25 % DXT5 / 15 % DXN_XY, packed by
the GPU dxt_hc and the project's writer.  Every texture's every level from
one batch transcode plan (crn2_batch_*) equals the oracle's transcode, bit
for bit; the whole set re-encoded by one texture-set encode
(crn2_encode_set_*, graph replay) gives every sampled image the blocks of the
per-image encoder, and DXT1 images the blocks of the oracle's optimizer."""
import ctypes as C
import time

import numpy as np
import pytest
import torch

import oracle
from crunch2_amd import crnd, dxt1, texture
from crunch2_amd._lib import FMT_DXN_XY, FMT_DXT1, FMT_DXT5, EncodeItem, lib

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def mixed(gpu):
    t = time.time()
    s = texture.mixed_set(512, writer="hc")
    print("512-texture set built in %.1fs, %d bytes" % (time.time() - t, sum(len(x[2]) for x in s)), flush=True)
    return s


def test_mixed_set_transcode_matches_oracle(mixed):
    ctxs, outs = [], []
    for _, _, data in mixed:
        ctx = crnd.unpack_begin(data)
        assert ctx is not None
        n = crnd.levels_size(ctx, 0, crnd.get_texture_info(data)["levels"])
        ctxs.append(ctx)
        outs.append(torch.zeros(n, dtype=torch.uint8, device="cuda"))
    k = len(ctxs)
    plan = C.c_void_p()
    assert lib().crn2_batch_create((C.c_void_p * k)(*[c.handle for c in ctxs]), k,
                                   (C.c_void_p * k)(*[o.data_ptr() for o in outs]),
                                   (C.c_uint64 * k)(*[o.numel() for o in outs]), C.byref(plan)) == 0
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib().crn2_batch_run(plan, st) == 0
    assert lib().crn2_batch_status(plan, st) == 0
    torch.cuda.synchronize()
    texels = 0
    for i, (fmt, side, data) in enumerate(mixed):
        ref = oracle.unpack_all(data)
        exp = np.concatenate([ref[l][0].reshape(-1) for l in sorted(ref)])
        assert np.array_equal(outs[i].cpu().numpy(), exp), (i, fmt, side)
        texels += sum(((max(1, side >> l) + 3) // 4) ** 2 * 16 for l in sorted(ref))
    print("transcoded %d textures, %.1f Mtexels, bit-exact" % (len(mixed), texels / 1e6), flush=True)
    assert lib().crn2_batch_destroy(plan) == 0
    for c in ctxs:
        crnd.unpack_end(c)


def test_mixed_set_encode_sweep(mixed):
    kinds = {FMT_DXT1: texture.RGB, FMT_DXT5: texture.RGBA, FMT_DXN_XY: texture.NORMAL}
    rng = np.random.default_rng(2024)
    images, items, outs = [], [], []
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=16) as pool:     # the texture synthesis runs in libcrn2, outside the GIL
        chains = list(pool.map(lambda j: texture.mip_chain(texture.synth(j[1][1], j[1][1], 5000 + j[0],
                                                                         kinds[j[1][0]])), enumerate(mixed)))
    for i, ((fmt, side, _), chain) in enumerate(zip(mixed, chains)):
        for lv in chain:
            src = torch.from_numpy(np.ascontiguousarray(lv).reshape(-1)).cuda()
            h, w = lv.shape[:2]
            bpb = 8 if fmt == FMT_DXT1 else 16
            out = torch.zeros(((w + 3) // 4) * ((h + 3) // 4) * bpb, dtype=torch.uint8, device="cuda")
            images.append((i, fmt, lv))
            items.append((src, out))
            outs.append(out)
    arr = (EncodeItem * len(items))()
    for k, ((i, fmt, lv), (src, out)) in enumerate(zip(images, items)):
        arr[k] = EncodeItem(src.data_ptr(), lv.shape[1], lv.shape[0], fmt, out.data_ptr())
    s = C.c_void_p()
    assert lib().crn2_encode_set_create(arr, len(items), 4, 1 | 2, C.byref(s)) == 0
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    t = time.time()
    assert lib().crn2_encode_set_run_graph(s, st) == 0
    torch.cuda.synchronize()
    print("encoded %d level images in %.2fs" % (len(items), time.time() - t), flush=True)
    # a sample of images against the per-image encoder; small DXT1 ones against the oracle
    pick = sorted(set(rng.integers(0, len(images), 40).tolist()))
    for k in pick:
        i, fmt, lv = images[k]
        h, w = lv.shape[:2]
        src = items[k][0]
        ref = torch.zeros_like(outs[k])
        assert lib().crn2_encode_image_device(C.c_void_p(src.data_ptr()), w, h, fmt, 4, 1 | 2,
                                              C.c_void_p(ref.data_ptr()), None) == 0
        torch.cuda.synchronize()
        got = outs[k].cpu().numpy()
        assert np.array_equal(got, ref.cpu().numpy()), (k, fmt, w, h)
        if fmt == FMT_DXT1 and w * h <= 128 * 128:
            sel, err, lo, hi = oracle.dxt1_blocks(dxt1.image_blocks(lv), 4,
                                                  oracle.DXT1_PERCEPTUAL | oracle.DXT1_ALPHA_BLOCKS)[:4]
            wv = got.view(np.uint16).reshape(-1, 4)
            assert np.array_equal(wv[:, 0], lo) and np.array_equal(wv[:, 1], hi), (k, w, h)
    assert lib().crn2_encode_set_destroy(s) == 0
