"""The learned DXT1 workload at its full size: one 8192x8192 RGB texture
(4,194,304 blocks), uber, perceptual, four-colour blocks.  The plain GPU
search agrees with the oracle's optimizer on a random sample of the blocks;
every learned result's error is its blocks' exact error; the default
(8 % fallback) stays within 0.05 dB of the plain search; the guarded mode
ends no block worse than the plain search and keeps the better of the two
on every block."""
import numpy as np
import pytest
import torch

import oracle
from crunch2_amd import dxt1, learned, texture
from crunch2_amd.train_learned import DEFAULT_OUT

pytestmark = pytest.mark.gpu


def decode_sse(blocks, ends, sels):
    """Uniform RGB squared error of DXT1 blocks under the standard decode."""
    c = torch.stack([ends & 0xFFFF, (ends >> 16) & 0xFFFF], 1)
    r, g, b = (c >> 11) & 31, (c >> 5) & 63, c & 31
    p = torch.stack([(r << 3) | (r >> 2), (g << 2) | (g >> 4), (b << 3) | (b >> 2)], 2)
    four = (c[:, 0] > c[:, 1])[:, None]
    p2 = torch.where(four, (2 * p[:, 0] + p[:, 1]) // 3, (p[:, 0] + p[:, 1]) // 2)
    p3 = torch.where(four, (p[:, 0] + 2 * p[:, 1]) // 3, torch.zeros_like(p[:, 0]))
    pal = torch.stack([p[:, 0], p[:, 1], p2, p3], 1)
    idx = (sels[:, None] >> (2 * torch.arange(16, device=sels.device))[None, :]) & 3
    dec = torch.gather(pal, 1, idx[:, :, None].expand(-1, -1, 3).long())
    d = dec - blocks[:, :, :3].long()
    return (d * d).sum((1, 2))


def psnr(sse, npix):
    mse = float(sse) / (3.0 * npix)
    return 20.0 * np.log10(255.0 / np.sqrt(mse))


def test_learned_encode_8192(gpu):
    theta = torch.from_numpy(np.load(DEFAULT_OUT)).cuda()
    img = texture.synth(8192, 8192, 4242, texture.RGB)
    blocks_np = dxt1.image_blocks(img)
    blocks = torch.from_numpy(blocks_np).cuda()
    n = blocks.shape[0]
    assert n == 4194304
    raw = dxt1.encode_blocks_device(blocks, dxt1.UBER, learned.LABEL_FLAGS)
    torch.cuda.synchronize()
    # the plain search at this size against the oracle, on a sample
    pick = np.sort(np.random.default_rng(9).choice(n, 3000, replace=False))
    got = dxt1.unpack_results(raw[torch.from_numpy(pick).cuda()].cpu().numpy())
    ref = oracle.dxt1_blocks(blocks_np[pick], 4, oracle.DXT1_PERCEPTUAL)
    for k in range(7):
        assert np.array_equal(np.asarray(got[k]), np.asarray(ref[k])), k
    pend = raw.view(torch.int32).reshape(-1, 6)[:, 2].contiguous()
    perr, psel = learned.block_errors(blocks, pend)
    # default: the network, refinement, the plain search on the 8 % least confident blocks
    ends, sels, errs, nfb = learned.encode(theta, blocks, fallback_frac=0.08)
    e2, _ = learned.block_errors(blocks, ends)
    assert torch.equal(e2, errs)
    sse_p = int(decode_sse(blocks, pend, psel).sum())
    sse_l = int(decode_sse(blocks, ends, sels).sum())
    d = psnr(sse_l, n * 16) - psnr(sse_p, n * 16)
    worse = int((errs > perr).sum())
    print("8192^2: fallback %d blocks, %d blocks worse than plain, delta %.4f dB" % (nfb, worse, d), flush=True)
    assert nfb == round(0.08 * n)
    assert d >= -0.05
    # guarded: every block also runs the plain search and keeps the better
    gends, gsels, gerrs, gnfb = learned.encode(theta, blocks, guarded=True)
    assert gnfb == n
    assert int((gerrs > perr).sum()) == 0
    lo = torch.minimum(perr, learned.encode(theta, blocks, fallback_frac=0.0)[2])
    assert torch.equal(gerrs, lo)
