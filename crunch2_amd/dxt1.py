"""DXT1 endpoint search on the GPU (crnlib's dxt1_endpoint_optimizer,
crnlib/crn_dxt1.h:163-352) -- Python face of crn2_dxt1_blocks_device."""
import ctypes as C

import numpy as np
import torch

from ._lib import lib, check

PERCEPTUAL, ALPHA_BLOCKS, GRAYSCALE, FORCE_ALPHA = 1, 2, 4, 8
# DXT1A: texels with alpha below the threshold (threshold(t), bits 8..15) are
# transparent; BLACK_AS_TRANSPARENT runs the transparent-black trial;
# DXT1A_PER_BLOCK sets HAVE_ALPHA for the blocks that have such a texel
HAVE_ALPHA, BLACK_AS_TRANSPARENT, DXT1A_PER_BLOCK = 16, 32, 64
UBER, BETTER, NORMAL, FAST, SUPERFAST = 4, 3, 2, 1, 0


def threshold(t=128):
    return (int(t) & 255) << 8


OUT_DTYPE = np.dtype([("error", "<u8"), ("lo_hi", "<u4"), ("selectors", "<u4"), ("flags", "<u4"), ("pad", "<u4")])


def encode_blocks_device(rgba_blocks, quality=UBER, flags=PERCEPTUAL | ALPHA_BLOCKS, out=None, stream=None):
    """rgba_blocks: uint8 CUDA tensor [n, 16, 4] (contiguous). Returns a uint8
    CUDA tensor [n, 24] of packed results (see include/crn2.h)."""
    n = rgba_blocks.shape[0]
    if out is None:
        out = torch.empty((n, 24), dtype=torch.uint8, device=rgba_blocks.device)
    s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
    check(lib().crn2_dxt1_blocks_device(C.c_void_p(rgba_blocks.data_ptr()), n, quality, flags,
                                        C.c_void_p(out.data_ptr()), C.c_void_p(s)), "crn2_dxt1_blocks_device")
    return out


def unpack_results(raw):
    """uint8 [n, 24] -> (selectors [n,16], error, low, high, alpha_block,
    reordered, alternate_rounding) numpy arrays, the layout of the oracle."""
    r = np.ascontiguousarray(raw).view(OUT_DTYPE).reshape(-1)
    sel = ((r["selectors"][:, None] >> (2 * np.arange(16, dtype=np.uint32))[None, :]) & 3).astype(np.uint8)
    f = r["flags"]
    return (sel, r["error"], (r["lo_hi"] & 0xFFFF).astype(np.uint16), (r["lo_hi"] >> 16).astype(np.uint16),
            (f & 1).astype(np.uint8), ((f >> 1) & 1).astype(np.uint8), ((f >> 2) & 1).astype(np.uint8))


def encode_blocks(rgba_blocks, quality=UBER, flags=PERCEPTUAL | ALPHA_BLOCKS):
    t = torch.from_numpy(np.ascontiguousarray(rgba_blocks, np.uint8).reshape(-1, 16, 4)).cuda()
    raw = encode_blocks_device(t, quality, flags)
    torch.cuda.synchronize()
    return unpack_results(raw.cpu().numpy())


def image_blocks(img):
    """RGBA8 image [h, w, 4] -> [nblocks, 16, 4] with edge clamp (crn_dxt_image.cpp:330-344)."""
    h, w = img.shape[:2]
    bx, by = (w + 3) // 4, (h + 3) // 4
    ys = np.minimum(np.arange(by * 4), h - 1)
    xs = np.minimum(np.arange(bx * 4), w - 1)
    p = img[ys][:, xs]
    return np.ascontiguousarray(p.reshape(by, 4, bx, 4, 4).transpose(0, 2, 1, 3, 4).reshape(-1, 16, 4))


CLUSTER_DTYPE = np.dtype([("error", "<u8"), ("lo_hi", "<u4"), ("nunique", "<u4"), ("flags", "<u4"), ("pad", "<u4")])


def encode_clusters_device(rgba_px, offsets, perceptual=True, selectors=True, alpha_blocks=False, alpha_threshold=0):
    """DXT1 endpoint search per pixel cluster at uber quality: the hc
    evaluators (dxt_hc endpoint clusters), or with alpha_blocks the uber
    evaluator with 3-colour blocks (qdxt1's clusters), where texels with alpha
    below alpha_threshold are transparent (DXT1A, selector 3).
    rgba_px: uint8 CUDA tensor [n, 4]; offsets: int32 CUDA tensor [k + 1].
    Returns (uint8 [k, 24] packed results, uint8 [n] selectors or None)."""
    k = offsets.shape[0] - 1
    out = torch.empty((k, 24), dtype=torch.uint8, device=rgba_px.device)
    sel = torch.empty(rgba_px.shape[0], dtype=torch.uint8, device=rgba_px.device) if selectors else None
    s = torch.cuda.current_stream().cuda_stream
    flags = int(bool(perceptual)) | (2 | threshold(alpha_threshold) if alpha_blocks else 0)
    check(lib().crn2_dxt1_clusters_device(C.c_void_p(rgba_px.data_ptr()), C.c_void_p(offsets.data_ptr()), k, flags,
                                          C.c_void_p(out.data_ptr()),
                                          C.c_void_p(sel.data_ptr() if sel is not None else 0), C.c_void_p(s)),
          "crn2_dxt1_clusters_device")
    return out, sel
