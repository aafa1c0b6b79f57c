"""DXT5 alpha endpoint search on the GPU (crnlib's dxt5_endpoint_optimizer,
crnlib/crn_dxt5a.h:31-88) -- Python face of crn2_dxt5a_blocks_device and
crn2_encode_dxt5a_image_device."""
import ctypes as C

import numpy as np
import torch

from ._lib import lib, check

UBER = 4


def compute_blocks(values, quality=UBER, use_both=True):
    """values: uint8 tensor/array [n, 16]. Returns (selectors [n,16] uint8 raw
    DXT5 codes, errors [n] int64, first, second, block_type, reordered) as
    CPU numpy arrays; the search runs on the current CUDA device."""
    v = torch.as_tensor(np.ascontiguousarray(values, np.uint8)).reshape(-1, 16).cuda()
    n = v.shape[0]
    sel = torch.empty(n, dtype=torch.int64, device=v.device)
    err = torch.empty(n, dtype=torch.int64, device=v.device)
    meta = torch.empty(n, dtype=torch.int32, device=v.device)
    stream = torch.cuda.current_stream().cuda_stream
    check(lib().crn2_dxt5a_blocks_device(C.c_void_p(v.data_ptr()), n, quality, int(use_both),
                                         C.c_void_p(sel.data_ptr()), C.c_void_p(err.data_ptr()),
                                         C.c_void_p(meta.data_ptr()), C.c_void_p(stream)), "crn2_dxt5a_blocks_device")
    s = sel.cpu().numpy().view(np.uint64)
    shifts = (3 * np.arange(16)).astype(np.uint64)
    sels = ((s[:, None] >> shifts[None, :]) & np.uint64(7)).astype(np.uint8)
    m = meta.cpu().numpy().view(np.uint32)
    return sels, err.cpu().numpy().view(np.uint64), m & 0xFF, (m >> 8) & 0xFF, (m >> 16) & 0xFF, (m >> 24) & 0xFF


def encode_image(rgba, comp, quality=UBER, use_both=True, stride=1, out=None, errors=None):
    """rgba: uint8 CUDA tensor [h, w, 4]. Returns int64 tensor of DXT5 alpha
    blocks (8 bytes each, row-major blocks)."""
    h, w = rgba.shape[:2]
    bx, by = (w + 3) // 4, (h + 3) // 4
    if out is None:
        out = torch.zeros(bx * by * stride, dtype=torch.int64, device=rgba.device)
    stream = torch.cuda.current_stream().cuda_stream
    check(lib().crn2_encode_dxt5a_image_device(C.c_void_p(rgba.data_ptr()), w, h, comp, quality, int(use_both),
                                               C.c_void_p(out.data_ptr()), stride,
                                               C.c_void_p(errors.data_ptr() if errors is not None else 0),
                                               C.c_void_p(stream)), "crn2_encode_dxt5a_image_device")
    return out
