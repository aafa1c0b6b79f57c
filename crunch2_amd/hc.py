"""dxt_hc colour selector assignment on the GPU (crnlib/crn_dxt_hc.cpp:1306-1360)
-- Python face of crn2_hc_color_selectors_device."""
import ctypes as C

import torch

from ._lib import lib, check


def _ptr(t):
    return C.c_void_p(t.data_ptr() if t is not None else 0)


def color_selectors(blocks, endpoint_index, cluster_colors, selectors, subblocks=False, perceptual=True,
                    details=None, used=None):
    """blocks: uint8 CUDA tensor [n, 16, 4]; endpoint_index: int32/uint32
    [n] or [n, 2]; cluster_colors: uint8 [k, 4, 4]; selectors: int32 [m]
    (bit patterns).  Returns (best_index int32 [n], best_error int32 [n]);
    details (int32 [m, 64]) and used (uint8 [m]) are accumulated in place
    when given."""
    n = blocks.shape[0]
    m = selectors.shape[0]
    dev = blocks.device
    best = torch.empty(n, dtype=torch.int32, device=dev)
    err = torch.empty(n, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    check(lib().crn2_hc_color_selectors_device(_ptr(blocks.contiguous()), _ptr(endpoint_index.contiguous()),
                                               _ptr(cluster_colors.contiguous()), n, int(bool(subblocks)),
                                               int(bool(perceptual)), _ptr(selectors.contiguous()), m, _ptr(best),
                                               _ptr(err), _ptr(details), _ptr(used), C.c_void_p(stream)),
          "crn2_hc_color_selectors_device")
    return best, err
