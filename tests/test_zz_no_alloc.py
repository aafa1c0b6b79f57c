"""crn2_dxt1_clusters_noalloc_device (the variant that allocates nothing)
against crn2_dxt1_clusters_device on the same clusters, through the C ABI:
every result record and every selector byte equal, on two calls in a row, the
second with fewer clusters and texels than the first (the fixed scratch keeps
the first call's contents)."""
import ctypes as C

import numpy as np
import pytest
import torch

from crunch2_amd import texture
from crunch2_amd._lib import lib

pytestmark = pytest.mark.gpu


def _clusters(seed, nclusters, max_size):
    rng = np.random.default_rng(seed)
    img = texture.synth(512, 512, seed, texture.RGB).reshape(-1, 4)
    sizes = rng.integers(1, max_size, nclusters)
    sizes[:3] = [1, 600, 2500]                       # touch three of the four size classes
    parts = []
    for n in sizes:
        if rng.random() < 0.5:
            s = int(rng.integers(0, img.shape[0] - n))
            parts.append(img[s:s + n])
        else:
            parts.append(img[rng.integers(0, img.shape[0], n)])
    px = np.ascontiguousarray(np.concatenate(parts))
    ofs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint32)
    return px, ofs


def _run(fn, px_d, ofs_d, k, perceptual):
    out = torch.full((k, 24), 0xAB, dtype=torch.uint8, device="cuda")
    sel = torch.full((px_d.shape[0],), 0xCD, dtype=torch.uint8, device="cuda")
    s = torch.cuda.current_stream().cuda_stream
    rc = fn(C.c_void_p(px_d.data_ptr()), C.c_void_p(ofs_d.data_ptr()), k, int(perceptual),
            C.c_void_p(out.data_ptr()), C.c_void_p(sel.data_ptr()), C.c_void_p(s))
    assert rc == 0, rc
    return out, sel


@pytest.mark.parametrize("perceptual", [True, False])
def test_noalloc_matches_allocating_entry(gpu, perceptual):
    L = lib()
    for call, (k, max_size) in enumerate([(700, 1200), (90, 300)]):
        px, ofs = _clusters(40 + call + 2 * perceptual, k, max_size)
        print("call", call, "clusters", k, "texels", len(px), "perceptual", perceptual, flush=True)
        px_d = torch.from_numpy(px).cuda()
        ofs_d = torch.from_numpy(ofs.view(np.int32)).cuda()
        ref_out, ref_sel = _run(L.crn2_dxt1_clusters_device, px_d, ofs_d, k, perceptual)
        got_out, got_sel = _run(L.crn2_dxt1_clusters_noalloc_device, px_d, ofs_d, k, perceptual)
        torch.cuda.synchronize()
        assert torch.equal(ref_out, got_out), call
        assert torch.equal(ref_sel, got_sel), call


def test_noalloc_rejects_too_many_clusters(gpu):
    px = torch.zeros((65537, 4), dtype=torch.uint8, device="cuda")
    ofs = torch.arange(65538, dtype=torch.int32, device="cuda")
    out = torch.zeros((65537, 24), dtype=torch.uint8, device="cuda")
    rc = lib().crn2_dxt1_clusters_noalloc_device(C.c_void_p(px.data_ptr()), C.c_void_p(ofs.data_ptr()), 65537, 1,
                                                 C.c_void_p(out.data_ptr()), None,
                                                 C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == -1
