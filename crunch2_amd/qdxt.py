"""Clustered DXT1 on the GPU (crnlib's qdxt1, crnlib/crn_qdxt1.cpp) and its
vector quantizers (crnlib/crn_clusterizer.h, crn_threaded_clusterizer.h) --
Python face of the crn2_qdxt1_* and crn2_vq_* entry points."""
import ctypes as C

import numpy as np
import torch

from ._lib import Qdxt1Params, Qdxt5Params, check, lib

PERCEPTUAL, ALPHA_BLOCKS, HIERARCHICAL = 1, 2, 4
UBER = 4


def level_blocks(levels):
    """The pixel blocks of a texture's levels as mipmapped_texture::qdxt_pack_init
    lays them out (crnlib/crn_mipmapped_texture.cpp:2410-2469): every level's
    4x4 blocks in raster order, texels past the edge clamped, levels one after
    another; returns (uint8 [n, 64] blocks, uint32 level descriptors: first
    block, width and height in blocks per level)."""
    blocks, desc, first = [], [], 0
    for img in levels:
        h, w = img.shape[:2]
        bw, bh = (w + 3) // 4, (h + 3) // 4
        ys = np.minimum(np.arange(bh * 4), h - 1)
        xs = np.minimum(np.arange(bw * 4), w - 1)
        pad = img[ys][:, xs]
        blocks.append(pad.reshape(bh, 4, bw, 4, 4).transpose(0, 2, 1, 3, 4).reshape(-1, 64))
        desc += [first, bw, bh]
        first += bw * bh
    return np.ascontiguousarray(np.concatenate(blocks), np.uint8), np.array(desc, np.uint32)


def profile(on):
    """Split timing (crn2_qdxt_profile): start with on=True; on=False returns
    (ms [3], launches [3], vector passes [3]) for the 2-, 6- and 16-D trees."""
    ms = (C.c_double * 3)()
    la = (C.c_uint64 * 3)()
    pa = (C.c_uint64 * 3)()
    check(lib().crn2_qdxt_profile(int(bool(on)), ms, la, pa), "crn2_qdxt_profile")
    return list(ms), list(la), list(pa)


def params(quality_level=128, dxt_quality=UBER, flags=PERCEPTUAL | HIERARCHICAL, alpha_threshold=128):
    return Qdxt1Params(quality_level, dxt_quality, flags, alpha_threshold)


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


class Qdxt1:
    """One qdxt1 object: init over the pixel blocks (a uint8 array or device
    tensor of [n, 64]) and their level layout, then any number of packs."""

    def __init__(self, blocks, level_desc, hierarchical=True):
        if isinstance(blocks, torch.Tensor):
            self.blocks = blocks.contiguous()
        else:
            self.blocks = torch.from_numpy(np.ascontiguousarray(blocks, np.uint8).reshape(-1, 64)).cuda()
        self.n = self.blocks.shape[0]
        self.desc = np.ascontiguousarray(level_desc, np.uint32).reshape(-1)
        h = C.c_void_p()
        check(lib().crn2_qdxt1_init(C.c_void_p(self.blocks.data_ptr()), self.n, len(self.desc) // 3,
                                    self.desc.ctypes.data_as(C.c_void_p), int(hierarchical), C.byref(h), _stream()),
              "crn2_qdxt1_init")
        self.h = h

    def pack(self, p=None, pow_mul=1.0, out=None):
        """The dxt1_blocks (uint64 device tensor [n]: low | high << 16 | selectors << 32)."""
        p = p if p is not None else params()
        if out is None:
            out = torch.zeros(self.n, dtype=torch.int64, device=self.blocks.device)
        check(lib().crn2_qdxt1_pack(self.h, C.byref(p), pow_mul, C.c_void_p(out.data_ptr()), _stream()),
              "crn2_qdxt1_pack")
        return out

    def info(self):
        v = [C.c_uint32() for _ in range(4)]
        check(lib().crn2_qdxt1_info(self.h, *[C.byref(x) for x in v]), "crn2_qdxt1_info")
        return dict(zip(("codebook_size", "max_selector_clusters", "endpoint_clusters", "selector_clusters"),
                        [x.value for x in v]))

    def training(self):
        vecs = np.zeros((self.n, 6), np.float32)
        w = np.zeros(self.n, np.uint32)
        enc = np.zeros(self.n, np.uint8)
        check(lib().crn2_qdxt1_training(self.h, vecs.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p),
                                        enc.ctypes.data_as(C.c_void_p)), "crn2_qdxt1_training")
        return vecs, w, enc

    def close(self):
        if self.h:
            lib().crn2_qdxt1_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


BOTH_BLOCK_TYPES = 1


def params5(quality_level=128, dxt_quality=UBER, flags=BOTH_BLOCK_TYPES):
    return Qdxt5Params(quality_level, dxt_quality, flags)


class Qdxt5:
    """One qdxt5 object over one channel (comp 3: alpha; 0 / 1: DXN X / Y)."""

    def __init__(self, blocks, level_desc, comp=3, hierarchical=True):
        if isinstance(blocks, torch.Tensor):
            self.blocks = blocks.contiguous()
        else:
            self.blocks = torch.from_numpy(np.ascontiguousarray(blocks, np.uint8).reshape(-1, 64)).cuda()
        self.n = self.blocks.shape[0]
        self.desc = np.ascontiguousarray(level_desc, np.uint32).reshape(-1)
        h = C.c_void_p()
        check(lib().crn2_qdxt5_init(C.c_void_p(self.blocks.data_ptr()), self.n, len(self.desc) // 3,
                                    self.desc.ctypes.data_as(C.c_void_p), int(hierarchical), comp, C.byref(h),
                                    _stream()), "crn2_qdxt5_init")
        self.h = h

    def pack(self, p=None, out=None):
        """The DXT5 alpha blocks (uint64 device tensor [n]: first | second << 8 | selectors << 16)."""
        p = p if p is not None else params5()
        if out is None:
            out = torch.zeros(self.n, dtype=torch.int64, device=self.blocks.device)
        check(lib().crn2_qdxt5_pack(self.h, C.byref(p), C.c_void_p(out.data_ptr()), _stream()), "crn2_qdxt5_pack")
        return out

    def info(self):
        v = [C.c_uint32() for _ in range(4)]
        check(lib().crn2_qdxt5_info(self.h, *[C.byref(x) for x in v]), "crn2_qdxt5_info")
        return dict(zip(("codebook_size", "max_selector_clusters", "endpoint_clusters", "selector_clusters"),
                        [x.value for x in v]))

    def training(self):
        vecs = np.zeros((self.n, 2), np.float32)
        w = np.zeros(self.n, np.uint32)
        enc = np.zeros(self.n, np.uint8)
        check(lib().crn2_qdxt5_training(self.h, vecs.ctypes.data_as(C.c_void_p), w.ctypes.data_as(C.c_void_p),
                                        enc.ctypes.data_as(C.c_void_p)), "crn2_qdxt5_training")
        return vecs, w, enc

    def close(self):
        if self.h:
            lib().crn2_qdxt5_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _clusters(idx, ofs, nc):
    return [idx[ofs[k]:ofs[k + 1]].copy() for k in range(nc)]


def vq_clusters(vecs, weights, max_size, max_clusters=0):
    """clusterizer<vec<d,float>>: (clusters as index arrays, codebook size)."""
    v = torch.as_tensor(np.ascontiguousarray(vecs, np.float32)).cuda()
    w = torch.as_tensor(np.ascontiguousarray(weights, np.uint32).view(np.int32)).cuda()
    n, d = v.shape
    idx = np.zeros(n, np.uint32)
    ofs = np.zeros(n + 1, np.uint32)
    nc, cbs = C.c_uint32(), C.c_uint32()
    check(lib().crn2_vq_clusters_device(C.c_void_p(v.data_ptr()), C.c_void_p(w.data_ptr()), n, d, max_size,
                                        max_clusters, idx.ctypes.data_as(C.c_void_p),
                                        ofs.ctypes.data_as(C.c_void_p), C.byref(nc), C.byref(cbs), _stream()),
          "crn2_vq_clusters_device")
    return _clusters(idx, ofs, nc.value), cbs.value


def vq_threaded_clusters(vecs, weights, max_clusters):
    v = torch.as_tensor(np.ascontiguousarray(vecs, np.float32)).cuda()
    w = torch.as_tensor(np.ascontiguousarray(weights, np.uint32).view(np.int32)).cuda()
    n, d = v.shape
    idx = np.zeros(n, np.uint32)
    ofs = np.zeros(n + 1, np.uint32)
    nc = C.c_uint32()
    check(lib().crn2_vq_threaded_clusters_device(C.c_void_p(v.data_ptr()), C.c_void_p(w.data_ptr()), n, d,
                                                 max_clusters, idx.ctypes.data_as(C.c_void_p),
                                                 ofs.ctypes.data_as(C.c_void_p), C.byref(nc), _stream()),
          "crn2_vq_threaded_clusters_device")
    return _clusters(idx, ofs, nc.value)
