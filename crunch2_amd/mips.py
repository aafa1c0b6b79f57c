"""Mip generation on the GPU (mipmapped_texture::generate_mipmaps,
crnlib/crn_mipmapped_texture.cpp:2140-2220) -- Python face of the
crn2_mip_plan_* entry points."""
import ctypes as C

import numpy as np
import torch

from ._lib import MipParams, check, lib

BOX, TENT, LANCZOS4, MITCHELL, KAISER = 0, 1, 2, 3, 4
SRGB, WRAP, ALPHA, RENORMALIZE, THREADED = 1, 2, 4, 8, 16


def params(filter=KAISER, flags=SRGB | ALPHA, filter_scale=0.9, source_gamma=2.2, max_levels=0, min_mip_size=1):
    """crn_mipmap_params defaults (inc/crnlib.h:478-505): Kaiser, gamma
    filtering, blurriness 0.9, clamp; generate_mipmaps keeps gamma 2.2."""
    return MipParams(filter, flags, filter_scale, source_gamma, max_levels, min_mip_size)


class Plan:
    """A crn2_mip_plan for one level-0 size."""

    def __init__(self, width, height, p=None):
        self.width, self.height = width, height
        self.p = p if p is not None else params()
        h = C.c_void_p()
        check(lib().crn2_mip_plan_create(width, height, C.byref(self.p), C.byref(h)), "crn2_mip_plan_create")
        self.h = h
        n, size = C.c_uint32(), C.c_uint64()
        check(lib().crn2_mip_plan_info(h, C.byref(n), C.byref(size), None, None), "crn2_mip_plan_info")
        self.levels, self.chain_bytes = n.value, size.value
        self.offsets = np.zeros(self.levels, np.uint64)
        self.y_first = np.zeros(self.levels, np.uint8)
        check(lib().crn2_mip_plan_info(h, None, None, self.offsets.ctypes.data, self.y_first.ctypes.data),
              "crn2_mip_plan_info")

    def sizes(self):
        return [(max(1, self.width >> l), max(1, self.height >> l)) for l in range(self.levels)]

    def run(self, rgba, out=None, stream=None):
        """rgba: uint8 CUDA tensor [h, w, 4]; returns uint8 [chain_bytes]."""
        if out is None:
            out = torch.empty(self.chain_bytes, dtype=torch.uint8, device=rgba.device)
        s = stream if stream is not None else torch.cuda.current_stream().cuda_stream
        check(lib().crn2_mip_plan_run(self.h, C.c_void_p(rgba.data_ptr()), C.c_void_p(out.data_ptr()),
                                      out.numel(), C.c_void_p(s)), "crn2_mip_plan_run")
        return out

    def profile(self, on):
        ms = (C.c_double * 3)()
        runs = C.c_uint64()
        check(lib().crn2_mip_plan_profile(self.h, int(on), ms, C.byref(runs)), "crn2_mip_plan_profile")
        return list(ms), runs.value

    def split(self, chain):
        """Chain bytes (numpy) -> list of [h, w, 4] levels."""
        out = []
        for (w, h), o in zip(self.sizes(), self.offsets):
            out.append(chain[int(o):int(o) + w * h * 4].reshape(h, w, 4))
        return out

    def close(self):
        if self.h:
            lib().crn2_mip_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def generate(img, p=None):
    """RGBA8 numpy [h, w, 4] -> list of levels (numpy), computed on the GPU."""
    h, w = img.shape[:2]
    plan = Plan(w, h, p)
    t = torch.from_numpy(np.ascontiguousarray(img, np.uint8)).cuda()
    chain = plan.run(t)
    torch.cuda.synchronize()
    levels = plan.split(chain.cpu().numpy())
    plan.close()
    return levels


def contributors(src, dst, filter=KAISER, filter_scale=0.9, wrap=False):
    """Host contributor list of one axis: (offs [dst+1], pix, weight)."""
    offs = np.zeros(dst + 1, np.uint32)
    n = lib().crn2_mip_contributors(src, dst, int(wrap), filter, filter_scale, offs.ctypes.data, None, None, 0)
    if n < 0:
        raise RuntimeError("crn2_mip_contributors failed with status %d" % n)
    pix = np.zeros(max(n, 1), np.uint32)
    wt = np.zeros(max(n, 1), np.float32)
    check(0 if lib().crn2_mip_contributors(src, dst, int(wrap), filter, filter_scale, offs.ctypes.data,
                                           pix.ctypes.data, wt.ctypes.data, n) == n else -1, "crn2_mip_contributors")
    return offs, pix[:n], wt[:n]


def gamma_tables(gamma=2.2):
    s2l = np.zeros(256, np.float32)
    l2s = np.zeros(8192, np.uint8)
    check(lib().crn2_mip_gamma_tables(gamma, s2l.ctypes.data, l2s.ctypes.data), "crn2_mip_gamma_tables")
    return s2l, l2s
