"""Learned DXT1 endpoint proposal on the GPU -- training, inference, refinement
and the fallback to the plain endpoint search.

There is no reference counterpart: crnlib's dxt1_endpoint_optimizer
(crnlib/crn_dxt1.cpp:2234-2246) is the plain search the network learns from,
the fallback for blocks the network is unsure of, and (guarded mode) the bar
every block must meet.  Every kernel is the project's own (include/crn2.h,
"learned DXT1 endpoint proposal"); torch only allocates device memory.
"""
import ctypes as C

import numpy as np
import torch

from . import dxt1
from ._lib import lib, check

N_IN, N_H, N_OUT = 57, 128, 7
SHAPES = [("W1", (N_IN, N_H)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)), ("W3", (N_H, N_OUT)),
          ("b3", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)
OFFSETS = {}
_o = 0
for _n, _s in SHAPES:
    OFFSETS[_n] = _o
    _o += int(np.prod(_s))

LABEL_FLAGS = dxt1.PERCEPTUAL     # four-colour blocks, perceptual metric


def _p(t):
    return C.c_void_p(t.data_ptr() if t is not None else 0)


def _s():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def init_params(seed=0, device="cuda"):
    rng = np.random.default_rng(seed)
    parts = []
    for name, shp in SHAPES:
        if name.startswith("W"):
            parts.append((rng.standard_normal(shp) * np.sqrt(2.0 / shp[0])).astype(np.float32).ravel())
        else:
            parts.append(np.zeros(int(np.prod(shp)), np.float32))
    return torch.from_numpy(np.concatenate(parts)).to(device)


def gemm(M, N, K, A, sam, sak, B, sbk, sbn, Cm, ldc, accumulate=False, round_a=False, round_b=False, ksplit=1,
         a_ofs=0, b_ofs=0, c_ofs=0):
    """Strided float32 GEMM on the device (offsets in elements)."""
    pa = C.c_void_p(A.data_ptr() + 4 * a_ofs)
    pb = C.c_void_p(B.data_ptr() + 4 * b_ofs)
    pc = C.c_void_p(Cm.data_ptr() + 4 * c_ofs)
    check(lib().crn2_gemm_f32_device(M, N, K, pa, sam, sak, pb, sbk, sbn, pc, ldc, int(accumulate), int(round_a),
                                     int(round_b), ksplit, _s()), "crn2_gemm_f32_device")


def features(rgba_blocks):
    n = rgba_blocks.shape[0]
    f = torch.empty((n, N_IN), dtype=torch.float32, device=rgba_blocks.device)
    check(lib().crn2_mlp_features_device(_p(rgba_blocks), n, _p(f), _s()), "crn2_mlp_features_device")
    return f


class Trainer:
    """Adam on float32 master weights; every forward rounds weights and layer
    inputs to bfloat16 (the network the inference kernel runs)."""

    def __init__(self, theta=None, lr=2e-3, seed=0, device="cuda"):
        self.dev = torch.device(device)
        self.theta = theta.clone() if theta is not None else init_params(seed, device)
        self.m = torch.zeros_like(self.theta)
        self.v = torch.zeros_like(self.theta)
        self.g = torch.zeros_like(self.theta)
        self.loss = torch.zeros(1, dtype=torch.float32, device=self.dev)
        self.ones = torch.ones(1, dtype=torch.float32, device=self.dev)
        self.lr = lr
        self.step_no = 0
        self._bufs = {}

    def _buf(self, name, shape):
        b = self._bufs.get(name)
        if b is None or b.shape != shape:
            b = torch.empty(shape, dtype=torch.float32, device=self.dev)
            self._bufs[name] = b
        return b

    def gradients(self, X, T, labels=None):
        """Loss and flat gradient of the batch (X: n x 57, T: n x 6 float32;
        labels: optional n float32 confidence labels)."""
        n = X.shape[0]
        th, o = self.theta, OFFSETS
        Z1, A1 = self._buf("Z1", (n, N_H)), self._buf("A1", (n, N_H))
        Z2, A2 = self._buf("Z2", (n, N_H)), self._buf("A2", (n, N_H))
        Y, dY = self._buf("Y", (n, N_OUT)), self._buf("dY", (n, N_OUT))
        dA, dZ = self._buf("dA", (n, N_H)), self._buf("dZ", (n, N_H))
        s = _s()
        L = lib()
        # forward
        gemm(n, N_H, N_IN, X, N_IN, 1, th, N_H, 1, Z1, N_H, round_a=True, round_b=True, b_ofs=o["W1"])
        check(L.crn2_mlp_bias_act_device(_p(Z1), _p(A1), C.c_void_p(th.data_ptr() + 4 * o["b1"]), n, N_H, 1, s),
              "bias_act")
        gemm(n, N_H, N_H, A1, N_H, 1, th, N_H, 1, Z2, N_H, round_b=True, b_ofs=o["W2"])
        check(L.crn2_mlp_bias_act_device(_p(Z2), _p(A2), C.c_void_p(th.data_ptr() + 4 * o["b2"]), n, N_H, 1, s),
              "bias_act")
        gemm(n, N_OUT, N_H, A2, N_H, 1, th, N_OUT, 1, Y, N_OUT, round_b=True, b_ofs=o["W3"])
        check(L.crn2_mlp_bias_act_device(_p(Y), _p(Y), C.c_void_p(th.data_ptr() + 4 * o["b3"]), n, N_OUT, 0, s),
              "bias_act")
        self.loss.zero_()
        lp = C.c_void_p(labels.data_ptr()) if labels is not None else C.c_void_p(0)
        check(L.crn2_mlp_loss_device(_p(Y), _p(T), lp, n, _p(dY), _p(self.loss), s), "loss")
        # backward
        g = self.g
        g.zero_()
        ks = max(1, min(64, n // 1024))
        gemm(N_H, N_OUT, n, A2, 1, N_H, dY, N_OUT, 1, g, N_OUT, ksplit=ks, c_ofs=o["W3"])
        gemm(1, N_OUT, n, self.ones, 0, 0, dY, N_OUT, 1, g, N_OUT, ksplit=ks, c_ofs=o["b3"])
        gemm(n, N_H, N_OUT, dY, N_OUT, 1, th, 1, N_OUT, dA, N_H, round_b=True, b_ofs=o["W3"])
        check(L.crn2_mlp_gelu_back_device(_p(dA), _p(Z2), _p(dZ), n * N_H, s), "gelu_back")
        gemm(N_H, N_H, n, A1, 1, N_H, dZ, N_H, 1, g, N_H, ksplit=ks, c_ofs=o["W2"])
        gemm(1, N_H, n, self.ones, 0, 0, dZ, N_H, 1, g, N_H, ksplit=ks, c_ofs=o["b2"])
        gemm(n, N_H, N_H, dZ, N_H, 1, th, 1, N_H, dA, N_H, round_b=True, b_ofs=o["W2"])
        check(L.crn2_mlp_gelu_back_device(_p(dA), _p(Z1), _p(dZ), n * N_H, s), "gelu_back")
        gemm(N_IN, N_H, n, X, 1, N_IN, dZ, N_H, 1, g, N_H, round_a=True, ksplit=ks, c_ofs=o["W1"])
        gemm(1, N_H, n, self.ones, 0, 0, dZ, N_H, 1, g, N_H, ksplit=ks, c_ofs=o["b1"])
        return self.loss, g

    def step(self, X, T, labels=None):
        loss, g = self.gradients(X, T, labels)
        self.step_no += 1
        check(lib().crn2_adam_device(_p(self.theta), _p(g), _p(self.m), _p(self.v), N_PARAMS, self.lr, 0.9, 0.999,
                                     1e-8, self.step_no, _s()), "adam")
        return loss


def infer(theta, rgba_blocks):
    """n x 8 network outputs (bfloat16 matrix-core forward)."""
    n = rgba_blocks.shape[0]
    y = torch.empty((n, 8), dtype=torch.float32, device=rgba_blocks.device)
    check(lib().crn2_mlp_infer_device(_p(rgba_blocks), n, _p(theta), _p(y), _s()), "crn2_mlp_infer_device")
    return y


START_NET, START_BOX, START_PCA, REFIT, POLISH = 1, 2, 4, 8, 16
# refinement of the learned encode: every start, refit and lattice polish
ENCODE_STARTS = START_NET | START_BOX | START_PCA | REFIT | POLISH
ENCODE_ITERS = 16


def refine(rgba_blocks, y, max_iters=8, starts=START_NET):
    """Refinement from the network proposal y and/or the bright/dark texel start."""
    n = rgba_blocks.shape[0]
    out = torch.empty((n, 4), dtype=torch.int32, device=rgba_blocks.device)
    check(lib().crn2_dxt1_refine_device(_p(rgba_blocks), n, _p(y), max_iters, starts, _p(out), _s()),
          "crn2_dxt1_refine_device")
    return out


def block_errors(rgba_blocks, endpoints, stride_u32=1):
    """Exact perceptual error and best selectors of given endpoint words."""
    n = rgba_blocks.shape[0]
    err = torch.empty(n, dtype=torch.int32, device=rgba_blocks.device)
    sel = torch.empty(n, dtype=torch.int32, device=rgba_blocks.device)
    check(lib().crn2_dxt1_block_error_device(_p(rgba_blocks), n, _p(endpoints), stride_u32, _p(err), _p(sel), _s()),
          "crn2_dxt1_block_error_device")
    return err, sel


def plain_targets(rgba_blocks, quality=dxt1.UBER):
    """Targets (n x 6, brighter endpoint first, components / (31, 63, 31))
    and the raw plain results for a batch of blocks."""
    raw = dxt1.encode_blocks_device(rgba_blocks, quality, LABEL_FLAGS)
    lohi = raw.view(torch.int32).reshape(-1, 6)[:, 2]
    lo = lohi & 0xFFFF
    hi = (lohi >> 16) & 0xFFFF

    def comps(c):
        return torch.stack([(c >> 11) & 31, (c >> 5) & 63, c & 31], 1)
    a, b = comps(lo), comps(hi)
    la = 2 * a[:, 0] * 33 // 4 + 4 * a[:, 1] * 65 // 16 + a[:, 2] * 33 // 4
    lb = 2 * b[:, 0] * 33 // 4 + 4 * b[:, 1] * 65 // 16 + b[:, 2] * 33 // 4
    swap = (lb > la) | ((lb == la) & (hi > lo))
    first = torch.where(swap[:, None], b, a).float()
    second = torch.where(swap[:, None], a, b).float()
    scale = torch.tensor([31.0, 63.0, 31.0], device=rgba_blocks.device)
    return torch.cat([first / scale, second / scale], 1).contiguous(), raw


def encode(theta, rgba_blocks, conf_threshold=0.0, max_iters=ENCODE_ITERS, guarded=False, quality=dxt1.UBER,
           starts=ENCODE_STARTS, fallback_frac=None):
    """Learned DXT1 encode of n blocks.  Blocks whose confidence logit is below
    `conf_threshold` (or, with `fallback_frac`, that fraction of the blocks with
    the lowest confidence) also run the plain search and keep the better of the
    two; with `guarded`, every block does (no block ends worse than plain).
    Returns (endpoints int32 [n] (c0 | c1 << 16), selectors int32 [n],
    errors int32 [n], number of blocks that ran the plain search)."""
    y = infer(theta, rgba_blocks)
    r = refine(rgba_blocks, y, max_iters, starts)
    ends, sels, errs = r[:, 0].clone(), r[:, 1].clone(), r[:, 2].clone()
    n = rgba_blocks.shape[0]
    if guarded:
        idx = torch.arange(n, device=rgba_blocks.device)
    elif fallback_frac is not None:
        k = int(round(fallback_frac * n))
        idx = torch.topk(y[:, 6], k, largest=False).indices if k else torch.zeros(0, dtype=torch.long,
                                                                                 device=rgba_blocks.device)
    else:
        idx = torch.nonzero(y[:, 6] < conf_threshold).flatten()
    if idx.numel():
        sub = rgba_blocks.index_select(0, idx).contiguous()
        raw = dxt1.encode_blocks_device(sub, quality, LABEL_FLAGS)
        w = raw.view(torch.int32).reshape(-1, 6)
        pend = w[:, 2].contiguous()
        perr, psel = block_errors(sub, pend)
        better = perr < errs[idx]
        bi = idx[better]
        ends[bi] = pend[better]
        sels[bi] = psel[better]
        errs[bi] = perr[better]
    return ends, sels, errs, int(idx.numel())
