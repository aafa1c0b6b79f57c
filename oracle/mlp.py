"""CPU oracle for the learned DXT1 endpoint proposal -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use
this module, and only as the checker of the HIP kernels in
crunch2_amd/csrc/mlp_gpu.hip.  The learned proposal has no counterpart in the
reference; the closest reference pieces are the block statistics of
dxt1_endpoint_optimizer::compute_vectors / compute_pca
(crnlib/crn_dxt1.cpp:155-256) that the covariance features restate, and the
exact DXT1 block error (crn_dxt1.cpp:1370-1592, weights crn_color.h:724-726)
that judges every proposal.

Model: 57 features (48 texel channels scaled to [-0.5, 0.5], 9 covariance
entries of the texel colours scaled by 4) -> 128 -> 128 -> 7 with tanh-form
GELU, weights and layer inputs rounded to bfloat16 (round to nearest even),
sums in float32.  Outputs 0-5 are the two endpoints (r, g, b of the brighter,
then the darker, each scaled to [0, 1] over its 5/6-bit range), output 6 a
confidence logit.

Parity pins (tests/test_mlp_oracle.py): the analytic gradients of
`loss_and_grads` against central finite differences of its float64 loss;
the covariance features against numpy.cov; `bf16` against torch's bfloat16
conversion; `adam` against torch.optim.Adam.
"""
import numpy as np

N_IN, N_H, N_OUT = 57, 128, 7
SHAPES = [("W1", (N_IN, N_H)), ("b1", (N_H,)), ("W2", (N_H, N_H)), ("b2", (N_H,)), ("W3", (N_H, N_OUT)),
          ("b3", (N_OUT,))]
N_PARAMS = sum(int(np.prod(s)) for _, s in SHAPES)   # 24,839
CONF_WEIGHT = 0.05
GELU_C = np.float32(0.7978845608028654)   # sqrt(2 / pi)
GELU_A = np.float32(0.044715)


def bf16(x):
    """Round float32 to the nearest bfloat16 (ties to even), returned as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).reshape(x.shape)


def unpack(theta):
    out, o = {}, 0
    for name, shp in SHAPES:
        n = int(np.prod(shp))
        out[name] = theta[o:o + n].reshape(shp)
        o += n
    return out


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    parts = []
    for name, shp in SHAPES:
        if name.startswith("W"):
            parts.append((rng.standard_normal(shp) * np.sqrt(2.0 / shp[0])).astype(np.float32).ravel())
        else:
            parts.append(np.zeros(int(np.prod(shp)), np.float32))
    return np.concatenate(parts)


def features(blocks):
    """blocks: (n, 16, 4) uint8 -> (n, 57) float32."""
    px = blocks[:, :, :3].astype(np.float32) / np.float32(255.0)
    f = np.empty((blocks.shape[0], N_IN), np.float32)
    f[:, :48] = (px - np.float32(0.5)).reshape(-1, 48)
    mean = px.mean(axis=1, keepdims=True, dtype=np.float32)
    d = px - mean
    cov = np.einsum("nki,nkj->nij", d, d).astype(np.float32) / np.float32(16.0)
    f[:, 48:] = cov.reshape(-1, 9) * np.float32(4.0)
    return f


def gelu(z):
    t = np.tanh(GELU_C * (z + GELU_A * z * z * z))
    return np.float32(0.5) * z * (np.float32(1.0) + t)


def gelu_grad(z):
    u = GELU_C * (z + GELU_A * z * z * z)
    t = np.tanh(u)
    du = GELU_C * (np.float32(1.0) + np.float32(3.0) * GELU_A * z * z)
    return np.float32(0.5) * (np.float32(1.0) + t) + np.float32(0.5) * z * (np.float32(1.0) - t * t) * du


def _rounders(bf16_round):
    """bf16_round: bfloat16 weights and layer inputs, float32 biases and sums
    (the GPU's network); otherwise everything in float64 (for finite
    differences).  Returns (operand rounding, bias conversion)."""
    if bf16_round:
        return bf16, (lambda v: v)
    f64 = (lambda a: np.asarray(a, np.float64))
    return f64, f64


def forward(theta, x, bf16_round=True):
    p = unpack(theta)
    r, b = _rounders(bf16_round)
    x = r(x)
    z1 = x @ r(p["W1"]) + b(p["b1"])
    a1 = r(gelu(z1))
    z2 = a1 @ r(p["W2"]) + b(p["b2"])
    a2 = r(gelu(z2))
    y = a2 @ r(p["W3"]) + b(p["b3"])
    return y, (x, z1, a1, z2, a2)


def conf_labels(y, t):
    """1 where every predicted endpoint component, rounded to its 5/6-bit
    grid, is within one step of the target's."""
    scale = np.array([31, 63, 31, 31, 63, 31], np.float32)
    q = np.clip(np.rint(y[:, :6] * scale), 0, scale)
    tq = np.rint(t[:, :6] * scale)
    return (np.abs(q - tq) <= 1).all(axis=1).astype(np.float32)


def loss_and_grads(theta, x, t, bf16_round=True, labels=None):
    """Mean over the batch of: squared error of outputs 0-5 against t[:, :6]
    plus CONF_WEIGHT * logistic loss of output 6 against conf_labels (or the
    given labels).  Returns (loss, flat gradient).  The bf16 rounding of the
    forward is passed straight through in the backward pass."""
    p = unpack(theta)
    r, _ = _rounders(bf16_round)
    y, (xr, z1, a1, z2, a2) = forward(theta, x, bf16_round)
    n = x.shape[0]
    lab = conf_labels(y, t) if labels is None else labels
    d = y[:, :6] - t[:, :6]
    s = 1.0 / (1.0 + np.exp(-y[:, 6].astype(np.float64)))
    bce = np.logaddexp(0.0, y[:, 6].astype(np.float64)) - lab * y[:, 6]
    loss = float((d.astype(np.float64) ** 2).sum() / n + CONF_WEIGHT * bce.sum() / n)
    dy = np.zeros_like(y)
    dy[:, :6] = (2.0 / n) * d
    dy[:, 6] = (CONF_WEIGHT / n) * (s - lab)
    g = {}
    g["W3"] = a2.T @ dy
    g["b3"] = dy.sum(0)
    dz2 = (dy @ r(p["W3"]).T) * gelu_grad(z2)
    g["W2"] = a1.T @ dz2
    g["b2"] = dz2.sum(0)
    dz1 = (dz2 @ r(p["W2"]).T) * gelu_grad(z1)
    g["W1"] = xr.T @ dz1
    g["b1"] = dz1.sum(0)
    flat = np.concatenate([g[name].ravel() for name, _ in SHAPES])
    return loss, flat.astype(np.float32 if bf16_round else np.float64)


def adam(theta, grad, m, v, step, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
    m[:] = b1 * m + (1 - b1) * grad
    v[:] = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    theta[:] = theta - lr * mh / (np.sqrt(vh) + eps)


def targets_from_endpoints(low, high):
    """Plain-optimizer endpoints (RGB565) -> (n, 6) targets: the brighter
    endpoint first, components scaled to [0, 1]."""
    def comps(c):
        c = c.astype(np.int64)
        return np.stack([(c >> 11) & 31, (c >> 5) & 63, c & 31], 1)
    a, b = comps(low), comps(high)
    la = 2 * a[:, 0] * 33 // 4 + 4 * a[:, 1] * 65 // 16 + a[:, 2] * 33 // 4
    lb = 2 * b[:, 0] * 33 // 4 + 4 * b[:, 1] * 65 // 16 + b[:, 2] * 33 // 4
    swap = (lb > la) | ((lb == la) & (high.astype(np.int64) > low.astype(np.int64)))
    first = np.where(swap[:, None], b, a)
    second = np.where(swap[:, None], a, b)
    scale = np.array([31, 63, 31], np.float32)
    return np.concatenate([first / scale, second / scale], 1).astype(np.float32)
