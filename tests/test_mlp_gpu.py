"""Learned DXT1 endpoint proposal on the GPU against the oracle
(oracle/mlp.py): features, the strided GEMM, forward/backward gradients of the
training kernels, the bfloat16 matrix-core inference, the exact refinement
error, and an ablation showing what the trained network gains."""
import numpy as np
import pytest
import torch

from oracle import mlp as omlp
from crunch2_amd import dxt1, learned, texture
from tests.hc_cases import blocks_from_texture

pytestmark = pytest.mark.gpu

W = np.array([8, 25, 1], np.int64)


def dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def test_features(gpu):
    blocks = blocks_from_texture(64, 1)
    f = learned.features(dev(blocks)).cpu().numpy()
    assert np.allclose(f, omlp.features(blocks), atol=2e-6)


@pytest.mark.parametrize("M,N,K,ks,ra,rb", [(1, 1, 1, 1, 0, 0), (70, 130, 57, 1, 1, 1), (128, 7, 5000, 8, 0, 0),
                                           (33, 128, 128, 1, 0, 1), (1, 128, 3000, 4, 0, 0)])
def test_gemm(gpu, M, N, K, ks, ra, rb):
    print("gemm", M, N, K, ks, flush=True)
    rng = np.random.default_rng(M * 7 + N)
    A = rng.standard_normal((K, M)).astype(np.float32)     # stored transposed: A(m, k) = A[k, m]
    B = rng.standard_normal((K, N)).astype(np.float32)
    C = torch.zeros((M, N), dtype=torch.float32, device="cuda")
    learned.gemm(M, N, K, dev(A), 1, M, dev(B), N, 1, C, N, round_a=ra, round_b=rb, ksplit=ks)
    a = omlp.bf16(A) if ra else A
    b = omlp.bf16(B) if rb else B
    ref = a.T.astype(np.float64) @ b.astype(np.float64)
    assert np.allclose(C.cpu().numpy(), ref, rtol=1e-4, atol=1e-3 * np.sqrt(K))


def training_batch(n, seed):
    blocks = dxt1.image_blocks(texture.synth(128, 128, seed, texture.RGB))[:n]
    rng = np.random.default_rng(seed)
    t = rng.random((len(blocks), 6)).astype(np.float32)
    return blocks, t


def test_gradients_match_oracle(gpu):
    blocks, t = training_batch(512, 3)
    theta = omlp.init_params(7)
    tr = learned.Trainer(dev(theta))
    x = learned.features(dev(blocks))
    loss, g = tr.gradients(x, dev(t))
    torch.cuda.synchronize()
    ref_loss, ref_g = omlp.loss_and_grads(theta, x.cpu().numpy(), t)
    g = g.cpu().numpy()
    print("loss", float(loss.item()), ref_loss, flush=True)
    assert abs(float(loss.item()) - ref_loss) <= 1e-4 * max(1.0, abs(ref_loss))
    cos = float(g @ ref_g / (np.linalg.norm(g) * np.linalg.norm(ref_g)))
    rel = np.abs(g - ref_g).max() / np.abs(ref_g).max()
    print("gradient cosine %.8f max rel %.2e" % (cos, rel), flush=True)
    assert cos > 0.99999 and rel < 1e-3


def test_gradients_with_given_labels_match_oracle(gpu):
    blocks, t = training_batch(512, 8)
    theta = omlp.init_params(9)
    lab = (np.random.default_rng(8).random(len(blocks)) < 0.3).astype(np.float32)
    tr = learned.Trainer(dev(theta))
    x = learned.features(dev(blocks))
    loss, g = tr.gradients(x, dev(t), dev(lab))
    torch.cuda.synchronize()
    ref_loss, ref_g = omlp.loss_and_grads(theta, x.cpu().numpy(), t, labels=lab)
    g = g.cpu().numpy()
    assert abs(float(loss.item()) - ref_loss) <= 1e-4 * max(1.0, abs(ref_loss))
    cos = float(g @ ref_g / (np.linalg.norm(g) * np.linalg.norm(ref_g)))
    assert cos > 0.99999 and np.abs(g - ref_g).max() / np.abs(ref_g).max() < 1e-3
    # the labels reach the confidence output: its bias gradient changes with them
    _, g2 = tr.gradients(x, dev(t), dev(1.0 - lab))
    assert abs(float(g2[-1].item()) - float(g[-1])) > 1e-6


def test_inference_matches_oracle(gpu):
    blocks, _ = training_batch(700, 4)
    theta = omlp.init_params(11)
    y = learned.infer(dev(theta), dev(blocks)).cpu().numpy()
    ref, _ = omlp.forward(theta, omlp.features(blocks))
    d = np.abs(y[:, :7] - ref)
    print("inference max abs diff %.3e" % d.max(), flush=True)
    assert d.max() < 2e-2 and np.median(d) < 1e-3


def np_block_error(px, c0, c1):
    def ex(c):
        r, g, b = (c >> 11) & 31, (c >> 5) & 63, c & 31
        return np.array([(r << 3) | (r >> 2), (g << 2) | (g >> 4), (b << 3) | (b >> 2)], np.int64)
    p0, p1 = ex(int(c0)), ex(int(c1))
    if c0 > c1:
        pal = [p0, p1, (2 * p0 + p1) // 3, (p0 + 2 * p1) // 3]
    else:
        pal = [p0, p1, (p0 + p1) // 2, np.zeros(3, np.int64)]
    d = np.stack([(((px[:, :3].astype(np.int64) - q) ** 2) * W).sum(1) for q in pal], 1)
    return int(d.min(1).sum())


def test_refinement_is_exact_and_locally_optimal(gpu):
    blocks, _ = training_batch(300, 5)
    rng = np.random.default_rng(6)
    y = np.zeros((len(blocks), 8), np.float32)
    y[:, :6] = rng.random((len(blocks), 6))
    out = learned.refine(dev(blocks), dev(y), max_iters=64).cpu().numpy().view(np.uint32)
    for b in range(len(blocks)):
        c0, c1 = out[b, 0] & 0xFFFF, out[b, 0] >> 16
        assert out[b, 2] == np_block_error(blocks[b], c0, c1)
        if out[b, 3] < 64:       # converged: no single +-1 channel step improves
            for ci, c in enumerate((c0, c1)):
                for sh, bits in ((11, 5), (5, 6), (0, 5)):
                    v = (int(c) >> sh) & ((1 << bits) - 1)
                    for d in (-1, 1):
                        if 0 <= v + d < (1 << bits):
                            nc = (int(c) & ~(((1 << bits) - 1) << sh)) | ((v + d) << sh)
                            e0, e1 = (nc, c1) if ci == 0 else (c0, nc)
                            e0, e1 = max(e0, e1), min(e0, e1)
                            assert np_block_error(blocks[b], e0, e1) >= out[b, 2]


def train_small(seed=0, steps=300, n=262144, batch=8192):
    imgs = [texture.synth(512, 512, 100 + s, texture.RGB) for s in range(n // 16384)]
    blocks = np.concatenate([dxt1.image_blocks(im) for im in imgs])
    tb = dev(blocks)
    t, _ = learned.plain_targets(tb)
    x = learned.features(tb)
    tr = learned.Trainer(seed=seed, lr=2e-3)
    g = torch.Generator(device="cuda").manual_seed(seed)
    for s in range(steps):
        idx = torch.randint(0, x.shape[0], (batch,), device="cuda", generator=g)
        tr.step(x[idx].contiguous(), t[idx].contiguous())
    return tr


def test_training_gains_over_untrained_and_box_start(gpu):
    """Ablation on held-out blocks.  The method without the network refines
    from the block's brightest/darkest texels; with it, the refinement also
    starts from the network's proposal and keeps the better.  The trained
    network must lower the total error and raise the share of blocks that
    reach the plain search's error; an untrained network must gain less."""
    tr = train_small(steps=1500)
    test = dev(dxt1.image_blocks(texture.synth(256, 256, 999, texture.RGB)))
    plain = dxt1.encode_blocks_device(test, dxt1.UBER, learned.LABEL_FLAGS)
    perr, _ = learned.block_errors(test, plain.view(torch.int32).reshape(-1, 6)[:, 2].contiguous())
    runs = {
        "box only": learned.refine(test, None, 8, learned.START_BOX),
        "untrained + box": learned.refine(test, learned.infer(learned.init_params(0), test), 8,
                                          learned.START_NET | learned.START_BOX),
        "trained + box": learned.refine(test, learned.infer(tr.theta, test), 8,
                                        learned.START_NET | learned.START_BOX),
        "trained only": learned.refine(test, learned.infer(tr.theta, test), 8, learned.START_NET),
    }
    res = {}
    for k, r in runs.items():
        res[k] = (float(r[:, 2].double().sum()), float((r[:, 2] <= perr).float().mean()))
        print(k, "total error %.5g, blocks at or below plain %.4f" % res[k], flush=True)
    assert res["trained + box"][0] < res["box only"][0]
    assert res["trained + box"][1] > res["box only"][1]
    gain_trained = res["box only"][0] - res["trained + box"][0]
    gain_untrained = res["box only"][0] - res["untrained + box"][0]
    assert gain_trained > gain_untrained


def test_guarded_encode_never_worse_than_plain(gpu):
    tr = train_small(steps=100)
    test = dev(dxt1.image_blocks(texture.synth(128, 128, 555, texture.RGB)))
    ends, sels, errs, nplain = learned.encode(tr.theta, test, guarded=True)
    plain = dxt1.encode_blocks_device(test, dxt1.UBER, learned.LABEL_FLAGS)
    perr, _ = learned.block_errors(test, plain.view(torch.int32).reshape(-1, 6)[:, 2].contiguous())
    assert nplain == len(test)
    assert bool((errs <= perr).all())
    e2, s2 = learned.block_errors(test, ends)
    assert bool((e2 == errs).all()) and bool((s2 == sels).all())
