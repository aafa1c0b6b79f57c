"""Trains the learned DXT1 endpoint proposal on the GPU with the project's own
kernels (crunch2_amd/learned.py) and writes its float32 parameters.

    python -m crunch2_amd.train_learned [--blocks 10000000] [--steps 6000] [--out crunch2_amd/data/dxt1_proposal.npy]

Training data: seeded synthetic RGB textures (crn2_synth_rgba), cut into
4x4 blocks; labels are the endpoints the plain GPU search
(crn2_dxt1_blocks_device, uber quality, perceptual, four-colour blocks)
chooses for them.
"""
import argparse
import os
import sys
import time

import numpy as np
import torch

from . import dxt1, learned, texture

DEFAULT_OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "dxt1_proposal.npy")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_dataset(n_blocks, side=1024, seed0=1000):
    """Blocks (n x 16 x 4 u8), features (n x 57), targets (n x 6) and the
    plain search's exact errors (n), on the device."""
    per = (side // 4) ** 2
    bs, xs, ts, es = [], [], [], []
    got = 0
    k = 0
    while got < n_blocks:
        img = texture.synth(side, side, seed0 + k, texture.RGB)
        b = torch.from_numpy(dxt1.image_blocks(img)).cuda()
        t, raw = learned.plain_targets(b)
        err, _ = learned.block_errors(b, raw.view(torch.int32).reshape(-1, 6)[:, 2].contiguous())
        bs.append(b)
        xs.append(learned.features(b))
        ts.append(t)
        es.append(err)
        got += per
        k += 1
        if k % 16 == 0:
            log("labelled %d blocks" % got)
    cut = lambda a: torch.cat(a)[:n_blocks].contiguous()
    return cut(bs), cut(xs), cut(ts), cut(es)


def confidence_labels(theta, blocks, plain_err, chunk=1 << 20):
    """1 where the learned encode (proposal + refinement, no fallback) ends no
    worse than the plain search, else 0."""
    out = []
    for i in range(0, blocks.shape[0], chunk):
        b = blocks[i:i + chunk]
        y = learned.infer(theta, b)
        r = learned.refine(b, y, learned.ENCODE_ITERS, learned.ENCODE_STARTS)
        out.append((r[:, 2] <= plain_err[i:i + chunk]).float())
    return torch.cat(out).contiguous()


def train(n_blocks=10_000_000, steps=6000, batch=8192, lr=2e-3, seed=0, conf_steps=2000):
    """Endpoint regression (with the one-step confidence label) for `steps`,
    then `conf_steps` more with the confidence output fitted to whether the
    refined proposal matches or beats the plain search."""
    t0 = time.time()
    blocks, x, t, perr = make_dataset(n_blocks)
    torch.cuda.synchronize()
    log("dataset: %d blocks in %.1fs" % (x.shape[0], time.time() - t0))
    tr = learned.Trainer(seed=seed, lr=lr)
    g = torch.Generator(device="cuda").manual_seed(seed)
    t1 = time.time()
    for s in range(steps):
        if s == int(steps * 0.7):
            tr.lr = lr * 0.3
        if s == int(steps * 0.9):
            tr.lr = lr * 0.1
        idx = torch.randint(0, x.shape[0], (batch,), device="cuda", generator=g)
        loss = tr.step(x[idx].contiguous(), t[idx].contiguous())
        if s % 500 == 0 or s == steps - 1:
            log("step %d loss %.5f (%.1fs)" % (s, float(loss.item()), time.time() - t1))
    torch.cuda.synchronize()
    log("trained %d steps of %d in %.1fs" % (steps, batch, time.time() - t1))
    if conf_steps:
        lab = confidence_labels(tr.theta, blocks, perr)
        log("refined proposal no worse than plain on %.4f of the blocks" % float(lab.mean().item()))
        tr.lr = lr * 0.3
        for s in range(conf_steps):
            if s == int(conf_steps * 0.8):
                tr.lr = lr * 0.1
            idx = torch.randint(0, x.shape[0], (batch,), device="cuda", generator=g)
            loss = tr.step(x[idx].contiguous(), t[idx].contiguous(), lab[idx].contiguous())
            if s % 500 == 0 or s == conf_steps - 1:
                log("confidence step %d loss %.5f" % (s, float(loss.item())))
        conf = torch.cat([learned.infer(tr.theta, blocks[i:i + (1 << 20)])[:, 6]
                          for i in range(0, blocks.shape[0], 1 << 20)])
        order = torch.argsort(conf)
        for q in (0.1, 0.2, 0.3):
            k = int(q * blocks.shape[0])
            caught = float((1.0 - lab[order[:k]]).sum().item()) / max(1.0, float((1.0 - lab).sum().item()))
            log("lowest-confidence %.0f%% of blocks hold %.3f of the blocks the proposal loses" % (100 * q, caught))
    return tr.theta


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--blocks", type=int, default=10_000_000)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--conf-steps", type=int, default=2000)
    ap.add_argument("--out", default=DEFAULT_OUT)
    a = ap.parse_args()
    theta = train(a.blocks, a.steps, a.batch, conf_steps=a.conf_steps)
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    np.save(a.out, theta.cpu().numpy())
    log("wrote", a.out)


if __name__ == "__main__":
    main()
