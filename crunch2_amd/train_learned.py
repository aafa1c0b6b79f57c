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
    """Features (n x 57) and targets (n x 6) on the device."""
    per = (side // 4) ** 2
    xs, ts = [], []
    got = 0
    k = 0
    while got < n_blocks:
        img = texture.synth(side, side, seed0 + k, texture.RGB)
        b = torch.from_numpy(dxt1.image_blocks(img)).cuda()
        t, _ = learned.plain_targets(b)
        xs.append(learned.features(b))
        ts.append(t)
        got += per
        k += 1
        if k % 16 == 0:
            log("labelled %d blocks" % got)
    return torch.cat(xs)[:n_blocks].contiguous(), torch.cat(ts)[:n_blocks].contiguous()


def train(n_blocks=10_000_000, steps=6000, batch=8192, lr=2e-3, seed=0):
    t0 = time.time()
    x, t = make_dataset(n_blocks)
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
    return tr.theta


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--blocks", type=int, default=10_000_000)
    ap.add_argument("--steps", type=int, default=6000)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--out", default=DEFAULT_OUT)
    a = ap.parse_args()
    theta = train(a.blocks, a.steps, a.batch)
    os.makedirs(os.path.dirname(a.out), exist_ok=True)
    np.save(a.out, theta.cpu().numpy())
    log("wrote", a.out)


if __name__ == "__main__":
    main()
