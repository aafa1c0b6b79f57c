"""Benchmark: CRN->DXTn transcode of a 4096x4096 RGBA8 texture with its full
13-level mip chain, in DXT1 and DXT5 (one step = both files, all levels, in
128 independent copies), inputs resident in HBM, on one MI355X.  Beside it,
the encode workloads: plain DXT1 of a 256^2 texture, the learned vs plain
DXT1 encode of an 8192^2 texture, dxt_hc of a 2048^2 texture and the
512-texture serving set (encode and transcode).

Output: one JSON line (see the keys below).  The CPU baselines are the
project's C restatements (oracle/) timed on this host.

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes the walk kernel moves per copy of the 4096^2 DXT1 + DXT5 pair
# (its FETCH_SIZE + WRITE_SIZE, scaled by a device copy of known size), for
# the default inputs (GPU dxt_hc + the project's writer)
WALK_TRAFFIC_PER_COPY = 370.3e6
STAGES = ("crnd_tables_kernel", "crnd_walk_kernel", "crnd_sym_kernel<even>", "crnd_sym_kernel<odd>",
          "crnd_index_{chunk,carry}_kernel", "crnd_index_final_kernel (+ block write)")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def decode_sse(blocks, ends, sels):
    """Uniform RGB squared error of DXT1 blocks (standard decode) per block."""
    import torch
    c = torch.stack([ends & 0xFFFF, (ends >> 16) & 0xFFFF], 1)
    r, g, b = (c >> 11) & 31, (c >> 5) & 63, c & 31
    p = torch.stack([(r << 3) | (r >> 2), (g << 2) | (g >> 4), (b << 3) | (b >> 2)], 2)     # n, 2, 3
    four = (c[:, 0] > c[:, 1])[:, None]
    p2 = torch.where(four, (2 * p[:, 0] + p[:, 1]) // 3, (p[:, 0] + p[:, 1]) // 2)
    p3 = torch.where(four, (p[:, 0] + 2 * p[:, 1]) // 3, torch.zeros_like(p[:, 0]))
    pal = torch.stack([p[:, 0], p[:, 1], p2, p3], 1)                                            # n, 4, 3
    idx = (sels[:, None] >> (2 * torch.arange(16, device=sels.device))[None, :]) & 3           # n, 16
    dec = torch.gather(pal, 1, idx[:, :, None].expand(-1, -1, 3).long())
    d = dec - blocks[:, :, :3].long()
    return (d * d).sum((1, 2))


def psnr(sse, npix):
    import math
    mse = float(sse) / (3.0 * npix)
    return 99.0 if mse == 0 else 20.0 * math.log10(255.0 / math.sqrt(mse))


VALU_PEAK_TOPS = 39.3      # int32 lane-ops/s: 256 CUs x 64 lanes x 2.4 GHz (no packed int32 ops)
MFMA_BF16_PEAK_TFLOPS = 2500.0
MFMA_INT8_PEAK_TOPS = 5000.0
OPS_PER_DISTANCE = 9      # 3 differences, 3 squares, 2 weight multiplies and adds folded, compare-select


def gpu_ms(fn, reps):
    """Average duration of fn's launches on torch's current stream (HIP events)."""
    import torch
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


def cpu_threads():
    return min(16, os.cpu_count() or 1)


def oracle_dxt1_rate(blocks, quality, flags, seconds):
    """The oracle's DXT1 search on host threads (one slice of blocks per
    thread, ctypes calls outside the GIL), repeated for about `seconds`."""
    import numpy as np
    import oracle
    from concurrent.futures import ThreadPoolExecutor
    th = cpu_threads()
    parts = np.array_split(blocks, th)
    done, t = 0, 0.0
    with ThreadPoolExecutor(max_workers=th) as pool:
        while t < seconds:
            s0 = time.perf_counter()
            list(pool.map(lambda b: oracle.dxt1_blocks(b, quality, flags), parts))
            t += time.perf_counter() - s0
            done += len(blocks)
    return done * 16 / t / 1e6, th, t


def plain_leg(size, cpu_seconds):
    """Plain DXT1 block-by-block compression of one size x size RGB texture,
    uber, perceptual, both block types: 8-byte blocks + per-block error from
    one image launch; every block checked against the oracle; VALU roofline
    from the oracle's count of colour distances."""
    import numpy as np
    import torch
    import oracle
    from crunch2_amd import dxt1, texture
    from crunch2_amd._lib import lib
    img = texture.synth(size, size, 1, texture.RGB)
    blocks = dxt1.image_blocks(img)
    n = len(blocks)
    flags = dxt1.PERCEPTUAL | dxt1.ALPHA_BLOCKS
    src = torch.from_numpy(np.ascontiguousarray(img)).cuda()
    out = torch.empty(n, dtype=torch.int64, device="cuda")
    err = torch.empty(n, dtype=torch.int64, device="cuda")
    st = torch.cuda.current_stream().cuda_stream

    def run():
        assert lib().crn2_encode_dxt1_image_device(src.data_ptr(), size, size, 4, flags, out.data_ptr(), 1,
                                                   err.data_ptr(), C.c_void_p(st)) == 0
    run()
    torch.cuda.synchronize()
    ms = gpu_ms(run, 20)
    ref = oracle.dxt1_blocks(blocks, 4, flags, count=True)
    w = out.cpu().numpy().view(np.uint64)
    rsel = (ref[0].astype(np.uint64) << (2 * np.arange(16, dtype=np.uint64))).sum(1)
    exp = ref[2].astype(np.uint64) | (ref[3].astype(np.uint64) << 16) | (rsel << 32)
    exact = float((w == exp).mean())
    assert exact == 1.0, "plain DXT1 differs from the oracle"
    assert np.array_equal(err.cpu().numpy().view(np.uint64), ref[1])
    ops = ref[7] * OPS_PER_DISTANCE
    cpu, th, ct = oracle_dxt1_rate(blocks, 4, flags, cpu_seconds)
    achieved = ops / (ms * 1e-3) / 1e12
    return {"workload": "plain DXT1 of one %dx%d RGB texture (%d blocks), uber, perceptual, both block types, "
                        "8-byte blocks + per-block error" % (size, size, n),
            "mtexels_s": round(n * 16 / ms * 1e-3, 2), "ms": round(ms, 4), "bit_exact_rate": exact,
            "roofline": {"bound": "valu", "kernel": "dxt1_image_kernel", "achieved": round(achieved, 4),
                         "peak": VALU_PEAK_TOPS, "unit": "Tops/s (int32 lane-ops)",
                         "frac": round(achieved / VALU_PEAK_TOPS, 6), "traffic": None,
                         "ops_per_launch": int(ops), "colour_distances": int(ref[7])},
            "cpu_baseline": {"value": round(cpu, 3), "unit": "Mtexels/s", "cores": th, "kind": "port",
                             "sample": "the oracle's search over the same %d blocks on %d threads (%.1fs)"
                                       % (n, th, ct)}}


def encode_leg(size, frac, cpu_seconds):
    """DXT1 encode of one size x size RGB texture: the plain GPU search against
    the learned proposal (+ refinement, + the plain search for the `frac` of
    blocks with the lowest confidence)."""
    import numpy as np
    import torch
    import oracle
    from crunch2_amd import dxt1, learned, texture
    from crunch2_amd.train_learned import DEFAULT_OUT
    if not os.path.exists(DEFAULT_OUT):
        return {"skipped": "no trained parameters at " + DEFAULT_OUT}
    theta = torch.from_numpy(np.load(DEFAULT_OUT)).cuda()
    img = texture.synth(size, size, 4242, texture.RGB)
    nb = dxt1.image_blocks(img)
    blocks = torch.from_numpy(nb).cuda()
    n = blocks.shape[0]
    texels = n * 16
    # warm both paths on a slice
    dxt1.encode_blocks_device(blocks[:4096].contiguous(), dxt1.UBER, learned.LABEL_FLAGS)
    learned.encode(theta, blocks[:4096].contiguous(), fallback_frac=frac)
    torch.cuda.synchronize()
    raw = torch.empty((n, 24), dtype=torch.uint8, device="cuda")
    t_plain = gpu_ms(lambda: dxt1.encode_blocks_device(blocks, dxt1.UBER, learned.LABEL_FLAGS, out=raw), 1) * 1e-3
    pend = raw.view(torch.int32).reshape(-1, 6)[:, 2].contiguous()
    perr, psel = learned.block_errors(blocks, pend)
    t0 = time.perf_counter()
    ends, sels, errs, nfb = learned.encode(theta, blocks, fallback_frac=frac)
    torch.cuda.synchronize()
    t_learn = time.perf_counter() - t0
    mlp_ms = gpu_ms(lambda: learned.infer(theta, blocks), 3)
    sse_p = int(decode_sse(blocks, pend, psel).sum())
    sse_l = int(decode_sse(blocks, ends, sels).sum())
    ps, pl = psnr(sse_p, texels), psnr(sse_l, texels)
    # algorithmic work of the plain search from the oracle on a sample of the blocks
    idx = np.random.default_rng(0).choice(n, 4096, replace=False)
    smp = nb[idx]
    ref = oracle.dxt1_blocks(smp, 4, oracle.DXT1_PERCEPTUAL, count=True)
    got = raw[torch.from_numpy(idx).cuda()].cpu().numpy().view(np.uint32).reshape(-1, 6)
    assert np.array_equal(got[:, 2] & 0xFFFF, ref[2]) and np.array_equal(got[:, 2] >> 16, ref[3]), \
        "plain search differs from the oracle on the sample"
    ops = ref[7] / 4096 * n * OPS_PER_DISTANCE
    valu = ops / t_plain / 1e12
    mlp_flops = 2.0 * n * (57 * 128 + 128 * 128 + 128 * 7)
    mlp_tf = mlp_flops / (mlp_ms * 1e-3) / 1e12
    cpu, th, ct = oracle_dxt1_rate(smp, 4, oracle.DXT1_PERCEPTUAL, cpu_seconds)
    return {"workload": "DXT1 encode of one %dx%d RGB texture (%d blocks), uber, perceptual, four-colour blocks"
                        % (size, size, n),
            "plain_mtexels_s": round(texels / t_plain / 1e6, 1), "learned_mtexels_s": round(texels / t_learn / 1e6, 1),
            "speedup": round(t_plain / t_learn, 3), "psnr_plain_db": round(ps, 4), "psnr_learned_db": round(pl, 4),
            "delta_db": round(pl - ps, 4), "blocks_worse_than_plain": int((errs > perr).sum()),
            "fallback_blocks": nfb, "fallback_frac": frac,
            "roofline": {"bound": "valu", "kernel": "dxt1_blocks_kernel (plain search)", "achieved": round(valu, 4),
                         "peak": VALU_PEAK_TOPS, "unit": "Tops/s (int32 lane-ops)",
                         "frac": round(valu / VALU_PEAK_TOPS, 6), "traffic": None,
                         "ops_per_launch": ops,
                         "ops_from": "the oracle's colour-distance count on 4,096 sampled blocks, scaled to %d" % n},
            "mlp_roofline": {"bound": "mfma", "kernel": "mlp_infer_kernel (bf16 MFMA, fp32 accumulation)",
                             "achieved": round(mlp_tf, 2), "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                             "frac": round(mlp_tf / MFMA_BF16_PEAK_TFLOPS, 5), "ms": round(mlp_ms, 3),
                             "flops_per_launch": mlp_flops},
            "cpu_baseline": {"value": round(cpu, 3), "unit": "Mtexels/s", "cores": th, "kind": "port",
                             "sample": "the oracle's plain search over 4,096 sampled blocks of the texture on %d "
                                       "threads (%.1fs)" % (th, ct)}}


def hc_leg(size, cpu):
    """dxt_hc of one size x size RGB texture (4096 colour endpoints, 8192
    colour selectors, hierarchical, no mips) on the GPU, the colour selector
    assignment on INT8 matrix cores timed alone, and (cpu) the oracle's
    single-thread composition of the same texture, whose output is also
    compared with the GPU's."""
    import numpy as np
    import torch
    from crunch2_amd import dxt1, hc, texture
    img = texture.synth(size, size, 11, texture.RGB)
    blocks = dxt1.image_blocks(img)
    n = blocks.shape[0]
    levels = [(0, n, size // 4, 1.0)]
    sizes = (4096, 8192, 3072, 3072)
    tb = torch.from_numpy(blocks).cuda()
    hc.compress(tb, levels, codebook_sizes=sizes)       # warm-up
    reps = 3
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = hc.compress(tb, levels, codebook_sizes=sizes)
    torch.cuda.synchronize()
    t = (time.perf_counter() - t0) / reps
    ep, sel = out[0].long(), out[1].long()
    ce, cs = out[2].long() & 0xFFFFFFFF, out[3].long() & 0xFFFFFFFF
    # decode (colour selectors in linear order -> DXT order) for the PSNR
    lin2dxt = torch.tensor([0, 2, 3, 1], device=tb.device)
    e = ce[ep[:, 0]]
    s_ = cs[sel[:, 0]]
    codes = (s_[:, None] >> (2 * torch.arange(16, device=tb.device))[None, :]) & 3
    raw = lin2dxt[codes]
    packed = torch.zeros_like(s_)
    for i in range(16):
        packed |= raw[:, i] << (2 * i)
    sse = int(decode_sse(tb, e, packed).sum())
    # the INT8 matrix-core assignment alone, on the final colour selectors against per-block cluster colours
    used_sel = out[3]
    m = used_sel.shape[0]
    best = torch.empty(n, dtype=torch.int32, device=tb.device)
    cols = torch.zeros((ce.shape[0], 4, 4), dtype=torch.uint8, device=tb.device)
    c565 = torch.stack([ce & 0xFFFF, ce >> 16], 1)
    r, g, b = (c565 >> 11) & 31, (c565 >> 5) & 63, c565 & 31
    p = torch.stack([(r << 3) | (r >> 2), (g << 2) | (g >> 4), (b << 3) | (b >> 2)], 2)
    lin = torch.stack([p[:, 0], (2 * p[:, 0] + p[:, 1]) // 3, (p[:, 0] + 2 * p[:, 1]) // 3, p[:, 1]], 1)
    cols[:, :, :3] = lin.to(torch.uint8)
    ei = out[0][:, 0].to(torch.int32).contiguous()
    hc.color_selectors(tb, ei, cols, used_sel)
    sel_ms = gpu_ms(lambda: hc.color_selectors(tb, ei, cols, used_sel), 5)
    ops = 2.0 * n * m * 64 * 4
    tops = ops / (sel_ms * 1e-3) / 1e12
    line = {"workload": "dxt_hc of one %dx%d RGB texture (%d blocks), 4096 colour endpoints / 8192 colour "
                        "selectors, perceptual, no mips" % (size, size, n),
            "mtexels_s": round(n * 16 / t / 1e6, 2), "ms": round(t * 1e3, 2),
            "color_endpoints": int(ce.shape[0]), "color_selectors": int(m),
            "psnr_db": round(psnr(sse, n * 16), 3),
            "roofline": {"bound": "mfma", "kernel": "colour selector assignment (v_mfma_i32_32x32x32_i8, blocks x "
                                                    "codewords x 64, exact int32)",
                         "achieved": round(tops, 1), "peak": MFMA_INT8_PEAK_TOPS, "unit": "TOP/s",
                         "frac": round(tops / MFMA_INT8_PEAK_TOPS, 4), "ms": round(sel_ms, 3), "ops": ops,
                         "traffic": None}}
    if cpu:
        import oracle
        t0 = time.perf_counter()
        ref = oracle.hc_compress(blocks, levels, codebook_sizes=sizes)
        ct = time.perf_counter() - t0
        same = (np.array_equal(out[0].cpu().numpy().view(np.uint16), ref[0]) and
                np.array_equal(out[1].cpu().numpy().view(np.uint16), ref[1]) and
                np.array_equal(out[2].cpu().numpy().view(np.uint32), ref[2]) and
                np.array_equal(out[3].cpu().numpy().view(np.uint32), ref[3]))
        assert same, "dxt_hc differs from the oracle"
        line["bit_exact"] = same
        line["cpu_baseline"] = {"value": round(n * 16 / ct / 1e6, 3), "unit": "Mtexels/s", "cores": 1,
                                "kind": "port",
                                "sample": "the oracle's dxt_hc composition of the same %dx%d texture on one thread "
                                          "(%.1fs); its output equals the GPU's" % (size, size, ct)}
    return line


def mixed_leg(count, cpu_seconds):
    """The serving-style set: `count` textures, sides log-uniform in [64, 2048]
    with full mip chains, 60 % RGB -> DXT1, 25 % RGBA -> DXT5, 15 % normal
    maps -> DXN_XY.  Encode: every level image of the set in one graph replay
    of the texture-set encoder (uber, perceptual, both block types).
    Transcode: the set's .crn files (simple-quantizer writer, to keep the leg
    short) in one batch plan."""
    import numpy as np
    import torch
    import oracle
    from concurrent.futures import ThreadPoolExecutor
    from crunch2_amd import crnd, texture
    from crunch2_amd._lib import FMT_DXN_XY, FMT_DXT1, FMT_DXT5, EncodeItem, lib
    rng = np.random.default_rng(2024)
    jobs = []
    for i in range(count):
        side = int(round(2 ** rng.uniform(6, 11)))
        u = rng.random()
        jobs.append((FMT_DXT1 if u < 0.60 else (FMT_DXT5 if u < 0.85 else FMT_DXN_XY), side, 5000 + i))
    kinds = {FMT_DXT1: texture.RGB, FMT_DXT5: texture.RGBA, FMT_DXN_XY: texture.NORMAL}
    t0 = time.time()
    with ThreadPoolExecutor(max_workers=cpu_threads()) as pool:
        chains = list(pool.map(lambda j: texture.mip_chain(texture.synth(j[1], j[1], j[2], kinds[j[0]])), jobs))
    log("mixed set: %d mip chains in %.1fs" % (len(chains), time.time() - t0))
    items, keep, texels, out_bytes = [], [], 0, 0
    for (fmt, side, _), chain in zip(jobs, chains):
        for lv in chain:
            h, w = lv.shape[:2]
            src = torch.from_numpy(np.ascontiguousarray(lv).reshape(-1)).cuda()
            nbytes = ((w + 3) // 4) * ((h + 3) // 4) * (8 if fmt == FMT_DXT1 else 16)
            dst = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
            keep.append((src, dst))
            items.append(EncodeItem(src.data_ptr(), w, h, fmt, dst.data_ptr()))
            texels += ((w + 3) // 4) * ((h + 3) // 4) * 16
            out_bytes += nbytes
    arr = (EncodeItem * len(items))(*items)
    es = C.c_void_p()
    assert lib().crn2_encode_set_create(arr, len(items), 4, 1 | 2, C.byref(es)) == 0
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib().crn2_encode_set_run_graph(es, st) == 0      # capture + first replay
    torch.cuda.synchronize()
    enc_ms = gpu_ms(lambda: lib().crn2_encode_set_run_graph(es, st), 2)
    assert lib().crn2_encode_set_destroy(es) == 0
    # transcode of the set's files
    t0 = time.time()
    with ThreadPoolExecutor(max_workers=cpu_threads()) as pool:
        files = list(pool.map(lambda j: texture.synth_crn((j[1], j[1]), j[2], j[0]), jobs))
    log("mixed set: %d .crn files in %.1fs" % (len(files), time.time() - t0))
    ctxs, outs = [], []
    tex_t, crn_bytes, dxt_bytes = 0, 0, 0
    for d in files:
        ctx = crnd.unpack_begin(d)
        assert ctx is not None
        inf = crnd.get_texture_info(d)
        nb = crnd.levels_size(ctx, 0, inf["levels"])
        outs.append(torch.empty(nb, dtype=torch.uint8, device="cuda"))
        ctxs.append(ctx)
        crn_bytes += len(d)
        dxt_bytes += nb
        for l in range(inf["levels"]):
            li = crnd.get_level_info(d, l)
            tex_t += li["blocks_x"] * li["blocks_y"] * 16
    k = len(ctxs)
    plan = C.c_void_p()
    assert lib().crn2_batch_create((C.c_void_p * k)(*[c.handle for c in ctxs]), k,
                                   (C.c_void_p * k)(*[o.data_ptr() for o in outs]),
                                   (C.c_uint64 * k)(*[o.numel() for o in outs]), C.byref(plan)) == 0
    assert lib().crn2_batch_run(plan, st) == 0
    torch.cuda.synchronize()
    for i in (0, k // 2, k - 1):
        ref = oracle.unpack_all(files[i])
        exp = np.concatenate([ref[l][0].reshape(-1) for l in sorted(ref)])
        assert np.array_equal(outs[i].cpu().numpy(), exp), "mixed-set transcode differs from the oracle"
    tr_ms = gpu_ms(lambda: lib().crn2_batch_run(plan, st), 3)
    assert lib().crn2_batch_status(plan, st) == 0
    lib().crn2_batch_destroy(plan)
    for c in ctxs:
        crnd.unpack_end(c)
    # CPU baselines: the oracle's DXT1 search on a sample of the set's DXT1 blocks
    from crunch2_amd import dxt1
    dx = [c[0] for j, c in zip(jobs, chains) if j[0] == FMT_DXT1][:8]
    smp = np.concatenate([dxt1.image_blocks(x) for x in dx])[:4096]
    cpu_enc, th, ct = oracle_dxt1_rate(smp, 4, oracle.DXT1_PERCEPTUAL | oracle.DXT1_ALPHA_BLOCKS, cpu_seconds)
    gbs = (crn_bytes + dxt_bytes) / (tr_ms * 1e-3) / 1e9
    return {"workload": "serving-style set of %d textures (sides 64..2048 log-uniform, full mips; %d DXT1 / %d DXT5 / "
                        "%d DXN_XY), %d level images" % (count, sum(j[0] == FMT_DXT1 for j in jobs),
                                                         sum(j[0] == FMT_DXT5 for j in jobs),
                                                         sum(j[0] == FMT_DXN_XY for j in jobs), len(items)),
            "encode": {"mtexels_s": round(texels / (enc_ms * 1e-3) / 1e6, 1), "ms": round(enc_ms, 2),
                       "texels": texels, "dxtn_bytes": out_bytes,
                       "how": "one HIP graph replay of the texture-set encoder (uber, perceptual, both block types)",
                       "cpu_baseline": {"value": round(cpu_enc, 3), "unit": "Mtexels/s", "cores": th, "kind": "port",
                                        "sample": "the oracle's DXT1 search on 4,096 blocks of the set's largest "
                                                  "DXT1 images on %d threads (%.1fs)" % (th, ct)}},
            "transcode": {"mtexels_s": round(tex_t / (tr_ms * 1e-3) / 1e6, 1), "ms": round(tr_ms, 3),
                          "gb_s": round(gbs, 2), "hbm_frac": round(gbs / HBM_PEAK_GBS, 5),
                          "crn_bytes": crn_bytes, "dxtn_bytes": dxt_bytes,
                          "inputs": "the simple block quantizer's .crn files at quality-128 palette sizes"}}


# HBM bytes mip_pass1_kernel moves per 8192^2 chain (FETCH_SIZE x 2 +
# WRITE_SIZE x 1, the factors a device copy of known size reads back)
MIP_PASS1_TRAFFIC = 17.834e9
FP32_PEAK_TFLOPS = 157.3   # MI355X dense fp32 (packed FMA); one multiply and one add per
                           # contributor, unfused and unpacked, reach a quarter of it


def mip_leg(size, cpu_seconds):
    """Mip generation of one size x size RGBA texture with crn_mipmap_params'
    defaults (Kaiser, gamma filtering, blurriness 0.9, clamp, alpha valid):
    levels 1..n-1, each resampled from level 0, on the GPU; the whole chain
    checked against the oracle (computed level by level on host threads, which
    is also the CPU baseline); fp32 roofline of the dominant launch (pass 1)
    from its multiply + add per contributor."""
    import numpy as np
    import torch
    import oracle
    from concurrent.futures import ThreadPoolExecutor
    from crunch2_amd import mips, texture
    img = texture.synth(size, size, 23, texture.RGBA)
    p = mips.params()
    plan = mips.Plan(size, size, p)
    src = torch.from_numpy(img).cuda()
    out = torch.empty(plan.chain_bytes, dtype=torch.uint8, device="cuda")
    for _ in range(2):
        plan.run(src, out)
    torch.cuda.synchronize()
    ms = gpu_ms(lambda: plan.run(src, out), 10)
    plan.profile(True)
    for _ in range(5):
        plan.run(src, out)
    stage, runs = plan.profile(False)
    stage = [t / runs for t in stage]
    got = plan.split(out.cpu().numpy())
    # algorithmic multiply-adds and HBM bytes of each pass: pass 1 reads the
    # level-0 byte planes once per level (4 components) and writes the float
    # intermediate; pass 2 reads the intermediate and writes the RGBA8 level
    nc = 4
    mac1 = mac2 = bytes1 = bytes2 = 0
    for l, (w, h) in enumerate(plan.sizes()[1:], 1):
        xo = len(mips.contributors(size, w)[1])
        yo = len(mips.contributors(size, h)[1])
        d = h if plan.y_first[l] else w
        bytes1 += nc * size * size + 4 * nc * size * d
        bytes2 += 4 * nc * size * d + 4 * w * h
        if plan.y_first[l]:
            mac1 += nc * size * yo
            mac2 += nc * h * xo
        else:
            mac1 += nc * size * xo
            mac2 += nc * w * yo
    achieved = bytes1 / (stage[1] * 1e-3) / 1e9
    tflops = 2 * mac1 / (stage[1] * 1e-3) / 1e12
    th = cpu_threads()
    sizes = plan.sizes()[1:]
    t0 = time.perf_counter()
    with ThreadPoolExecutor(max_workers=th) as pool:
        ref = list(pool.map(lambda wh: oracle.mip_resample(img, wh[0], wh[1])[0], sizes))
    cpu_s = time.perf_counter() - t0
    exact = all(np.array_equal(a, b) for a, b in zip(got[1:], ref)) and np.array_equal(got[0], img)
    assert exact, "GPU mip chain differs from the oracle"
    plan.close()
    return {"workload": "mip chain of one %dx%d RGBA8 texture (%d levels): Kaiser, sRGB gamma filtering, "
                        "blurriness 0.9, clamp, 4 components, every level resampled from level 0"
                        % (size, size, plan.levels),
            "level0_mtexels_s": round(size * size / (ms * 1e-3) / 1e6, 1), "ms": round(ms, 4),
            "bit_exact": exact,
            "stage_ms": {"mip_planes_kernel": round(stage[0], 4), "mip_pass1_kernel": round(stage[1], 4),
                         "mip_pass2_kernel": round(stage[2], 4)},
            "roofline": {"bound": "hbm", "kernel": "mip_pass1_kernel", "achieved": round(achieved, 1),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 5),
                         "traffic": MIP_PASS1_TRAFFIC if (size == 8192 and MIP_PASS1_TRAFFIC) else None,
                         "algorithmic_bytes_per_launch": bytes1, "launch_ms": round(stage[1], 4),
                         "flops_per_launch": 2 * mac1, "achieved_tflops": round(tflops, 3),
                         "valu_frac": round(tflops / FP32_PEAK_TFLOPS, 5),
                         "valu_ceiling_unfused_unpacked": round(FP32_PEAK_TFLOPS / 4, 2),
                         "pass2": {"algorithmic_bytes": bytes2, "flops": 2 * mac2, "launch_ms": round(stage[2], 4),
                                   "achieved": round(bytes2 / (stage[2] * 1e-3) / 1e9, 1)}},
            "cpu_baseline": {"value": round(size * size / cpu_s / 1e6, 3), "unit": "level-0 Mtexels/s",
                             "cores": th, "kind": "port",
                             "sample": "the oracle's resample of all %d generated levels of the same image, one "
                                       "level per thread on %d threads (%.1fs)" % (len(sizes), th, cpu_s)}}


def qdxt_leg(size):
    """Clustered DXT5 -- crnlib's DDS path below the top quality
    (mipmapped_texture::qdxt_pack_init / qdxt_pack,
    crnlib/crn_mipmapped_texture.cpp:2310-2560) -- of one size x size RGBA
    texture with its box mip chain: qdxt1 for the colour (no 3-colour blocks,
    codebook curve 0.75, as DXT5 asks) and qdxt5 for the alpha, quality 128,
    uber; a step is init + pack of both on the GPU.  Checked bit for bit
    against the oracle, whose single-thread run of the same step is the CPU
    baseline.  Roofline of the dominant launch (the 16-D selector-tree
    splits) from the vector bytes each of its passes reads."""
    import numpy as np
    import torch
    import oracle
    from crunch2_amd import qdxt, texture
    import threading
    blocks, desc = qdxt.level_blocks(texture.mip_chain(texture.synth(size, size, 77, texture.RGBA)))
    n = len(blocks)
    b = torch.from_numpy(blocks).cuda()
    torch.cuda.synchronize()
    P, H = qdxt.PERCEPTUAL, qdxt.HIERARCHICAL
    s_colour, s_alpha = torch.cuda.Stream(), torch.cuda.Stream()

    def colour(res):
        try:
            with torch.cuda.stream(s_colour):
                q1 = qdxt.Qdxt1(b, desc)
                res["c"] = q1.pack(qdxt.params(128, 4, P | H), 0.75)
                res["i"] = q1.info()
                q1.close()
                s_colour.synchronize()
        except BaseException as e:          # re-raised by the step
            res["e"] = e

    def step(concurrent=True):
        # colour and alpha are independent (qdxt_pack runs them one after the
        # other): one host thread and stream each, so that their serial tree
        # rounds overlap, as crn2_compress_dds_clustered runs them
        res = {}
        t = threading.Thread(target=colour, args=(res,))
        t.start()
        if not concurrent:
            t.join()
        try:
            with torch.cuda.stream(s_alpha):
                q5 = qdxt.Qdxt5(b, desc, 3)
                a = q5.pack(qdxt.params5(128, 4, qdxt.BOTH_BLOCK_TYPES))
                i5 = q5.info()
                q5.close()
                s_alpha.synchronize()
        finally:
            t.join()
        if "e" in res:
            raise res["e"]
        return res["c"], a, (res["i"], i5)

    step()
    torch.cuda.synchronize()
    reps = 3
    t0 = time.perf_counter()
    for _ in range(reps):
        c, a, info = step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    t0 = time.perf_counter()
    step(concurrent=False)
    torch.cuda.synchronize()
    serial_ms = (time.perf_counter() - t0) * 1e3
    qdxt.profile(True)
    step(concurrent=False)        # launch times of the splits without the other stream's work beside them
    torch.cuda.synchronize()
    sms, launches, passes = qdxt.profile(False)
    got_c = c.cpu().numpy().view(np.uint64)
    got_a = a.cpu().numpy().view(np.uint64)
    t0 = time.perf_counter()
    exp_c, _, _ = oracle.qdxt1(blocks, desc, 128, 4, oracle.QDXT_PERCEPTUAL | oracle.QDXT_HIERARCHICAL, 128, 0.75)
    exp_a, _, _ = oracle.qdxt5(blocks, desc, 128, 4, oracle.QDXT5_BOTH | oracle.QDXT_HIERARCHICAL, 3)
    cpu_s = time.perf_counter() - t0
    exact = bool(np.array_equal(got_c, exp_c) and np.array_equal(got_a, exp_a))
    assert exact, "GPU clustered DXT5 differs from the oracle"
    texels = 16 * n
    # one vector pass of a 16-D split reads the vector (64 B), its weight and its index
    k = 2
    bytes16 = passes[k] * (16 * 4 + 4 + 4)
    achieved = bytes16 / (sms[k] * 1e-3) / 1e9 if sms[k] else 0.0
    return {"workload": "clustered DXT5 (qdxt1 colour + qdxt5 alpha) of one %dx%d RGBA8 texture with its %d-level "
                        "mip chain (%d blocks), quality 128, uber, init + pack" % (size, size, len(desc) // 3, n),
            "mtexels_s": round(texels / (ms * 1e-3) / 1e6, 2), "ms": round(ms, 2), "bit_exact": exact,
            "how": "colour and alpha on two streams and host threads at once", "serial_ms": round(serial_ms, 2),
            "clusters": {"colour_endpoint": info[0]["endpoint_clusters"], "colour_selector":
                         info[0]["selector_clusters"], "alpha_endpoint": info[1]["endpoint_clusters"],
                         "alpha_selector": info[1]["selector_clusters"]},
            "tree_split_ms": {"2d": round(sms[0], 2), "6d": round(sms[1], 2), "16d": round(sms[2], 2)},
            "roofline": {"bound": "hbm", "kernel": "vq_split_kernel<16>", "achieved": round(achieved, 2),
                         "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6),
                         "traffic": None, "algorithmic_bytes_per_step": bytes16, "vector_passes": passes[k],
                         "launches": launches[k], "launch_ms_per_step": round(sms[k], 2),
                         "note": "the splits' float sums run in list order, one add per vector and component on a "
                                 "serial chain: latency-bound, not HBM-bound"},
            "cpu_baseline": {"value": round(texels / cpu_s / 1e6, 3), "unit": "Mtexels/s", "cores": 1,
                             "kind": "port", "sample": "the oracle's qdxt1 + qdxt5 of the same texture, one thread "
                                                       "(%.1fs)" % cpu_s}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch", type=int, default=1, help="texture pairs (distinct seeds)")
    ap.add_argument("--copies", type=int, default=128,
                    help="independent copies of every file per step (separate contexts and outputs), so that one "
                         "launch sequence moves hundreds of MB")
    ap.add_argument("--mixed", type=int, default=0,
                    help="instead of the 4096^2 pair, transcode a mixed set of this many textures (sides "
                         "log-uniform in [64, 2048], full mips; 60%% DXT1 RGB, 25%% DXT5 RGBA, 15%% DXN normal maps)")
    ap.add_argument("--mixed-set", type=int, default=512,
                    help="textures of the serving-set leg (encode + transcode); 0 skips it")
    ap.add_argument("--plain-size", type=int, default=256,
                    help="side of the RGB texture of the plain DXT1 leg; 0 skips it")
    ap.add_argument("--streams", type=int, default=4, help="plans (groups of copies) per step, one stream each")
    ap.add_argument("--writer", choices=("hc", "simple"), default="hc",
                    help="how the input .crn files are made: the GPU dxt_hc with the project's CRN writer, or the "
                         "simple CPU block quantizer")
    ap.add_argument("--size", type=int, default=4096)
    ap.add_argument("--cpu-seconds", type=float, default=10.0, help="CPU baseline sample length")
    ap.add_argument("--encode-size", type=int, default=8192,
                    help="side of the RGB texture of the DXT1 encode leg (learned vs plain); 0 skips it")
    ap.add_argument("--hc-size", type=int, default=2048,
                    help="side of the RGB texture of the dxt_hc leg; 0 skips it")
    ap.add_argument("--hc-cpu", type=int, default=1,
                    help="1: time the oracle's dxt_hc composition on the same texture (about 15 s); 0 skips it")
    ap.add_argument("--encode-cpu-seconds", type=float, default=5.0, help="CPU baseline sample of each encode leg")
    ap.add_argument("--mip-size", type=int, default=8192, help="side of the mip generation leg (0 = skip)")
    ap.add_argument("--qdxt-size", type=int, default=2048, help="side of the clustered DXT5 leg (0 = skip)")
    ap.add_argument("--fallback-frac", type=float, default=0.08,
                    help="fraction of blocks (lowest confidence first) that also run the plain search in the encode leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (each GPU would transcode its "
              "own textures independently); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from crunch2_amd import crnd, texture
    from crunch2_amd._lib import FMT_DXT1, FMT_DXT5, lib

    torch.cuda.set_device(0)
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()

    # ---- inputs: seeded textures compressed by the project's CRN writer
    t0 = time.time()
    files = []
    if args.mixed:
        from crunch2_amd._lib import FMT_DXN_XY
        files = [d for _, _, d in texture.mixed_set(args.mixed, writer=args.writer)]
        log("mixed set of %d textures built in %.1fs (%d bytes)" % (len(files), time.time() - t0,
                                                                    sum(len(f) for f in files)))
    else:
        for b in range(args.batch):
            for fmt in (FMT_DXT1, FMT_DXT5):
                pal = texture.Q128_DXT1 if fmt == FMT_DXT1 else texture.Q128_DXT5
                if args.writer == "hc":
                    pal = tuple(v if v else 8 for v in pal)
                files.append(texture.synth_crn(args.size, 1 + b, fmt, palette_sizes=pal, writer=args.writer))
        log("inputs built in %.1fs: %s bytes" % (time.time() - t0, [len(f) for f in files]))

    # crnd_unpack_begin of every file on the host (header, models, palettes):
    # the part of the example2 rate (begin + levels) outside the GPU pipeline
    b0 = time.perf_counter()
    for _ in range(3):
        for d in files:
            crnd.unpack_end(crnd.unpack_begin(d))
    begin_ms = (time.perf_counter() - b0) * 1e3 / 3
    ctxs, outs = [], []
    texels = 0         # of one copy of the files
    crn_bytes = 0
    out_bytes = 0
    for d in files:
        inf = crnd.get_texture_info(d)
        for l in range(inf["levels"]):
            li = crnd.get_level_info(d, l)
            texels += li["blocks_x"] * li["blocks_y"] * 16 * li["faces"]
        crn_bytes += len(d)
    for c in range(args.copies):
        for d in files:
            ctx = crnd.unpack_begin(d)
            assert ctx is not None
            inf = crnd.get_texture_info(d)
            n = crnd.levels_size(ctx, 0, inf["levels"])
            outs.append(torch.empty(n, dtype=torch.uint8, device=dev))
            ctxs.append(ctx)
            if c == 0:
                out_bytes += n

    # plans: the copies split into `streams` groups, one plan and one stream
    # each; a step runs every plan once, the plans overlapping on their streams
    # (one plan's serial walk beside another's bulk stages)
    nf = len(files)
    groups = max(1, min(args.streams, args.copies))
    per = [(args.copies * (g + 1)) // groups - (args.copies * g) // groups for g in range(groups)]
    plans, streams = [], []
    at = 0
    for g in range(groups):
        k = per[g] * nf
        pl = C.c_void_p()
        rc = lib().crn2_batch_create((C.c_void_p * k)(*[c.handle for c in ctxs[at:at + k]]), k,
                                     (C.c_void_p * k)(*[o.data_ptr() for o in outs[at:at + k]]),
                                     (C.c_uint64 * k)(*[o.numel() for o in outs[at:at + k]]), C.byref(pl))
        assert rc == 0, rc
        plans.append(pl)
        streams.append(stream if g == 0 else torch.cuda.Stream(device=dev))
        at += k

    def step():
        for pl, st in zip(plans, streams):
            assert lib().crn2_batch_run(pl, C.c_void_p(st.cuda_stream)) == 0

    # ---- parity spot check before timing (oracle restatement, first file)
    import oracle
    ref = oracle.unpack_all(files[0])
    step()
    torch.cuda.synchronize()
    exp = np.concatenate([np.concatenate([f.reshape(-1) for f in ref[l]]) for l in sorted(ref)])
    assert np.array_equal(outs[0].cpu().numpy(), exp), "GPU output differs from the oracle"
    assert np.array_equal(outs[len(outs) - nf].cpu().numpy(), exp), "last copy differs from the oracle"

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps

    # stage times: the first plan alone, profiled (stages one after another)
    lib().crn2_profile_reset()
    lib().crn2_profile_enable(1)
    for _ in range(3):
        assert lib().crn2_batch_run(plans[0], C.c_void_p(stream.cuda_stream)) == 0
    torch.cuda.synchronize()
    lib().crn2_profile_enable(0)
    prof = {}
    for k, name in enumerate(STAGES):
        t = C.c_double()
        n = C.c_uint64()
        lib().crn2_profile_read(k, C.byref(t), C.byref(n))
        prof[name] = t.value / max(1, n.value)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)

    # latency of one copy alone (a plan over the first copy's contexts)
    plan1 = C.c_void_p()
    rc = lib().crn2_batch_create((C.c_void_p * nf)(*[c.handle for c in ctxs[:nf]]), nf,
                                 (C.c_void_p * nf)(*[o.data_ptr() for o in outs[:nf]]),
                                 (C.c_uint64 * nf)(*[o.numel() for o in outs[:nf]]), C.byref(plan1))
    assert rc == 0, rc
    for _ in range(2):
        assert lib().crn2_batch_run(plan1, C.c_void_p(stream.cuda_stream)) == 0
    torch.cuda.synchronize()
    e0.record(stream)
    for _ in range(5):
        assert lib().crn2_batch_run(plan1, C.c_void_p(stream.cuda_stream)) == 0
    e1.record(stream)
    torch.cuda.synchronize()
    one_ms = e0.elapsed_time(e1) / 5
    lib().crn2_batch_destroy(plan1)

    mtex = args.copies * texels / (ms * 1e-3) / 1e6
    alg_bytes = args.copies * (crn_bytes + out_bytes)
    # algorithmic bytes of one step (all copies): compressed in + DXTn out
    step_gbs = alg_bytes / (ms * 1e-3) / 1e9
    # dominant kernel: the walk (the longest stage); one launch covers the
    # first plan's copies, timed alone with HIP events on its stream
    walk_ms = prof["crnd_walk_kernel"]
    plan_bytes = per[0] * (crn_bytes + out_bytes)
    achieved = plan_bytes / (walk_ms * 1e-3) / 1e9
    default_inputs = not args.mixed and args.size == 4096 and args.batch == 1 and args.writer == "hc"

    # ---- CPU baseline: the C restatement on the host cores, one context per
    # level of every file (the reference transcoder is single-threaded per
    # context; its multithreaded use is one context per level or texture), so
    # the levels of all files decode concurrently; contexts are opened (palettes
    # decoded) outside the timed region, as are the destination buffers.
    from concurrent.futures import ThreadPoolExecutor
    jobs = []
    for d in files:
        inf = crnd.get_texture_info(d)
        for l in range(inf["levels"]):
            li = crnd.get_level_info(d, l)
            c = oracle.Context(d)
            bpb = 8 if inf["format"] in (0, 9) else 16
            dst = [np.zeros(li["blocks_x"] * li["blocks_y"] * bpb, np.uint8) for _ in range(inf["faces"])]
            jobs.append((li["blocks_x"] * li["blocks_y"] * inf["faces"], c, l, dst))
    jobs.sort(key=lambda j: -j[0])
    threads = min(16, os.cpu_count() or 1, len(jobs))

    def run_job(j):
        assert j[1].unpack_level(j[2], dst=j[3]) is not False

    cpu_t = 0.0
    cpu_tex = 0
    reps = 0
    with ThreadPoolExecutor(max_workers=threads) as pool:
        while cpu_t < args.cpu_seconds:
            s0 = time.perf_counter()
            list(pool.map(run_job, jobs))
            cpu_t += time.perf_counter() - s0
            cpu_tex += texels
            reps += 1
    cpu_mtex = cpu_tex / cpu_t / 1e6
    s0 = time.perf_counter()
    for j in jobs:
        run_job(j)
    cpu1_mtex = texels / (time.perf_counter() - s0) / 1e6
    for j in jobs:
        j[1].close()

    line = {
        "metric": "Mtexels/s: CRN->DXTn transcode (DXT1 + DXT5, 4096^2, 13 levels)",
        "value": round(mtex, 1),
        "unit": "Mtexels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": ("synthetic (seeded %s, full box-filter mip chains, %s)"
                 % ("textures of the mixed set" if args.mixed else "4096^2 RGB and RGBA8 textures",
                    "compressed by the GPU dxt_hc at crnlib's quality-128 palette sizes and packed by the "
                    "project's writer with crnlib's palette ordering" if args.writer == "hc" else
                    "packed to .crn by the project's simple block quantizer at crnlib's quality-128 palette sizes")),
        "config": {"workload": ("CRN->DXTn transcode, mixed set of %d textures (64^2..2048^2, full mips, "
                                "DXT1/DXT5/DXN_XY)" % args.mixed) if args.mixed else
                               ("CRN->DXTn transcode, %dx%d, 13 levels, DXT1 + DXT5, %d texture pair(s) x %d "
                                "copies per step" % (args.size, args.size, args.batch, args.copies)),
                   "copies": args.copies, "streams": groups, "one_copy_ms": round(one_ms, 3),
                   "one_copy_mtexels_s": round(texels / (one_ms * 1e-3) / 1e6, 1),
                   "begin_ms_per_copy": round(begin_ms, 3),
                   "one_copy_with_begin_mtexels_s": round(texels / ((one_ms + begin_ms) * 1e-3) / 1e6, 1),
                   "step_with_begin_mtexels_s": round(args.copies * texels /
                                                      ((ms + args.copies * begin_ms) * 1e-3) / 1e6, 1),
                   "crn_bytes": args.copies * crn_bytes,
                   "dxtn_bytes": args.copies * out_bytes, "texels": args.copies * texels},
        "roofline": {"bound": "hbm", "kernel": "crnd_walk_kernel (one launch per plan of %d copies)" % per[0],
                     "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 6),
                     "traffic": (WALK_TRAFFIC_PER_COPY * per[0]) if (default_inputs and WALK_TRAFFIC_PER_COPY)
                     else None,
                     "algorithmic_bytes_per_launch": plan_bytes, "launch_ms": round(walk_ms, 4),
                     "step": {"achieved": round(step_gbs, 2), "frac": round(step_gbs / HBM_PEAK_GBS, 6),
                              "what": "algorithmic bytes of the whole step / step time (%d plans of %s copies on "
                                      "%d streams)" % (groups, "/".join(str(x) for x in per), groups)},
                     "stage_ms_one_plan": {k: round(v, 4) for k, v in prof.items()}},
        "cpu_baseline": {"value": round(cpu_mtex, 1), "unit": "Mtexels/s", "cores": threads, "kind": "port",
                         "sample": "%d full passes over the same %d files, %d level jobs on %d threads (%.1fs of "
                                   "wall time); one thread alone: %.1f Mtexels/s"
                                   % (reps, len(files), len(jobs), threads, cpu_t, cpu1_mtex)},
    }
    for pl in plans:
        assert lib().crn2_batch_status(pl, C.c_void_p(stream.cuda_stream)) == 0
    if args.plain_size:
        line["plain_dxt1"] = plain_leg(args.plain_size, args.encode_cpu_seconds)
        log("plain DXT1 leg done")
    if args.encode_size:
        line["encode"] = encode_leg(args.encode_size, args.fallback_frac, args.encode_cpu_seconds)
        log("learned encode leg done")
    if args.hc_size:
        line["dxt_hc"] = hc_leg(args.hc_size, args.hc_cpu)
        log("dxt_hc leg done")
    if args.mixed_set:
        line["mixed_set"] = mixed_leg(args.mixed_set, args.encode_cpu_seconds)
        log("serving-set leg done")
    if args.mip_size:
        line["mips"] = mip_leg(args.mip_size, args.encode_cpu_seconds)
    if args.qdxt_size:
        line["qdxt"] = qdxt_leg(args.qdxt_size)
        log("mip generation leg done")
    print(json.dumps(line))
    for pl in plans:
        lib().crn2_batch_destroy(pl)
    for c in ctxs:
        crnd.unpack_end(c)


if __name__ == "__main__":
    main()
