"""GPU clustered DXT1 (crn2_qdxt1_*) and its vector quantizers (crn2_vq_*)
against the oracle's restatement of crnlib's qdxt1 / clusterizer /
threaded_clusterizer, bit for bit: the clusters of random 2-, 6- and 16-D
vectors at several leaf budgets and cuts, the threaded selector clusters,
the training vectors (hierarchical tiles over mip chains, odd level sizes,
and per block), and whole packs at several qualities, codebook-curve factors
and textures, with and without 3-colour blocks, up to a 512 x 512 mip chain;
and qdxt5 the same way."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle
from crunch2_amd import qdxt, texture
from crunch2_amd._lib import CRN2_EUNSUPPORTED, lib
from tests.qdxt_cases import level_blocks, texture_blocks

pytestmark = pytest.mark.gpu

P, A, H = qdxt.PERCEPTUAL, qdxt.ALPHA_BLOCKS, qdxt.HIERARCHICAL


def _same_clusters(got, exp, what):
    assert len(got) == len(exp), (what, len(got), len(exp))
    for k, (a, b) in enumerate(zip(got, exp)):
        assert np.array_equal(a, b), (what, k, a[:8], b[:8])


def _vectors(d, n, seed):
    rng = np.random.default_rng(seed)
    if d == 16:
        v = rng.integers(0, 4, (n, d)).astype(np.float32)
        v[: n // 3, :6] = 0
        w = rng.integers(1, 2049, n).astype(np.uint32)
    else:
        centres = rng.integers(0, 256, (24, d))
        v = np.clip(centres[rng.integers(0, 24, n)] + rng.normal(0, 7, (n, d)), 0, 255).round().astype(np.float32)
        v[: n // 10] = v[0]
        w = rng.integers(1, 9, n).astype(np.uint32)
    return v, w


@pytest.mark.parametrize("d,n,max_size", [(6, 5000, 1000), (6, 3000, 65535), (16, 4000, 700), (2, 2000, 300),
                                          (6, 1, 10), (6, 2, 10),
                                          (16, 100000, 3000)])   # large 16-D nodes: in-order sums, three waves
def test_vq_clusters_match_oracle(gpu, d, n, max_size):
    v, w = _vectors(d, n, d * 7 + n)
    q = oracle.VQ(v, w, max_size)
    for cut in (0, 1, 7, 96, max_size // 3):
        got, cbs = qdxt.vq_clusters(v, w, max_size, cut)
        assert cbs == q.codebook_size
        _same_clusters(got, q.retrieve(cut if cut else q.codebook_size), (d, n, max_size, cut))
    q.close()


@pytest.mark.parametrize("d,n,maxc", [(16, 6000, 500), (16, 3000, 100), (6, 2500, 300), (16, 100000, 1000)])
def test_vq_threaded_clusters_match_oracle(gpu, d, n, maxc):
    v, w = _vectors(d, n, 3 * n + maxc)
    _same_clusters(qdxt.vq_threaded_clusters(v, w, maxc), oracle.vq_threaded_clusters(v, w, maxc), (d, n, maxc))


@pytest.mark.parametrize("size,height,kind,hier", [(128, 128, texture.RGB, True), (72, 40, texture.RGBA, True),
                                                   (64, 64, texture.NORMAL, False)])
def test_training_vectors_match_oracle(gpu, size, height, kind, hier):
    blocks, desc = level_blocks(texture.mip_chain(texture.synth(size, height, 5, kind)))
    ev, ew, ee = oracle.qdxt1_train(blocks, desc, hier)
    q = qdxt.Qdxt1(blocks, desc, hier)
    gv, gw, ge = q.training()
    assert np.array_equal(gv, ev) and np.array_equal(gw, ew) and np.array_equal(ge, ee)
    assert q.info()["max_selector_clusters"] == oracle.qdxt1_distinct_selectors(blocks) + 128
    q.close()


CASES = [
    (128, 11, texture.RGB, 128, P | H, 1.0),
    (128, 12, texture.RGB, 40, P | H, 1.0),
    (128, 13, texture.RGB, 220, H, 1.0),
    (96, 14, texture.NORMAL, 128, H, 1.5),
    (128, 15, texture.RGBA, 90, P | H, 0.75),
    (128, 16, texture.RGB, 128, P, 1.0),
    (128, 17, texture.RGBA, 255, P | A | H, 1.0),
    (64, 18, texture.RGB, 255, H, 1.0),
    (128, 19, texture.RGB, 128, P | A | H, 1.0),
    (128, 20, texture.RGB, 60, A | H, 1.0),
    (96, 22, texture.NORMAL, 200, P | A, 1.0),
    (128, 23, texture.RGBA, 128, P | A | H, 1.0),
    (96, 24, texture.RGBA, 50, A | H, 1.0),
]


@pytest.mark.parametrize("case", CASES, ids=["%d-s%d-k%d-q%d-f%d-m%g" % c for c in CASES])
def test_pack_matches_oracle(gpu, case):
    size, seed, kind, ql, flags, mul = case
    blocks, desc = texture_blocks(size, seed, kind)
    exp, _, st = oracle.qdxt1(blocks, desc, ql, 4, flags, 128, mul)
    q = qdxt.Qdxt1(blocks, desc, bool(flags & H))
    got = q.pack(qdxt.params(ql, 4, flags, 128), mul).cpu().numpy().view(np.uint64)
    info = q.info()
    assert info["codebook_size"] == st["codebook_size"]
    assert info["max_selector_clusters"] == st["max_selector_clusters"]
    assert info["endpoint_clusters"] == st["endpoint_clusters"]
    if ql < 255:
        assert info["selector_clusters"] == st["selector_clusters"]
    bad = np.nonzero(got != exp)[0]
    assert not len(bad), (len(bad), bad[:5], got[bad[:3]], exp[bad[:3]])
    q.close()


def test_pack_512_chain_and_repacks(gpu):
    blocks, desc = texture_blocks(512, 21)
    q = qdxt.Qdxt1(blocks, desc)
    for ql in (128, 64):
        exp, _, st = oracle.qdxt1(blocks, desc, ql, 4, P | H, 128, 1.0)
        got = q.pack(qdxt.params(ql, 4, P | H)).cpu().numpy().view(np.uint64)
        assert np.array_equal(got, exp), ql
    q.close()


def test_unsupported_configurations(gpu):
    # qualities below uber under 255, superfast
    blocks, desc = texture_blocks(32, 3, texture.RGBA)
    q = qdxt.Qdxt1(blocks, desc)
    out = torch.zeros(len(blocks), dtype=torch.int64, device="cuda")
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    for p in (qdxt.params(128, 3, P | H), qdxt.params(255, 0, P | H)):
        assert lib().crn2_qdxt1_pack(q.h, C.byref(p), 1.0, C.c_void_p(out.data_ptr()), s) == CRN2_EUNSUPPORTED
    q.close()


def test_transparent_clusters(gpu):
    # DXT1A inside multi-block clusters: fully transparent blocks, clusters
    # with one opaque colour, and mixed ones (crn_qdxt1.cpp:590-626)
    img = texture.synth(64, 64, 9, texture.RGBA)
    img[:16, :32, 3] = 0                       # transparent area
    img[16:32, :32, :3] = (40, 90, 200)        # one colour ...
    img[16:32, :16, 3] = 10                    # ... partly transparent
    blocks, desc = level_blocks(texture.mip_chain(img))
    for ql, flags in ((100, P | A | H), (30, A | H), (200, P | A)):
        exp, _, _ = oracle.qdxt1(blocks, desc, ql, 4, flags, 128, 1.0)
        q = qdxt.Qdxt1(blocks, desc, bool(flags & H))
        got = q.pack(qdxt.params(ql, 4, flags, 128)).cpu().numpy().view(np.uint64)
        q.close()
        assert np.array_equal(got, exp), (ql, flags, np.nonzero(got != exp)[0][:5])


# ------------------------------------------------------------------ qdxt5
@pytest.mark.parametrize("size,kind,comp,hier", [(96, texture.RGBA, 3, True), (64, texture.NORMAL, 1, False)])
def test_qdxt5_training_matches_oracle(gpu, size, kind, comp, hier):
    blocks, desc = texture_blocks(size, 8, kind)
    ev, ew, ee = oracle.qdxt5_train(blocks, desc, comp, hier)
    q = qdxt.Qdxt5(blocks, desc, comp, hier)
    gv, gw, ge = q.training()
    assert np.array_equal(gv, ev) and np.array_equal(gw, ew) and np.array_equal(ge, ee)
    assert q.info()["max_selector_clusters"] == oracle.qdxt5_distinct_selectors(blocks, comp) + 128
    q.close()


CASES5 = [
    (128, 31, texture.RGBA, 3, 128, 4, 1),
    (128, 32, texture.RGBA, 3, 60, 4, 0),
    (128, 33, texture.NORMAL, 0, 200, 3, 1),
    (96, 34, texture.NORMAL, 1, 128, 2, 1),
    (64, 35, texture.RGBA, 3, 255, 4, 1),
    (96, 36, texture.RGBA, 3, 128, 0, 1),       # superfast: dxt_fast's alpha compressor per cluster
    (64, 37, texture.NORMAL, 1, 255, 0, 0),
]


@pytest.mark.parametrize("case", CASES5, ids=["%d-s%d-k%d-c%d-q%d-d%d-b%d" % c for c in CASES5])
def test_qdxt5_pack_matches_oracle(gpu, case):
    size, seed, kind, comp, ql, dq, both = case
    blocks, desc = texture_blocks(size, seed, kind)
    exp, _, st = oracle.qdxt5(blocks, desc, ql, dq, both | oracle.QDXT_HIERARCHICAL, comp)
    q = qdxt.Qdxt5(blocks, desc, comp)
    got = q.pack(qdxt.params5(ql, dq, both)).cpu().numpy().view(np.uint64)
    info = q.info()
    assert info["codebook_size"] == st["codebook_size"]
    assert info["endpoint_clusters"] == st["endpoint_clusters"]
    if ql < 255:
        assert info["selector_clusters"] == st["selector_clusters"]
    bad = np.nonzero(got != exp)[0]
    assert not len(bad), (len(bad), bad[:5], got[bad[:3]], exp[bad[:3]])
    q.close()


def test_colour_and_alpha_on_two_streams_at_once(gpu):
    # handles on different streams run at once (stream-ordered device
    # buffers, one host thread each) and give what the oracle gives
    import threading
    blocks, desc = texture_blocks(128, 41, texture.RGBA)
    b = torch.from_numpy(blocks).cuda()
    torch.cuda.synchronize()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    res = {}

    def colour():
        try:
            with torch.cuda.stream(s1):
                q = qdxt.Qdxt1(b, desc)
                res["c"] = q.pack(qdxt.params(100, 4, P | H), 0.75)
                q.close()
                s1.synchronize()
        except BaseException as e:
            res["e"] = e

    for _ in range(2):
        t = threading.Thread(target=colour)
        t.start()
        with torch.cuda.stream(s2):
            q5 = qdxt.Qdxt5(b, desc, 3)
            a = q5.pack(qdxt.params5(100, 4, qdxt.BOTH_BLOCK_TYPES))
            q5.close()
            s2.synchronize()
        t.join()
        assert "e" not in res, res.get("e")
        exp_c, _, _ = oracle.qdxt1(blocks, desc, 100, 4, P | H, 128, 0.75)
        exp_a, _, _ = oracle.qdxt5(blocks, desc, 100, 4, oracle.QDXT5_BOTH | H, 3)
        assert np.array_equal(res["c"].cpu().numpy().view(np.uint64), exp_c)
        assert np.array_equal(a.cpu().numpy().view(np.uint64), exp_a)
