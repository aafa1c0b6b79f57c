"""GPU DXTn decode (crn2_decode_dxt_device, crn2_decompress_blocks and the
crn_decompress_block facade) and its error sums (crn2_dxt_error_device)
against the oracle's restatement of crn_decompress_block, every format,
random blocks covering both colour modes and both alpha modes, odd image
sizes and row pitches."""
import ctypes as C

import numpy as np
import pytest
import torch

from crunch2_amd._lib import lib
from oracle import dxt_decode as dd

pytestmark = pytest.mark.gpu
FORMATS = list(range(10))


def random_blocks(fmt, n, seed):
    rng = np.random.default_rng(seed)
    b = rng.integers(0, 256, (n, dd.bytes_per_block(fmt)), dtype=np.uint8)
    # colour words in both orders (four- and three-colour blocks); random alpha
    # bytes cover l > h and l <= h
    cols = [8] if fmt in (1, 2, 3, 4, 5, 6) else ([0] if fmt == 0 else [])
    for c in cols:
        b[: n // 2, c:c + 2], b[: n // 2, c + 2:c + 4] = b[: n // 2, c + 2:c + 4].copy(), b[: n // 2, c:c + 2].copy()
        b[:4, c + 2:c + 4] = b[:4, c:c + 2]           # equal endpoints: the three-colour / six-value mode
    return b


@pytest.mark.parametrize("fmt", FORMATS)
def test_decode_blocks_match_oracle(gpu, fmt):
    n = 4096
    b = random_blocks(fmt, n, 100 + fmt)
    exp = dd.decode_blocks(b.tobytes(), fmt)
    d_b = torch.from_numpy(b.reshape(-1)).cuda()
    out = torch.zeros(n * 64, dtype=torch.uint8, device="cuda")
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib().crn2_decode_dxt_device(C.c_void_p(d_b.data_ptr()), fmt, 4, 4 * n, 0, C.c_void_p(out.data_ptr()), 0,
                                        s) == 0
    torch.cuda.synchronize()
    assert np.array_equal(out.cpu().numpy().reshape(n, 16, 4), exp)
    # host form
    host = np.zeros((n, 16), np.uint32)
    assert lib().crn2_decompress_blocks(b.ctypes.data_as(C.c_void_p), n, fmt, host.ctypes.data_as(C.c_void_p)) == 0
    assert np.array_equal(host.view(np.uint8).reshape(n, 16, 4), exp)


@pytest.mark.parametrize("fmt,w,h,pitch_pad", [(0, 13, 7, 0), (2, 61, 37, 16), (7, 8, 8, 0), (9, 5, 3, 8),
                                               (0, 256, 256, 0)])
def test_decode_image_with_pitch(gpu, fmt, w, h, pitch_pad):
    bpb = dd.bytes_per_block(fmt)
    bx, by = (w + 3) // 4, (h + 3) // 4
    pitch = bx * bpb + pitch_pad
    rng = np.random.default_rng(w * h)
    raw = rng.integers(0, 256, (by, pitch), dtype=np.uint8)
    blocks = raw[:, :bx * bpb].reshape(-1)
    exp = dd.decode_image(blocks.tobytes(), fmt, w, h)
    d_raw = torch.from_numpy(raw.reshape(-1)).cuda()
    rp = w * 4 + 12
    out = torch.full((h * rp,), 7, dtype=torch.uint8, device="cuda")
    s = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    assert lib().crn2_decode_dxt_device(C.c_void_p(d_raw.data_ptr()), fmt, w, h, pitch, C.c_void_p(out.data_ptr()),
                                        rp, s) == 0
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(h, rp)
    assert np.array_equal(got[:, :w * 4].reshape(h, w, 4), exp)
    assert (got[:, w * 4:] == 7).all()                # nothing past the image edge
    # error sums against a random source image
    src = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    d_src = torch.from_numpy(src.reshape(-1)).cuda()
    sse = torch.zeros(4, dtype=torch.int64, device="cuda")
    assert lib().crn2_dxt_error_device(C.c_void_p(d_raw.data_ptr()), fmt, w, h, pitch, C.c_void_p(d_src.data_ptr()),
                                       0, C.c_void_p(sse.data_ptr()), s) == 0
    torch.cuda.synchronize()
    e = ((exp.astype(np.int64) - src.astype(np.int64)) ** 2).reshape(-1, 4).sum(0)
    assert sse.cpu().numpy().tolist() == e.tolist()


@pytest.mark.parametrize("fmt", FORMATS)
def test_crn_decompress_block_facade(gpu, fmt):
    f = getattr(lib(), "_Z20crn_decompress_blockPKvPj10crn_format")
    f.restype, f.argtypes = C.c_bool, [C.c_void_p, C.c_void_p, C.c_int64]
    b = random_blocks(fmt, 8, 7 + fmt)
    for i in range(8):
        out = np.zeros(16, np.uint32)
        assert f(b[i].ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), fmt)
        assert np.array_equal(out.view(np.uint8).reshape(16, 4), dd.decode_blocks(b[i].tobytes(), fmt)[0])
    out = np.zeros(16, np.uint32)
    assert not f(b[0].ctypes.data_as(C.c_void_p), out.ctypes.data_as(C.c_void_p), 10)   # ETC1: not built

