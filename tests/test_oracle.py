"""Pins the CPU oracle (oracle/crnd_oracle.c) against checks that do not come
from the oracle itself: catalogue CRC values, a textbook canonical-code
construction, an independent model-transmission encoder written here, and
known-answer .crn files whose expected DXTn bytes are computed from the
palettes (tests/crn_cases.py)."""
import heapq

import numpy as np
import pytest

import oracle
from crunch2_amd import texture
from crunch2_amd._lib import lib, FMT_DXT1
from tests import crn_cases


# ------------------------------------------------------------------ crc16
def crc16_bitwise(data):
    """CRC-16/GENIBUS (poly 0x1021, init 0xFFFF, xorout 0xFFFF, MSB first)."""
    c = 0xFFFF
    for byte in data:
        c ^= byte << 8
        for _ in range(8):
            c = ((c << 1) ^ 0x1021) if c & 0x8000 else (c << 1)
            c &= 0xFFFF
    return c ^ 0xFFFF


def test_crc16_catalogue_check_value():
    assert oracle.crc16(b"123456789") == 0xD64E
    assert crc16_bitwise(b"123456789") == 0xD64E


def test_crc16_matches_bitwise_and_writer_form():
    import ctypes as C
    rng = np.random.default_rng(0)
    for n in (0, 1, 2, 7, 64, 1000):
        d = rng.integers(0, 256, n).astype(np.uint8).tobytes()
        ref = crc16_bitwise(d)
        assert oracle.crc16(d) == ref
        # the writer's table-less byte form (crn_decomp.h:2048-2069 shape)
        assert lib().crn2_crc16(C.c_char_p(d), n, 0) == ref


# ------------------------------------------------------- canonical codes
def huffman_lengths(freq, limit=16):
    """Plain Huffman code lengths (heap), rejecting results over the limit."""
    items = [(f, i) for i, f in enumerate(freq) if f > 0]
    if len(items) == 1:
        out = [0] * len(freq)
        out[items[0][1]] = 1
        return out
    heap = [(f, k, [i]) for k, (f, i) in enumerate(items)]
    heapq.heapify(heap)
    depth = [0] * len(freq)
    k = len(heap)
    while len(heap) > 1:
        f1, _, a = heapq.heappop(heap)
        f2, _, b = heapq.heappop(heap)
        for i in a + b:
            depth[i] += 1
        heapq.heappush(heap, (f1 + f2, k, a + b))
        k += 1
    assert max(depth) <= limit
    return depth


def rfc1951_codes(lengths):
    """Canonical codes per RFC 1951 3.2.2: shorter codes first, equal lengths
    in increasing symbol order."""
    maxlen = max(lengths)
    bl_count = [0] * (maxlen + 1)
    for L in lengths:
        if L:
            bl_count[L] += 1
    code, next_code = 0, [0] * (maxlen + 2)
    for bits in range(1, maxlen + 1):
        code = (code + bl_count[bits - 1]) << 1 if bits > 1 else 0
        next_code[bits] = code
    codes = {}
    for s, L in enumerate(lengths):
        if L:
            codes[s] = (next_code[L], L)
            next_code[L] += 1
    return codes


def pack_bits(pairs):
    bits = []
    for v, n in pairs:
        bits.extend((v >> (n - 1 - i)) & 1 for i in range(n))
    while len(bits) % 8:
        bits.append(0)
    return np.packbits(np.array(bits, np.uint8)).tobytes()


@pytest.mark.parametrize("nsyms,seed", [(2, 0), (21, 1), (300, 2), (8192, 3)])
def test_canonical_decode_matches_textbook(nsyms, seed):
    rng = np.random.default_rng(seed)
    freq = ((rng.pareto(2.5, nsyms) * 20).astype(np.int64) + 1) * (rng.random(nsyms) < 0.7)
    freq[0] = 1
    lengths = huffman_lengths(list(freq))
    codes = rfc1951_codes(lengths)
    used = [s for s in range(nsyms) if lengths[s]]
    msg = list(rng.choice(used, 4000))
    data = pack_bits([codes[s] for s in msg])
    got = oracle.decode_symbols(lengths, data, len(msg))
    assert list(got) == msg


def test_no_matching_code_returns_zero_without_consuming():
    # one code of length 1 ("0"); a "1" bit matches nothing: the reference
    # returns symbol 0 and consumes no bits (crn_decomp.h:3239-3244)
    lengths = [0, 0, 1]
    data = bytes([0b01111111, 0xFF])
    got = oracle.decode_symbols(lengths, data, 5)
    assert list(got) == [2, 0, 0, 0, 0]


# --------------------------------------------------- model transmission
CL_ORDER = [17, 18, 19, 20, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15, 16]


def encode_model(sizes):
    """Independent encoder for the code-length transmission of
    crn_decomp.h:3004-3123: 14-bit used count, 5-bit count of code-length
    code sizes sent in CL_ORDER, 3 bits each, then literals 0..16 and runs
    17 (3..10 zeros, +3 bits), 18 (11..138 zeros, +7), 19 (3..6 repeats, +2),
    20 (7..70 repeats, +6)."""
    used = max([i + 1 for i, s in enumerate(sizes) if s] + [0])
    out = [(used, 14)]
    if not used:
        return pack_bits(out)
    syms = []
    i = 0
    while i < used:
        v = sizes[i]
        r = 1
        while i + r < used and sizes[i + r] == v:
            r += 1
        left = r
        if v == 0:
            while left >= 11:
                k = min(left, 138)
                syms.append((18, k - 11, 7))
                left -= k
            if left >= 3:
                syms.append((17, left - 3, 3))
                left = 0
            syms.extend([(0, 0, 0)] * left)
        else:
            syms.append((v, 0, 0))
            left -= 1
            while left >= 7:
                k = min(left, 70)
                syms.append((20, k - 7, 6))
                left -= k
            if left >= 3:
                syms.append((19, left - 3, 2))
                left = 0
            syms.extend([(v, 0, 0)] * left)
        i += r
    freq = [0] * 21
    for s, _, _ in syms:
        freq[s] += 1
    cl = huffman_lengths(freq, limit=7)
    codes = rfc1951_codes(cl)
    nsend = max(k + 1 for k in range(21) if cl[CL_ORDER[k]])
    out.append((nsend, 5))
    out += [(cl[CL_ORDER[k]], 3) for k in range(nsend)]
    for s, extra, nb in syms:
        out.append(codes[s])
        if nb:
            out.append((extra, nb))
    return pack_bits(out)


@pytest.mark.parametrize("seed", range(6))
def test_model_transmission(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.choice([3, 40, 257, 2000, 8192]))
    sizes = np.zeros(n, np.uint8)
    pos = 0
    while pos < n:  # runs of zeros and of equal lengths, plus singles
        r = int(rng.integers(1, 160))
        v = int(rng.integers(0, 17)) if rng.random() < 0.7 else 0
        sizes[pos:pos + r] = v
        pos += r
    sizes[int(rng.integers(0, n))] = 5
    got = oracle.receive_model(encode_model(list(sizes)))
    last = max(i + 1 for i in range(n) if sizes[i])
    assert got is not None and list(got) == list(sizes[:last])


def test_model_transmission_empty():
    assert list(oracle.receive_model(encode_model([0, 0, 0]))) == []


# ---------------------------------------------------- known-answer files
@pytest.mark.parametrize("i", range(len(crn_cases.CASES)))
def test_oracle_known_answer(i):
    data, expected = crn_cases.case(i)
    assert oracle.validate_file(data)
    got = oracle.unpack_all(data)
    assert got is not None
    for l, faces in expected.items():
        for f, exp in enumerate(faces):
            assert np.array_equal(got[l][f], exp), (i, l, f)


def test_selector_map_is_dxt_code_table():
    # palette entry with linear selectors 0,1,2,3 repeated -> DXT1 codes 0,2,3,1
    lin = sum(((i % 4) << (2 * i)) for i in range(16))
    data = texture.pack(4, 4, 1, 1, FMT_DXT1, color_endpoints=[0x1234ABCD], color_selectors=[lin],
                        ce_idx=np.zeros(4, np.uint32), cs_idx=np.zeros(4, np.uint32))
    with oracle.Context(data) as c:
        raw = int(c.palette(1)[0])
    assert [(raw >> (2 * i)) & 3 for i in range(4)] == [0, 2, 3, 1]


# --------------------------------------------------------- error paths
def test_error_paths():
    data, _ = crn_cases.case(2)  # 64x64 DXT1, 7 levels
    with oracle.Context(data) as c:
        assert c.h
        assert c.unpack_level(0, row_pitch=130) is False            # not a multiple of 4
        assert c.unpack_level(0, row_pitch=64) is False             # below 16 blocks * 8 bytes
        assert c.unpack_level(0, dst=[np.zeros(64, np.uint8)], dst_size=64) is False  # too small
        assert c.unpack_level(7) is False                           # no such level
        assert c.unpack_level(0, row_pitch=256) is not False        # padded pitch is fine
    assert oracle.Context(data[:61]).h is None                      # shorter than the minimum header
    bad = bytearray(data)
    bad[0] = ord("X")
    assert oracle.Context(bytes(bad)).h is None                     # signature
    flipped = bytearray(data)
    flipped[len(data) // 2] ^= 0x10
    assert not oracle.validate_file(bytes(flipped))                 # data CRC


def test_padded_pitch_keeps_gap_bytes():
    data, expected = crn_cases.case(1)  # 13x7: 4x2 blocks
    with oracle.Context(data) as c:
        dst = [np.full(2 * 48, 0xEE, np.uint8)]
        r = c.unpack_level(0, row_pitch=48, dst=dst)
    rows = r[0].reshape(2, 48)
    assert np.array_equal(rows[:, :32].reshape(-1), expected[0][0])
    assert (rows[:, 32:] == 0xEE).all()


@pytest.mark.parametrize("factor", [0.5, 0.9, 0.99])
def test_selector_range_check_matches_trace(factor):
    """With the colour-selector palette cut to m entries, a level fails exactly
    when one of its blocks uses a selector index >= m; the indices come from
    the trace of the intact file, the failures from the cut one."""
    from crunch2_amd._lib import FMT_DXT1 as fmt
    data = texture.synth_crn((256, 256), 31, fmt)
    d = crn_cases.shrink_palette(data, 1, factor, oracle.crc16)
    m = (d[47] << 8) | d[48]
    assert m < ((data[47] << 8) | data[48])
    levels = oracle.info(data)["levels"]
    with oracle.Context(data) as c:
        used = [int(c.trace_level(l)["cs"].max()) for l in range(levels)]
    with oracle.Context(d) as c:
        assert c.h
        got = [c.unpack_level(l) is not False for l in range(levels)]
    assert got == [u < m for u in used]
    assert not all(got)
