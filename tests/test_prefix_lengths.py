"""The transcoder's per-prefix code lengths (crnd_host.cpp: prefix_lengths),
pinned on the CPU against canonical decoding done from scratch: canonical
codes assigned by length then symbol, each length's codes one interval of
16-bit windows.  Every prefix that names a length must start that length in
all 32 of its windows (0 where no code matches), only prefixes whose windows
differ may be left to the canonical walk (0xFF), and there are at most five
of those (one per length boundary past the 11-bit direct table)."""
import numpy as np
import pytest

from crunch2_amd._lib import lib


def window_lengths(sizes):
    """Length of the code each 16-bit window starts (0: none), canonical order."""
    sizes = np.asarray(sizes)
    count = np.bincount(sizes[sizes > 0], minlength=17)
    out = np.zeros(65536, np.int64)
    code = 0
    for L in range(1, 17):
        lo, hi = code << (16 - L), (code + int(count[L])) << (16 - L)
        out[lo:min(hi, 65536)] = L
        code = (code + int(count[L])) << 1
    return out


def prefix_lengths(sizes):
    s = np.ascontiguousarray(sizes, np.uint8)
    out = np.zeros(2048, np.uint8)
    assert lib().crn2_prefix_lengths(s.ctypes.data, s.size, out.ctypes.data) == 0
    return out


def kraft_fit(rng, n, max_len, unused=0.0):
    """Random code sizes with the Kraft sum at most 1 (lengthening at random)."""
    L = rng.integers(1, max_len + 1, n)
    L[rng.random(n) < unused] = 0
    while (2.0 ** -L[L > 0]).sum() > 1:
        i = rng.choice(np.flatnonzero((L > 0) & (L < 16)))
        L[i] += 1
    return L


def huffman_sizes(rng, n, skew):
    """Code sizes of a length-limited (16) Huffman code of random frequencies."""
    f = rng.pareto(skew, n) + 1e-3
    import heapq
    heap = [(w, i, [i]) for i, w in enumerate(f)]
    heapq.heapify(heap)
    depth = np.zeros(n, np.int64)
    k = n
    while len(heap) > 1:
        a, _, sa = heapq.heappop(heap)
        b, _, sb = heapq.heappop(heap)
        for i in sa + sb:
            depth[i] += 1
        heapq.heappush(heap, (a + b, k, sa + sb))
        k += 1
    while depth.max() > 16:   # fold over-long codes back (keeps the Kraft sum <= 1)
        depth = np.minimum(depth, 16)
        while (2.0 ** -depth).sum() > 1:
            depth[np.argmin(np.where(depth < 16, depth, 99))] += 1
    return np.maximum(depth, 1)


def check(sizes):
    wl = window_lengths(sizes)
    got = prefix_lengths(sizes).astype(np.int64)
    per = wl.reshape(2048, 32)
    uniform = (per == per[:, :1]).all(axis=1)
    named = got != 0xFF
    assert np.array_equal(got[named], per[named, 0])
    assert not np.any(named & ~uniform) and np.all(named | ~uniform)
    assert (~named).sum() <= 5
    return (~named).sum()


@pytest.mark.parametrize("seed", range(6))
def test_huffman_codes_of_large_alphabets(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1500, 8192))
    check(huffman_sizes(rng, n, 0.6 + 0.4 * seed))


@pytest.mark.parametrize("seed", range(6))
def test_random_incomplete_codes(seed):
    rng = np.random.default_rng(100 + seed)
    check(kraft_fit(rng, int(rng.integers(2, 3000)), 16, unused=0.2))


def test_short_codes_only_and_one_symbol():
    assert check([1, 2, 3, 3]) == 0
    assert check([1]) == 0
    assert check([5] * 32) == 0


def test_boundaries_off_prefix_alignment():
    # after the 1-bit code (windows 0..32767) the three 12-bit codes take
    # windows 32768..32815: prefix 1024 is all 12 bits, prefix 1025 holds the
    # 12 -> 13 boundary, prefix 1026 the 13 -> 14 -> 15 -> 16 -> no-code ones
    # (windows 32840, 32852, 32858, 32861): two prefixes for the walk
    sizes = [1] + [12] * 3 + [13] * 3 + [14] * 3 + [15] * 3 + [16] * 3
    got = prefix_lengths(sizes)
    assert check(sizes) == 2
    assert got[1024] == 12 and got[1025] == 0xFF and got[1026] == 0xFF and got[1027] == 0
