// crn_pack.cpp -- host-side CRN writer and seeded synthetic textures.
//
// The writer produces .crn files in the format the transcoder consumes
// (reference format: inc/crn_defs.h:227-341, decoder inc/crn_decomp.h).  It is
// the encode direction of the format, written for this project:
//   * prefix codes: length-limited Huffman lengths (limit 16, palettes 15, as
//     crnlib/crn_comp.cpp:100,1573-1584), canonical codes assigned in
//     increasing symbol order within a length (crn_decomp.h:2150-2235);
//   * model transmission with the 21-symbol code-length alphabet and its
//     run codes 17..20 (crn_decomp.h:3004-3123);
//   * palettes: colour endpoints as six running deltas mod 32/64
//     (crn_comp.cpp:60-118), colour selectors as XOR-chained nibbles of the
//     linear selector word (crn_comp.cpp:238-266), alpha endpoints as two
//     running deltas mod 256 (crn_comp.cpp:155-214), alpha selectors as
//     XOR-chained 6-bit groups (crn_comp.cpp:268-293);
//   * level streams: reference group per 2x2 blocks, endpoint delta only for
//     reference 0, one selector symbol per block (crn_comp.cpp:295-422), with
//     references chosen as in crnlib/crn_dxt_hc.cpp:282-308 (left, then top);
//   * header and CRC16s as crnlib/crn_comp.cpp:1414-1496.
// The block quantizer in front of it (crn2_compress_rgba) is a simple
// deterministic CPU encoder used to produce inputs; it is not crnlib's dxt_hc.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <numeric>
#include <unordered_map>
#include <vector>

#include "crn2_internal.h"

namespace crn2 {

// ---------------------------------------------------------------- bit writer
struct BitWriter {
    std::vector<uint8_t> buf;
    uint64_t acc = 0;
    int n = 0;
    void put(uint32_t v, int bits)
    {
        if (!bits) return;
        acc = (acc << bits) | (v & ((bits == 32) ? 0xFFFFFFFFu : ((1u << bits) - 1)));
        n += bits;
        while (n >= 8) {
            buf.push_back((uint8_t)(acc >> (n - 8)));
            n -= 8;
        }
    }
    void finish()
    {
        if (n) buf.push_back((uint8_t)(acc << (8 - n)));
        n = 0;
        acc = 0;
    }
};

// --------------------------------------------------- length-limited Huffman
struct PrefixCode {
    std::vector<uint8_t> len;
    std::vector<uint32_t> code;
};

static void huffman_lengths(const std::vector<uint32_t>& freq, int limit, std::vector<uint8_t>& len)
{
    const size_t n = freq.size();
    len.assign(n, 0);
    std::vector<uint32_t> used;
    for (size_t i = 0; i < n; i++)
        if (freq[i]) used.push_back((uint32_t)i);
    if (used.empty()) return;
    if (used.size() == 1) {
        len[used[0]] = 1;
        return;
    }
    // two-queue Huffman over leaves sorted by (freq, symbol)
    std::sort(used.begin(), used.end(), [&](uint32_t a, uint32_t b) {
        return freq[a] != freq[b] ? freq[a] < freq[b] : a < b;
    });
    const size_t m = used.size();
    std::vector<uint64_t> w(2 * m);
    std::vector<int32_t> parent(2 * m, -1);
    for (size_t i = 0; i < m; i++) w[i] = freq[used[i]];
    size_t leaf = 0, inner = m, next = m;
    auto take = [&]() -> size_t {
        if (leaf < m && (inner >= next || w[leaf] <= w[inner])) return leaf++;
        return inner++;
    };
    while (next < 2 * m - 1) {
        size_t a = take(), b = take();
        w[next] = w[a] + w[b];
        parent[a] = parent[b] = (int32_t)next;
        next++;
    }
    std::vector<uint32_t> depth(2 * m, 0);
    for (size_t i = 2 * m - 2; i-- > 0;) depth[i] = depth[parent[i]] + 1;
    int maxlen = 0;
    for (size_t i = 0; i < m; i++) maxlen = std::max<int>(maxlen, (int)depth[i]);
    if (maxlen > limit) {
        // clamp, then restore the Kraft inequality by lengthening the least
        // frequent codes that are still shorter than the limit
        for (size_t i = 0; i < m; i++) depth[i] = std::min<uint32_t>(depth[i], (uint32_t)limit);
        uint64_t kraft = 0;
        for (size_t i = 0; i < m; i++) kraft += 1ull << (limit - depth[i]);
        const uint64_t full = 1ull << limit;
        while (kraft > full) {
            for (size_t i = 0; i < m && kraft > full; i++) {
                if ((int)depth[i] < limit) {
                    kraft -= 1ull << (limit - depth[i] - 1);
                    depth[i]++;
                }
            }
        }
    }
    for (size_t i = 0; i < m; i++) len[used[i]] = (uint8_t)depth[i];
}

static void canonical_codes(PrefixCode& pc)
{
    const size_t n = pc.len.size();
    pc.code.assign(n, 0);
    uint32_t count[17] = {0};
    for (size_t i = 0; i < n; i++)
        if (pc.len[i]) count[pc.len[i]]++;
    uint32_t next_code[17] = {0};
    uint32_t c = 0;
    for (int L = 1; L <= 16; L++) {
        next_code[L] = c;
        c = (c + count[L]) << 1;
    }
    for (size_t i = 0; i < n; i++)
        if (pc.len[i]) pc.code[i] = next_code[pc.len[i]]++;
}

static PrefixCode build_code(const std::vector<uint32_t>& freq, int limit)
{
    PrefixCode pc;
    huffman_lengths(freq, limit, pc.len);
    canonical_codes(pc);
    return pc;
}

static inline void put_sym(BitWriter& bw, const PrefixCode& pc, uint32_t s)
{
    bw.put(pc.code[s], pc.len[s]);
}

static const uint8_t k_cl_order[21] = {17, 18, 19, 20, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15, 16};

// code-length run coding: literal 0..16, 17 = 3..10 zeros (+3 bits),
// 18 = 11..138 zeros (+7), 19 = 3..6 repeats of the previous length (+2),
// 20 = 7..70 repeats (+6)
struct ClSym { uint8_t sym; uint8_t extra_bits; uint8_t extra; };

static void run_code_lengths(const std::vector<uint8_t>& len, uint32_t used, std::vector<ClSym>& out)
{
    uint32_t i = 0;
    while (i < used) {
        uint8_t v = len[i];
        uint32_t r = 1;
        while (i + r < used && len[i + r] == v) r++;
        if (v == 0) {
            uint32_t left = r;
            while (left) {
                if (left >= 11) {
                    uint32_t k = std::min<uint32_t>(left, 138);
                    out.push_back({18, 7, (uint8_t)(k - 11)});
                    left -= k;
                } else if (left >= 3) {
                    out.push_back({17, 3, (uint8_t)(left - 3)});
                    left = 0;
                } else {
                    out.push_back({0, 0, 0});
                    left--;
                }
            }
        } else {
            out.push_back({v, 0, 0});
            uint32_t left = r - 1;
            while (left) {
                if (left >= 7) {
                    uint32_t k = std::min<uint32_t>(left, 70);
                    out.push_back({20, 6, (uint8_t)(k - 7)});
                    left -= k;
                } else if (left >= 3) {
                    out.push_back({19, 2, (uint8_t)(left - 3)});
                    left = 0;
                } else {
                    out.push_back({v, 0, 0});
                    left--;
                }
            }
        }
        i += r;
    }
}

static void transmit_model(BitWriter& bw, const PrefixCode& pc)
{
    uint32_t used = 0;
    for (uint32_t i = (uint32_t)pc.len.size(); i > 0; i--)
        if (pc.len[i - 1]) { used = i; break; }
    bw.put(used, 14);
    if (!used) return;
    std::vector<ClSym> rl;
    run_code_lengths(pc.len, used, rl);
    std::vector<uint32_t> f(21, 0);
    for (auto& s : rl) f[s.sym]++;
    PrefixCode cl = build_code(f, 7);
    uint32_t nsend = 1;
    for (uint32_t i = 0; i < 21; i++)
        if (cl.len[k_cl_order[i]]) nsend = i + 1;
    bw.put(nsend, 5);
    for (uint32_t i = 0; i < nsend; i++) bw.put(cl.len[k_cl_order[i]], 3);
    for (auto& s : rl) {
        put_sym(bw, cl, s.sym);
        if (s.extra_bits) bw.put(s.extra, s.extra_bits);
    }
}

// ---------------------------------------------------------------- palettes
static std::vector<uint8_t> pack_color_endpoints(const std::vector<uint32_t>& ep)
{
    // component order per entry: r0 g0 b0 r1 g1 b1; table 1 for green
    std::vector<uint32_t> syms;
    std::vector<uint32_t> f0(32, 0), f1(64, 0);
    int prev[6] = {0};
    for (uint32_t e : ep) {
        int cur[6] = {(int)((e >> 11) & 31), (int)((e >> 5) & 63), (int)(e & 31),
                      (int)((e >> 27) & 31), (int)((e >> 21) & 63), (int)((e >> 16) & 31)};
        for (int k = 0; k < 6; k++) {
            const int mask = (k == 1 || k == 4) ? 63 : 31;
            uint32_t s = (uint32_t)((cur[k] - prev[k]) & mask);
            syms.push_back(s);
            ((k == 1 || k == 4) ? f1 : f0)[s]++;
            prev[k] = cur[k];
        }
    }
    PrefixCode c0 = build_code(f0, 15), c1 = build_code(f1, 15);
    BitWriter bw;
    transmit_model(bw, c0);
    transmit_model(bw, c1);
    for (size_t i = 0; i < syms.size(); i++) put_sym(bw, (i % 3 == 1) ? c1 : c0, syms[i]);
    bw.finish();
    return bw.buf;
}

static std::vector<uint8_t> pack_color_selectors(const std::vector<uint32_t>& lin)
{
    std::vector<uint32_t> f(16, 0), syms;
    uint32_t prev = 0;
    for (uint32_t s : lin) {
        uint32_t d = prev ^ s;
        prev = s;
        for (int c = 0; c < 8; c++, d >>= 4) {
            syms.push_back(d & 15);
            f[d & 15]++;
        }
    }
    PrefixCode pc = build_code(f, 15);
    BitWriter bw;
    transmit_model(bw, pc);
    for (uint32_t s : syms) put_sym(bw, pc, s);
    bw.finish();
    return bw.buf;
}

static std::vector<uint8_t> pack_alpha_endpoints(const std::vector<uint32_t>& ep)
{
    std::vector<uint32_t> f(256, 0), syms;
    int prev[2] = {0, 0};
    for (uint32_t e : ep) {
        int cur[2] = {(int)(e & 255), (int)((e >> 8) & 255)};
        for (int j = 0; j < 2; j++) {
            uint32_t s = (uint32_t)((cur[j] - prev[j]) & 255);
            syms.push_back(s);
            f[s]++;
            prev[j] = cur[j];
        }
    }
    PrefixCode pc = build_code(f, 15);
    BitWriter bw;
    transmit_model(bw, pc);
    for (uint32_t s : syms) put_sym(bw, pc, s);
    bw.finish();
    return bw.buf;
}

static std::vector<uint8_t> pack_alpha_selectors(const std::vector<uint64_t>& lin)
{
    std::vector<uint32_t> f(64, 0), syms;
    uint64_t prev = 0;
    for (uint64_t s : lin) {
        uint64_t d = prev ^ s;
        prev = s;
        for (int c = 0; c < 8; c++, d >>= 6) {
            syms.push_back((uint32_t)(d & 63));
            f[d & 63]++;
        }
    }
    PrefixCode pc = build_code(f, 15);
    BitWriter bw;
    transmit_model(bw, pc);
    for (uint32_t s : syms) put_sym(bw, pc, s);
    bw.finish();
    return bw.buf;
}

// ----------------------------------------------------------- palette order
// The entry order crnlib's compressor gives each palette before packing
// (crnlib/crn_comp.cpp:767-1352): the new index of every entry.
//   * selectors: a greedy chain from the all-zero word, each step taking the
//     remaining word nearest the last one taken (first on ties, remaining
//     words kept by swap-removal), nearness summed over texel pairs from a
//     table on the XOR of the two codes (colour: d = 0, 5, 14, 10 per 2-bit
//     code, optimize_color_selectors :935-986; alpha: d = 0, 2, 3, 3, 5, 5,
//     4, 4 per 3-bit code, optimize_alpha_selectors :1228-1273);
//   * endpoints: four candidate orders -- the same greedy chain on squared
//     endpoint distance from a zero entry (sort_color_endpoints :767-797,
//     sort_alpha_endpoints :1068-1097), and three chains grown at both ends
//     from the entry with the most endpoint changes, each step taking the
//     entry whose similarity to an end times its transition count with the
//     chain is largest (remap_color_endpoints :799-870 with similarity scale
//     1 + weight, weight 0, 1/6, 1/2; remap_alpha_endpoints :1099-1160) --
//     and the one whose packed palette plus coded index deltas is smallest
//     is kept (optimize_color :988-1047, optimize_alpha :1275-1352).
// Sums that the reference accumulates in 32-bit unsigned arithmetic wrap the
// same way here.
struct ColorEp { int lo[3], hi[3]; };

static uint32_t ep_dist(const ColorEp& a, const ColorEp& b)
{
    uint32_t e = 0;
    for (int k = 0; k < 3; k++) {
        const int dl = a.lo[k] - b.lo[k], dh = a.hi[k] - b.hi[k];
        e += (uint32_t)(dl * dl) + (uint32_t)(dh * dh);
    }
    return e;
}

static std::vector<uint16_t> chain_order(uint32_t n, const std::function<uint32_t(uint32_t, uint32_t)>& dist,
                                         uint32_t origin)
{
    // dist(i, j): entry i against entry j; `origin` = n stands for the zero entry
    std::vector<uint16_t> remap(n), idx(n);
    for (uint32_t i = 0; i < n; i++) idx[i] = (uint16_t)i;
    uint32_t sel = origin;
    for (uint32_t left = n; left;) {
        uint32_t best = 0, min_err = 0xFFFFFFFFu;
        for (uint32_t i = 0; i < left; i++) {
            const uint32_t e = dist(idx[i], sel);
            if (e < min_err) {
                min_err = e;
                best = i;
            }
        }
        sel = idx[best];
        remap[idx[best]] = (uint16_t)(n - left);
        left--;
        idx[best] = idx[left];
    }
    return remap;
}

static std::vector<uint16_t> remap_color_chain(const std::vector<ColorEp>& e, const std::vector<uint32_t>& hist,
                                               uint32_t n, uint32_t selected, float weight)
{
    struct Node { uint32_t index, frequency, front_sim, back_sim; };
    std::vector<Node> rem(n);
    for (uint32_t i = 0; i < n; i++) rem[i] = {i, 0, 0, 0};
    std::vector<uint32_t> chosen(2 * (size_t)n + 1);
    uint32_t count = n, cf = n, cb = n;
    chosen[cf] = selected;
    ColorEp fe = e[selected], be = fe;
    bool fu = true, bu = true;
    rem[selected] = rem[--count];
    const uint32_t* freq = &hist[(size_t)selected * n];
    const uint32_t base = (uint32_t)(4000 * (1.0f + weight));
    uint32_t norm = 0;
    while (count) {
        uint64_t best_value = 0;
        uint32_t best = 0;
        for (uint32_t i = 0; i < count; i++) {
            Node& nd = rem[i];
            nd.frequency += freq[nd.index];
            if (fu) nd.front_sim = base - std::min<uint32_t>(4000, ep_dist(e[nd.index], fe));
            if (bu) nd.back_sim = base - std::min<uint32_t>(4000, ep_dist(e[nd.index], be));
            const uint64_t v = (uint64_t)(uint32_t)(std::max(nd.front_sim, nd.back_sim) * (nd.frequency + norm) + 1u);
            if (v > best_value || (v == best_value && nd.index < selected)) {
                best_value = v;
                best = i;
                selected = nd.index;
            }
        }
        freq = &hist[(size_t)selected * n];
        uint32_t ff = 0, fb = 0;
        for (int front = (int)cf, back = (int)cb, scale = back - front; scale > 0; front++, back--, scale -= 2) {
            ff += (uint32_t)scale * freq[chosen[front]];
            fb += (uint32_t)scale * freq[chosen[back]];
        }
        fu = bu = false;
        const Node nd = rem[best];
        norm = nd.frequency << 3;
        if ((uint64_t)nd.front_sim * ff > (uint64_t)nd.back_sim * fb) {
            chosen[--cf] = selected;
            fe = e[selected];
            fu = true;
        } else {
            chosen[++cb] = selected;
            be = e[selected];
            bu = true;
        }
        rem[best] = rem[--count];
    }
    std::vector<uint16_t> remap(n);
    for (uint32_t i = cf; i <= cb; i++) remap[chosen[i]] = (uint16_t)(i - cf);
    return remap;
}

static std::vector<uint16_t> remap_alpha_chain(const std::vector<uint32_t>& ep, const std::vector<uint32_t>& hist,
                                               uint32_t n, uint32_t selected, float weight)
{
    auto dist = [&](uint32_t a, uint32_t b) {
        const int dl = (int)(ep[a] & 255) - (int)(ep[b] & 255), dh = (int)((ep[a] >> 8) & 255) - (int)((ep[b] >> 8) & 255);
        return (uint32_t)(dl * dl + dh * dh);
    };
    const uint32_t* freq = &hist[(size_t)selected * n];
    std::vector<uint32_t> chosen_front, chosen_back, remaining, total(n, 0);   // chain = reverse(front) + back
    chosen_back.push_back(selected);
    for (uint32_t i = 0; i < n; i++)
        if (i != selected) {
            remaining.push_back(i);
            total[i] = freq[i];
        }
    auto chain_at = [&](size_t k) -> uint32_t {
        return k < chosen_front.size() ? chosen_front[chosen_front.size() - 1 - k]
                                       : chosen_back[k - chosen_front.size()];
    };
    const uint32_t base = (uint32_t)(1000 * (1.0f + weight));
    uint32_t norm = 0;
    while (!remaining.empty()) {
        const uint32_t ef = chain_at(0), eb = chain_at(chosen_front.size() + chosen_back.size() - 1);
        uint32_t sel_i = 0;
        uint64_t best = 0, sf = 0, sb = 0;
        for (uint32_t i = 0; i < remaining.size(); i++) {
            const uint32_t r = remaining[i];
            const uint64_t simf = base - std::min<uint32_t>(dist(r, ef), 1000);
            const uint64_t simb = base - std::min<uint32_t>(dist(r, eb), 1000);
            const uint64_t v = std::max(simf, simb) * (uint64_t)(uint32_t)(total[r] + norm) + 1;
            if (v > best) {
                best = v;
                sel_i = i;
                sf = simf;
                sb = simb;
            }
        }
        selected = remaining[sel_i];
        freq = &hist[(size_t)selected * n];
        norm = total[selected];
        uint32_t ff = 0, fb = 0;
        const int len = (int)(chosen_front.size() + chosen_back.size());
        for (int front = 0, back = len - 1, scale = back; scale > 0; front++, back--, scale -= 2) {
            ff += (uint32_t)scale * freq[chain_at((size_t)front)];
            fb += (uint32_t)scale * freq[chain_at((size_t)back)];
        }
        if (sf * ff > sb * fb) chosen_front.push_back(selected);
        else chosen_back.push_back(selected);
        remaining.erase(remaining.begin() + sel_i);
        for (uint32_t r : remaining) total[r] += freq[r];
    }
    std::vector<uint16_t> remap(n);
    const size_t len = chosen_front.size() + chosen_back.size();
    for (size_t k = 0; k < len; k++) remap[chain_at(k)] = (uint16_t)k;
    return remap;
}

static std::vector<uint16_t> order_color_selectors(const std::vector<uint32_t>& sel)
{
    static const uint8_t d[4] = {0, 5, 14, 10};
    std::vector<uint8_t> D4(256), D8(65536);
    for (uint32_t i = 0; i < 256; i++) D4[i] = (uint8_t)(d[(i ^ (i >> 4)) & 3] + d[((i >> 2) ^ (i >> 6)) & 3]);
    for (uint32_t i = 0; i < 65536; i++)
        D8[i] = (uint8_t)(D4[((i >> 8) & 0xF0) | ((i >> 4) & 0xF)] + D4[((i >> 4) & 0xF0) | (i & 0xF)]);
    const uint32_t n = (uint32_t)sel.size();
    return chain_order(n, [&](uint32_t i, uint32_t j) {
        const uint32_t s = sel[i], t = j == n ? 0u : sel[j];
        return (uint32_t)D8[((s >> 16) & 0xFF00) | ((t >> 24) & 0xFF)] + D8[((s >> 8) & 0xFF00) | ((t >> 16) & 0xFF)] +
               D8[(s & 0xFF00) | ((t >> 8) & 0xFF)] + D8[((s << 8) & 0xFF00) | (t & 0xFF)];
    }, n);
}

static std::vector<uint16_t> order_alpha_selectors(const std::vector<uint64_t>& sel)
{
    static const uint8_t d[8] = {0, 2, 3, 3, 5, 5, 4, 4};
    std::vector<uint8_t> D6(4096);
    for (uint32_t i = 0; i < 4096; i++) D6[i] = (uint8_t)(d[(i ^ (i >> 6)) & 7] + d[((i >> 3) ^ (i >> 9)) & 7]);
    const uint32_t n = (uint32_t)sel.size();
    return chain_order(n, [&](uint32_t i, uint32_t j) {
        uint64_t s = sel[i] << 6, t = j == n ? 0ull : sel[j];
        uint32_t e = 0;
        for (int k = 0; k < 8; k++, s >>= 6, t >>= 6) e += D6[(s & 0xFC0) | (t & 0x3F)];
        return e;
    }, n);
}

static inline uint16_t crc16_update(uint16_t crc, const uint8_t* p, size_t len)
{
    uint32_t c = (uint16_t)~crc;
    for (size_t i = 0; i < len; i++) {
        uint32_t q = (uint32_t)(p[i] ^ (c >> 8)) & 0xFF;
        c = (c << 8) & 0xFFFF;
        uint32_t r = (q >> 4) ^ q;
        c ^= r;
        r <<= 5;
        c ^= r;
        r <<= 7;
        c ^= r;
        c &= 0xFFFF;
    }
    return (uint16_t)~c;
}

static void put_be(uint8_t* p, uint32_t v, int n)
{
    for (int i = n - 1; i >= 0; i--) {
        p[i] = (uint8_t)v;
        v >>= 8;
    }
}

}  // namespace crn2

using namespace crn2;

extern "C" uint16_t crn2_crc16(const void* p, uint32_t len, uint16_t crc)
{
    return crc16_update(crc, (const uint8_t*)p, len);
}

// Packs already-quantized blocks into a .crn file.  See crn2_internal.h.
extern "C" int crn2_pack(const crn2_pack_input* in, uint8_t** out, uint32_t* out_size)
{
    if (!in || !out || !out_size) return CRN2_EINVAL;
    *out = nullptr;
    *out_size = 0;
    const uint32_t fmt = in->format;
    const bool dxt5_family = fmt == CRN2_FMT_DXT5 || fmt == CRN2_FMT_DXT5_CCXY || fmt == CRN2_FMT_DXT5_XGXR ||
                             fmt == CRN2_FMT_DXT5_XGBR || fmt == CRN2_FMT_DXT5_AGBR;
    const bool has_color = (fmt == CRN2_FMT_DXT1 || dxt5_family);
    const int n_alpha = (fmt == CRN2_FMT_DXT1) ? 0 : (fmt == CRN2_FMT_DXN_XY || fmt == CRN2_FMT_DXN_YX) ? 2 : 1;
    if (fmt != CRN2_FMT_DXT1 && !dxt5_family && fmt != CRN2_FMT_DXT5A && fmt != CRN2_FMT_DXN_XY &&
        fmt != CRN2_FMT_DXN_YX)
        return CRN2_EUNSUPPORTED;
    if (in->levels < 1 || in->levels > 16 || (in->faces != 1 && in->faces != 6)) return CRN2_EINVAL;
    if (in->width < 1 || in->width > 4096 || in->height < 1 || in->height > 4096) return CRN2_EINVAL;
    if (has_color && (in->n_color_endpoints < 1 || in->n_color_endpoints > 8192 || in->n_color_selectors < 1 ||
                      in->n_color_selectors > 8192))
        return CRN2_EINVAL;
    if (n_alpha && (in->n_alpha_endpoints < 1 || in->n_alpha_endpoints > 8192 || in->n_alpha_selectors < 1 ||
                    in->n_alpha_selectors > 8192))
        return CRN2_EINVAL;

    // block grid per level (even-padded, faces stacked)
    std::vector<uint32_t> lw(in->levels), lh(in->levels), first(in->levels + 1, 0);
    for (uint32_t l = 0; l < in->levels; l++) {
        uint32_t w = std::max<uint32_t>(1, in->width >> l), h = std::max<uint32_t>(1, in->height >> l);
        lw[l] = (((w + 3) >> 2) + 1) & ~1u;
        lh[l] = (((h + 3) >> 2) + 1) & ~1u;
        first[l + 1] = first[l] + lw[l] * lh[l] * in->faces;
    }
    const uint32_t nblocks = first[in->levels];

    // histograms: reference groups, endpoint deltas, selector indices
    const uint32_t nce = has_color ? in->n_color_endpoints : 0, ncs = has_color ? in->n_color_selectors : 0;
    const uint32_t nae = n_alpha ? in->n_alpha_endpoints : 0, nas = n_alpha ? in->n_alpha_selectors : 0;
    std::vector<uint32_t> fref(256, 0), fce(std::max<uint32_t>(nce, 1), 0), fcs(std::max<uint32_t>(ncs, 1), 0);
    std::vector<uint32_t> fae(std::max<uint32_t>(nae, 1), 0), fas(std::max<uint32_t>(nas, 1), 0);
    std::vector<uint8_t> refs(nblocks, 0);

    auto idx_ok = [&](uint32_t b) {
        if (has_color && (in->color_endpoint_idx[b] >= nce || in->color_selector_idx[b] >= ncs)) return false;
        if (n_alpha >= 1 && (in->alpha0_endpoint_idx[b] >= nae || in->alpha0_selector_idx[b] >= nas)) return false;
        if (n_alpha == 2 && (in->alpha1_endpoint_idx[b] >= nae || in->alpha1_selector_idx[b] >= nas)) return false;
        return true;
    };
    auto same_ep = [&](uint32_t a, uint32_t b) {
        if (has_color && in->color_endpoint_idx[a] != in->color_endpoint_idx[b]) return false;
        if (n_alpha >= 1 && in->alpha0_endpoint_idx[a] != in->alpha0_endpoint_idx[b]) return false;
        if (n_alpha == 2 && in->alpha1_endpoint_idx[a] != in->alpha1_endpoint_idx[b]) return false;
        return true;
    };
    for (uint32_t b = 0; b < nblocks; b++)
        if (!idx_ok(b)) return CRN2_EINVAL;

    for (uint32_t l = 0; l < in->levels; l++) {
        const uint32_t W = lw[l], rows = lh[l] * in->faces;
        for (uint32_t y = 0; y < rows; y++)
            for (uint32_t x = 0; x < W; x++) {
                const uint32_t b = first[l] + y * W + x;
                bool top = y != 0, left = top || x != 0;
                uint8_t r = 0;
                if (left && same_ep(b, b - 1)) r = 1;
                else if (top && same_ep(b, b - W)) r = 2;
                refs[b] = r;
            }
    }

    // working copies of the palettes and per-block indices (reordered below when asked)
    std::vector<uint32_t> ce, cs, ae, ice, ics, ia[4];
    std::vector<uint64_t> as;
    if (has_color) {
        ce.assign(in->color_endpoints, in->color_endpoints + nce);
        cs.assign(in->color_selectors_linear, in->color_selectors_linear + ncs);
        ice.assign(in->color_endpoint_idx, in->color_endpoint_idx + nblocks);
        ics.assign(in->color_selector_idx, in->color_selector_idx + nblocks);
    }
    if (n_alpha) {
        ae.assign(in->alpha_endpoints, in->alpha_endpoints + nae);
        as.assign(in->alpha_selectors_linear, in->alpha_selectors_linear + nas);
        ia[0].assign(in->alpha0_endpoint_idx, in->alpha0_endpoint_idx + nblocks);
        ia[1].assign(in->alpha0_selector_idx, in->alpha0_selector_idx + nblocks);
        if (n_alpha == 2) {
            ia[2].assign(in->alpha1_endpoint_idx, in->alpha1_endpoint_idx + nblocks);
            ia[3].assign(in->alpha1_selector_idx, in->alpha1_selector_idx + nblocks);
        }
    }
    if (in->order_palettes) {
        // bits of an endpoint order: its packed palette plus the level streams' index deltas under it
        auto delta_bits = [&](const std::vector<uint16_t>& remap, uint32_t n,
                              std::initializer_list<const std::vector<uint32_t>*> chans, size_t palette_bytes) {
            std::vector<uint32_t> hist(n, 0);
            for (uint32_t l = 0; l < in->levels; l++)
                for (const std::vector<uint32_t>* ch : chans) {
                    uint32_t prev = 0;
                    for (uint32_t b = first[l]; b < first[l + 1]; b++) {
                        const uint32_t idx = remap[(*ch)[b]];
                        if (!refs[b]) hist[(idx + n - prev) % n]++;
                        prev = idx;
                    }
                }
            PrefixCode pc = build_code(hist, 16);
            uint64_t bits = (uint64_t)palette_bytes * 8;
            for (uint32_t sy = 0; sy < n; sy++) bits += (uint64_t)hist[sy] * pc.len[sy];
            BitWriter bw;
            transmit_model(bw, pc);
            return bits + bw.buf.size() * 8 + (uint64_t)bw.n;
        };
        // transition counts between consecutive blocks' endpoint indices (blocks with deltas only)
        auto transitions = [&](uint32_t n, std::initializer_list<const std::vector<uint32_t>*> chans,
                               std::vector<uint32_t>& hist) {
            hist.assign((size_t)n * n, 0);
            std::vector<uint32_t> sum(n, 0);
            for (const std::vector<uint32_t>* ch : chans) {
                uint32_t prev = 0;
                for (uint32_t b = 0; b < nblocks; b++) {
                    const uint32_t i = (*ch)[b];
                    if (!refs[b] && i != prev) {
                        hist[(size_t)i * n + prev]++;
                        hist[(size_t)prev * n + i]++;
                        sum[i]++;
                        sum[prev]++;
                    }
                    prev = i;
                }
            }
            uint32_t selected = 0, best = 0;
            for (uint32_t i = 0; i < n; i++)
                if (best < sum[i]) {
                    best = sum[i];
                    selected = i;
                }
            return selected;
        };
        auto apply = [&](const std::vector<uint16_t>& remap, std::vector<uint32_t>* idx) {
            for (uint32_t& v : *idx) v = remap[v];
        };
        const float weights[3] = {0.0f, 1.0f / 6.0f, 0.5f};
        if (has_color) {
            std::vector<ColorEp> un(nce);
            for (uint32_t i = 0; i < nce; i++)
                for (int h = 0; h < 2; h++) {
                    const uint32_t c = (ce[i] >> (16 * h)) & 0xFFFF;
                    int* o = h ? un[i].hi : un[i].lo;
                    const int r = (c >> 11) & 31, g = (c >> 5) & 63, b = c & 31;
                    o[0] = (r << 3) | (r >> 2);
                    o[1] = (g << 2) | (g >> 4);
                    o[2] = (b << 3) | (b >> 2);
                }
            std::vector<uint32_t> hist;
            const uint32_t sel = transitions(nce, {&ice}, hist);
            const ColorEp zero = {{0, 0, 0}, {0, 0, 0}};
            std::vector<std::vector<uint16_t>> trial;
            trial.push_back(chain_order(nce, [&](uint32_t i, uint32_t j) { return ep_dist(un[i], j == nce ? zero : un[j]); },
                                        nce));
            for (float w : weights) trial.push_back(remap_color_chain(un, hist, nce, sel, w));
            uint64_t best_bits = ~0ull;
            std::vector<uint16_t> best;
            for (auto& r : trial) {
                std::vector<uint32_t> pal(nce);
                for (uint32_t i = 0; i < nce; i++) pal[r[i]] = ce[i];
                const uint64_t bits = delta_bits(r, nce, {&ice}, pack_color_endpoints(pal).size());
                if (bits < best_bits) {
                    best_bits = bits;
                    best = r;
                }
            }
            std::vector<uint32_t> pal(nce);
            for (uint32_t i = 0; i < nce; i++) pal[best[i]] = ce[i];
            ce.swap(pal);
            apply(best, &ice);
            const std::vector<uint16_t> rs = order_color_selectors(cs);
            std::vector<uint32_t> spal(ncs);
            for (uint32_t i = 0; i < ncs; i++) spal[rs[i]] = cs[i];
            cs.swap(spal);
            apply(rs, &ics);
        }
        if (n_alpha) {
            std::vector<uint32_t> hist;
            uint32_t sel;
            if (n_alpha == 2) sel = transitions(nae, {&ia[0], &ia[2]}, hist);
            else sel = transitions(nae, {&ia[0]}, hist);
            std::vector<std::vector<uint16_t>> trial;
            trial.push_back(chain_order(nae, [&](uint32_t i, uint32_t j) {
                const int l0 = (int)(ae[i] & 255), h0 = (int)((ae[i] >> 8) & 255);
                const int l1 = j == nae ? 0 : (int)(ae[j] & 255), h1 = j == nae ? 0 : (int)((ae[j] >> 8) & 255);
                return (uint32_t)((l0 - l1) * (l0 - l1) + (h0 - h1) * (h0 - h1));
            }, nae));
            for (float w : weights) trial.push_back(remap_alpha_chain(ae, hist, nae, sel, w));
            uint64_t best_bits = ~0ull;
            std::vector<uint16_t> best;
            for (auto& r : trial) {
                std::vector<uint32_t> pal(nae);
                for (uint32_t i = 0; i < nae; i++) pal[r[i]] = ae[i];
                const size_t pb = pack_alpha_endpoints(pal).size();
                const uint64_t bits = n_alpha == 2 ? delta_bits(r, nae, {&ia[0], &ia[2]}, pb)
                                                   : delta_bits(r, nae, {&ia[0]}, pb);
                if (bits < best_bits) {
                    best_bits = bits;
                    best = r;
                }
            }
            std::vector<uint32_t> pal(nae);
            for (uint32_t i = 0; i < nae; i++) pal[best[i]] = ae[i];
            ae.swap(pal);
            apply(best, &ia[0]);
            if (n_alpha == 2) apply(best, &ia[2]);
            const std::vector<uint16_t> rs = order_alpha_selectors(as);
            std::vector<uint64_t> spal(nas);
            for (uint32_t i = 0; i < nas; i++) spal[rs[i]] = as[i];
            as.swap(spal);
            apply(rs, &ia[1]);
            if (n_alpha == 2) apply(rs, &ia[3]);
        }
    }

    struct Level { std::vector<uint8_t> bytes; };
    std::vector<Level> lv(in->levels);
    PrefixCode cref, cce, ccs, cae, cas;
    for (int pass = 0; pass < 2; pass++) {
        for (uint32_t l = 0; l < in->levels; l++) {
            BitWriter bw;
            const uint32_t W = lw[l], rows = lh[l] * in->faces;
            uint32_t run[3] = {0, 0, 0};
            for (uint32_t y = 0; y < rows; y++)
                for (uint32_t x = 0; x < W; x++) {
                    const uint32_t b = first[l] + y * W + x;
                    if (!(y & 1) && !(x & 1)) {
                        uint32_t g = refs[b] | (refs[b + W] << 2) | (refs[b + 1] << 4) | (refs[b + W + 1] << 6);
                        if (pass) put_sym(bw, cref, g); else fref[g]++;
                    }
                    uint32_t cur[3] = {has_color ? ice[b] : 0, n_alpha ? ia[0][b] : 0, n_alpha == 2 ? ia[2][b] : 0};
                    if (refs[b] == 0) {
                        if (has_color) {
                            uint32_t d = (cur[0] + nce - run[0]) % nce;
                            if (pass) put_sym(bw, cce, d); else fce[d]++;
                        }
                        for (int a = 0; a < n_alpha; a++) {
                            uint32_t d = (cur[1 + a] + nae - run[1 + a]) % nae;
                            if (pass) put_sym(bw, cae, d); else fae[d]++;
                        }
                    }
                    run[0] = cur[0];
                    run[1] = cur[1];
                    run[2] = cur[2];
                    if (has_color) {
                        uint32_t s = ics[b];
                        if (pass) put_sym(bw, ccs, s); else fcs[s]++;
                    }
                    if (n_alpha >= 1) {
                        uint32_t s = ia[1][b];
                        if (pass) put_sym(bw, cas, s); else fas[s]++;
                    }
                    if (n_alpha == 2) {
                        uint32_t s = ia[3][b];
                        if (pass) put_sym(bw, cas, s); else fas[s]++;
                    }
                }
            bw.finish();
            if (pass) {
                if (bw.buf.empty()) bw.buf.push_back(0);
                lv[l].bytes.swap(bw.buf);
            }
        }
        if (!pass) {
            cref = build_code(fref, 16);
            if (has_color) { cce = build_code(fce, 16); ccs = build_code(fcs, 16); }
            if (n_alpha) { cae = build_code(fae, 16); cas = build_code(fas, 16); }
        }
    }

    BitWriter tb;
    transmit_model(tb, cref);
    if (has_color) { transmit_model(tb, cce); transmit_model(tb, ccs); }
    if (n_alpha) { transmit_model(tb, cae); transmit_model(tb, cas); }
    tb.finish();

    std::vector<uint8_t> p_ce, p_cs, p_ae, p_as;
    if (has_color) {
        p_ce = pack_color_endpoints(ce);
        p_cs = pack_color_selectors(cs);
    }
    if (n_alpha) {
        p_ae = pack_alpha_endpoints(ae);
        p_as = pack_alpha_selectors(as);
    }

    const uint32_t hsize = 74 + 4 * (in->levels - 1);
    std::vector<uint8_t> f(hsize, 0);
    auto append = [&](const std::vector<uint8_t>& v, uint32_t& ofs) {
        ofs = (uint32_t)f.size();
        f.insert(f.end(), v.begin(), v.end());
    };
    uint32_t o_ce = 0, o_cs = 0, o_ae = 0, o_as = 0, o_tab = 0;
    if (has_color) { append(p_ce, o_ce); append(p_cs, o_cs); }
    if (n_alpha) { append(p_ae, o_ae); append(p_as, o_as); }
    append(tb.buf, o_tab);
    std::vector<uint32_t> lofs(in->levels);
    for (uint32_t l = 0; l < in->levels; l++) append(lv[l].bytes, lofs[l]);
    if (f.size() >= (1u << 24) && (has_color ? o_cs : o_as) >= (1u << 24)) return CRN2_EINVAL;

    uint8_t* h = f.data();
    put_be(h + 0, ('H' << 8) | 'x', 2);
    put_be(h + 2, hsize, 2);
    put_be(h + 6, (uint32_t)f.size(), 4);
    put_be(h + 12, in->width, 2);
    put_be(h + 14, in->height, 2);
    h[16] = (uint8_t)in->levels;
    h[17] = (uint8_t)in->faces;
    h[18] = (uint8_t)fmt;
    put_be(h + 25, in->userdata0, 4);
    put_be(h + 29, in->userdata1, 4);
    auto put_pal = [&](int i, uint32_t ofs, size_t size, uint32_t num) {
        put_be(h + 33 + 8 * i, size ? ofs : 0, 3);
        put_be(h + 36 + 8 * i, (uint32_t)size, 3);
        put_be(h + 39 + 8 * i, size ? num : 0, 2);
    };
    put_pal(0, o_ce, p_ce.size(), nce);
    put_pal(1, o_cs, p_cs.size(), ncs);
    put_pal(2, o_ae, p_ae.size(), nae);
    put_pal(3, o_as, p_as.size(), nas);
    put_be(h + 65, (uint32_t)tb.buf.size(), 2);
    put_be(h + 67, o_tab, 3);
    for (uint32_t l = 0; l < in->levels; l++) put_be(h + 70 + 4 * l, lofs[l], 4);
    put_be(h + 10, crc16_update(0, h + hsize, f.size() - hsize), 2);
    put_be(h + 4, crc16_update(0, h + 6, hsize - 6), 2);

    uint8_t* o = (uint8_t*)malloc(f.size());
    if (!o) return CRN2_ENOMEM;
    memcpy(o, f.data(), f.size());
    *out = o;
    *out_size = (uint32_t)f.size();
    return CRN2_OK;
}

extern "C" void crn2_free(void* p) { free(p); }
