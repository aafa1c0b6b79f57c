// crnd_host.cpp -- host half of the transcoder: header fields, prefix-code
// models and their device lookup tables, and palette decoding.
//
// Reference semantics followed (inc/crn_decomp.h):
//   header fields ........... crn_defs.h:227-341, crnd_get_header 2657-2670
//   canonical code tables ... decoder_tables::init 2150-2326
//   bit reader / decode ..... symbol_codec 3125-3253 (zero bytes past the end;
//                             a pattern matching no code yields symbol 0 and
//                             consumes nothing)
//   model transmission ...... decode_receive_static_data_model 3004-3123
//   palettes ................ decode_color_endpoints 3715-3761,
//                             decode_color_selectors 3763-3798,
//                             decode_alpha_endpoints 3800-3827,
//                             decode_alpha_selectors 3829-3851
#include "crnd_host.h"

#include <algorithm>
#include <cstring>

#include "crn2_internal.h"

namespace crn2 {

static uint32_t be(const uint8_t* p, int n)
{
    uint32_t v = 0;
    for (int i = 0; i < n; i++) v = (v << 8) | p[i];
    return v;
}

bool parse_header(const uint8_t* d, uint32_t size, Header& h)
{
    if (!d || size < 74) return false;
    if (be(d, 2) != (('H' << 8) | 'x')) return false;
    h.header_size = be(d + 2, 2);
    h.header_crc = be(d + 4, 2);
    h.data_size = be(d + 6, 4);
    h.data_crc = be(d + 10, 2);
    h.width = be(d + 12, 2);
    h.height = be(d + 14, 2);
    h.levels = d[16];
    h.faces = d[17];
    h.format = d[18];
    h.flags = be(d + 19, 2);
    h.userdata0 = be(d + 25, 4);
    h.userdata1 = be(d + 29, 4);
    for (int i = 0; i < 4; i++) {
        h.pal_ofs[i] = be(d + 33 + 8 * i, 3);
        h.pal_size[i] = be(d + 36 + 8 * i, 3);
        h.pal_num[i] = be(d + 39 + 8 * i, 2);
    }
    h.tables_size = be(d + 65, 2);
    h.tables_ofs = be(d + 67, 3);
    // the CRCs run over [6, header_size) and [header_size, data_size): both must lie inside the buffer
    if (h.header_size < 74 || size < h.data_size || h.data_size < h.header_size) return false;
    for (uint32_t l = 0; l < kMaxLevels; l++)
        h.level_ofs[l] = (l < h.levels && 70 + 4 * l + 4 <= h.header_size) ? be(d + 70 + 4 * l, 4) : 0;
    return true;
}

bool build_model(HostModel& m, uint32_t n, const uint8_t* sizes)
{
    m.valid = false;
    if (!n || n > kMaxSyms) return false;
    m.num_syms = n;
    m.sizes.assign(sizes, sizes + n);
    uint32_t count[17] = {0};
    for (uint32_t i = 0; i < n; i++) {
        if (sizes[i] > 16) return false;
        if (sizes[i]) count[sizes[i]]++;
    }
    uint32_t first[17] = {0};
    uint32_t code = 0, pos = 0;
    for (uint32_t L = 1; L <= 16; L++) {
        m.limit[L] = 0;
        m.base[L] = 0;
        if (count[L]) {
            const uint32_t last = code + count[L] - 1;
            m.limit[L] = 1 + ((last << (16 - L)) | ((1u << (16 - L)) - 1));
            m.base[L] = (int32_t)pos - (int32_t)code;
            first[L] = pos;
            pos += count[L];
            code += count[L];
        }
        code <<= 1;
    }
    m.sorted.assign(pos ? pos : 1, 0);
    for (uint32_t i = 0; i < n; i++)
        if (sizes[i]) m.sorted[first[sizes[i]]++] = (uint16_t)i;
    m.valid = true;
    return true;
}

void to_device_model(const HostModel& m, DevModel& dm)
{
    memset(&dm, 0, sizeof(dm));
    dm.num_syms = m.num_syms;
    for (int L = 0; L <= 16; L++) {
        dm.limit[L] = m.limit[L];
        dm.base[L] = m.base[L];
    }
    // every 11-bit prefix whose code is <= 11 bits long resolves directly
    for (uint32_t v = 0; v < kLutSize; v++) {
        const uint32_t k = (v << (16 - kLutBits)) + 1;  // smallest 16-bit window with this prefix, plus one
        uint32_t L = 1;
        while (L <= kLutBits && k > m.limit[L]) L++;
        if (L > kLutBits) continue;
        // the window's top L bits decide the code; all windows sharing the
        // 11-bit prefix agree on them because L <= 11
        const uint32_t idx = (uint32_t)(m.base[L] + (int32_t)(v >> (kLutBits - L)));
        if (idx >= m.num_syms) continue;
        dm.lut[v] = (uint32_t)m.sorted[idx] | (L << 16);
        dm.len8[v] = (uint8_t)L;
    }
}

void prefix_lengths(const DevModel& dm, uint8_t* out)
{
    // The model's canonical limits of lengths 12..16, made non-decreasing,
    // give each long window's length as 12 + the number of limits at or below
    // it (0 from the last limit on), the canonical walk's length; the walk's
    // index, base[L] + code, grows with the window, so its check holds for
    // every window once it holds for the first and last window of each length
    // (from the first window the direct table leaves open).
    uint32_t lim[5], mx = 0;
    for (uint32_t L = kLutBits + 1; L <= 16; L++) lim[L - kLutBits - 1] = mx = std::max(mx, dm.limit[L]);
    auto len_at = [&](uint32_t top) -> uint32_t {
        if (top >= lim[4]) return 0;
        uint32_t L = kLutBits + 1;
        for (uint32_t i = 0; i < 4; i++) L += top >= lim[i] ? 1u : 0u;
        return L;
    };
    uint32_t h = 0;
    while (h < kLutSize && dm.len8[h]) h++;
    bool exact = true;
    uint32_t lo = h << (16 - kLutBits);
    for (uint32_t L = kLutBits + 1; L <= 16; L++) {
        const uint32_t hi = lim[L - kLutBits - 1];
        if (lo < hi) {
            const int64_t i0 = (int64_t)dm.base[L] + (lo >> (16 - L));
            const int64_t i1 = (int64_t)dm.base[L] + ((hi - 1) >> (16 - L));
            exact &= i0 >= 0 && i1 < (int64_t)dm.num_syms;
        }
        lo = std::max(lo, hi);
    }
    for (uint32_t v = 0; v < kLutSize; v++) {
        if (dm.len8[v]) {
            out[v] = dm.len8[v];
            continue;
        }
        const uint32_t w = v << (16 - kLutBits), a = len_at(w), b = len_at(w + (1u << (16 - kLutBits)) - 1);
        out[v] = (exact && a == b) ? (uint8_t)a : (uint8_t)0xFF;
    }
}

uint32_t HostBits::get_bits(uint32_t n)
{
    while (count < (int)n) {
        uint32_t c = (next != end) ? *next++ : 0;
        count += 8;
        buf |= c << (32 - count);
    }
    uint32_t r = buf >> (32 - n);
    buf <<= n;
    count -= n;
    return r;
}

uint32_t HostBits::decode_bits(uint32_t n)
{
    if (!n) return 0;
    if (n > 16) {
        uint32_t hi = get_bits(n - 16);
        return (hi << 16) | get_bits(16);
    }
    return get_bits(n);
}

uint32_t HostBits::decode(const HostModel& m)
{
    if (count < 24) {
        if (count < 16) {
            uint32_t c0 = 0, c1 = 0;
            if (next < end) c0 = *next++;
            if (next < end) c1 = *next++;
            count += 16;
            buf |= ((c0 << 8) | c1) << (32 - count);
        } else {
            uint32_t c = (next < end) ? *next++ : 0;
            count += 8;
            buf |= c << (32 - count);
        }
    }
    const uint32_t k = (buf >> 16) + 1;
    uint32_t L = 1;
    while (L <= 16 && k > m.limit[L]) L++;
    if (L > 16) return 0;
    const uint32_t idx = (uint32_t)(m.base[L] + (int32_t)(buf >> (32 - L)));
    if (idx >= m.num_syms) return 0;
    buf <<= L;
    count -= L;
    return m.sorted[idx];
}

static const uint8_t k_cl_order[21] = {17, 18, 19, 20, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15, 16};

bool receive_model(HostBits& b, HostModel& m)
{
    m.valid = false;
    const uint32_t used = b.decode_bits(14);
    if (!used) return true;
    if (used > kMaxSyms) return false;
    std::vector<uint8_t> sizes(used, 0);
    const uint32_t ncl = b.decode_bits(5);
    if (ncl < 1 || ncl > 21) return false;
    uint8_t cls[21] = {0};
    for (uint32_t i = 0; i < ncl; i++) cls[k_cl_order[i]] = (uint8_t)b.decode_bits(3);
    HostModel cl;
    if (!build_model(cl, 21, cls)) return false;
    uint32_t ofs = 0;
    while (ofs < used) {
        const uint32_t left = used - ofs;
        const uint32_t c = b.decode(cl);
        if (c <= 16) {
            sizes[ofs++] = (uint8_t)c;
        } else if (c == 17 || c == 18) {
            const uint32_t len = (c == 17) ? b.decode_bits(3) + 3 : b.decode_bits(7) + 11;
            if (len > left) return false;
            ofs += len;
        } else if (c == 19 || c == 20) {
            const uint32_t len = (c == 19) ? b.decode_bits(2) + 3 : b.decode_bits(6) + 7;
            if (!ofs || len > left) return false;
            const uint8_t prev = sizes[ofs - 1];
            if (!prev) return false;
            for (uint32_t e = ofs + len; ofs < e;) sizes[ofs++] = prev;
        } else {
            return false;
        }
    }
    return build_model(m, used, sizes.data());
}

}  // namespace crn2

// Test hook: prefix_lengths of the prefix code with the given code sizes.
extern "C" int crn2_prefix_lengths(const uint8_t* code_sizes, uint32_t n, uint8_t* out)
{
    if (!code_sizes || !n || n > crn2::kMaxSyms || !out) return CRN2_EINVAL;
    crn2::HostModel m;
    if (!crn2::build_model(m, n, code_sizes)) return CRN2_EINVAL;
    crn2::DevModel dm;
    crn2::to_device_model(m, dm);
    crn2::prefix_lengths(dm, out);
    return CRN2_OK;
}
