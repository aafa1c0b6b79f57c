// crnd_gpu.hip -- CRN->DXTn transcoder on gfx950.
//
// Reference: inc/crn_decomp.h (crnd::crn_unpacker).  A level of a .crn file is
// one prefix-coded stream (unpack_dxt1 and friends, crn_decomp.h:3944-4223):
// which code decodes the next symbol depends on reference codes decoded
// earlier, and the odd row of each row pair takes its references from the
// even row above, so where a block starts is only known by following the
// stream from its first bit.  The pipeline keeps that serial part as short as
// it can be and runs everything else in parallel:
//   1. crnd_tables_kernel (all bit positions in parallel): for every bit
//      position, the bits eight even-row block pairs (with their odd-row
//      flags), one pair, or two odd-row blocks (four variants by the two
//      blocks' reference flags) would take if they started there; built in
//      LDS per tile from per-position code lengths.
//   2. crnd_walk_kernel (two waves per level): one wave follows the stream
//      through the composed tables, one lookup per eight pairs or two odd-row
//      blocks, in scalar registers; the other streams the tables into an LDS
//      ring ahead of it.  Records the start bit of every group of 8 pairs /
//      8 odd-row blocks.
//   3. crnd_sym_kernel (one thread per group; even rows, then odd rows):
//      endpoint deltas, selectors and reference codes from the start bits.
//   4. crnd_index_{chunk,carry}_kernel: running endpoint indices by pointer
//      jumping over the blocks' parent links (previous block or block above)
//      in LDS per chunk of rows, then chunk after chunk per level.
//   5. crnd_index_final_kernel (one thread per block): each block's indices
//      from its chunk's resolved row, the range checks, and for a visible
//      block the palette gather and the 8- or 16-byte DXTn block write.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <vector>

#include "crn2_internal.h"
#include "crnd_host.h"

using namespace crn2;

namespace {

enum : uint32_t { M_REF = 0, M_EP0 = 1, M_SEL0 = 2, M_EP1 = 3, M_SEL1 = 4, M_COUNT = 5 };

// kind 0 = DXT1 (colour), 1 = DXT5 family (alpha + colour), 2 = DXN (two
// alpha channels), 3 = DXT5A (one alpha channel)
static int format_kind(uint32_t f)
{
    switch (f) {
    case CRN2_FMT_DXT1: return 0;
    case CRN2_FMT_DXT5: case CRN2_FMT_DXT5_CCXY: case CRN2_FMT_DXT5_XGXR: case CRN2_FMT_DXT5_XGBR:
    case CRN2_FMT_DXT5_AGBR: return 1;
    case CRN2_FMT_DXN_XY: case CRN2_FMT_DXN_YX: return 2;
    case CRN2_FMT_DXT5A: return 3;
    }
    return -1;
}

static bool kind_uses(int kind, uint32_t m)
{
    if (m == M_REF) return true;
    if (m == M_EP0 || m == M_SEL0) return kind == 0 || kind == 1;
    return kind != 0;
}
// ------------------------------------------------------------ shared types
// Per-texture constants of a launch.
struct TexInfo {
    const DevModel* models;      // M_COUNT models (device)
    const uint16_t* sorted;      // M_COUNT x kMaxSyms canonical symbol orders
    // Per 11-bit prefix (crnd_tables_kernel): cw1 = GL | F << 8 | DL0 << 16 |
    // SL0 << 24, cw2 = DL1 | SL1 << 8 (second channel of kinds 1 and 2).  GL
    // is 0 when the group code is longer than the prefix; a DL / SL field is
    // the code length that every window with the prefix starts (0: no code
    // matches) or 0xFF where the windows differ, which the canonical walk
    // then decides.
    const uint32_t* cw1;
    const uint16_t* cw2;
    int32_t* status;             // set to CRN2_ECORRUPT on an out-of-range palette index
    uint32_t kind, n_ce, n_cs, n_ae, n_as;
};

// One level stream (all faces of one level of one texture).
struct LevelJob {
    const uint8_t* src;          // level bytes; the 4-byte word holding src[0] and every word up to
                                 // src + src_size are readable
    uint32_t src_size;
    uint32_t tex;                // index into the TexInfo array
    uint32_t gw, gh, faces;      // even-padded block grid
    uint32_t tlen;               // table entries: multiple of kTabPerWG, >= 8 * src_size + 1024
    uint32_t wg_trans;           // first workgroup of the job in crnd_tables_kernel
    uint32_t wg_sym;             // first workgroup of the job in crnd_sym_kernel<false> (even rows)
    uint32_t wg_odd;             // first workgroup of the job in crnd_sym_kernel<true> (odd rows)
    uint32_t* O2;                // two odd-row blocks, one entry per bit position
    uint32_t* E8;                // eight even-row pairs, one entry per bit position
    uint16_t* X;                 // one even-row pair, one entry per bit position
    uint32_t* posE;              // start bit of each group of kGroup pairs of the even rows
    uint32_t* posO;              // start bit of each group of kGroup blocks of the odd rows
    uint8_t* gvals;              // faces * gh/2 * gw/2 reference group symbols
    uint64_t* pre;               // faces * gh * gw: decoded symbols of each block
    uint32_t crows;              // block rows per index chunk
    uint32_t nchunks;            // index chunks of the level
    uint32_t wg_idx;             // first workgroup of the job in crnd_index_chunk_kernel
    uint32_t need_x;             // 1: the even rows end in a partial group of pairs, the only place the
                                 // walk reads X; 0: X is neither written nor read
    uint64_t blk_first;          // first block of the job in crnd_index_final_kernel
    uint64_t* ent;               // faces * gh * gw: (parent, index residues) after the chunk pass
    uint64_t* bnd;               // nchunks * gw: the chunks' last rows (resolved in place by crnd_index_carry_kernel)
};

// The four facts the transition tables need of a reference group symbol G
// (four 2-bit reference codes: the pair's two blocks in bits 0-1 and 4-5, the
// blocks below them in bits 2-3 and 6-7): bit 0 / bit 1 -- the pair's first /
// second block carries endpoint deltas (its code is 0); bit 2 / bit 3 -- the
// block below the first / second one does.
__host__ __device__ inline uint32_t group_flags(uint32_t g)
{
    return (uint32_t)((g & 0x03u) == 0) | ((uint32_t)((g & 0x30u) == 0) << 1) | ((uint32_t)((g & 0x0Cu) == 0) << 2) |
           ((uint32_t)((g & 0xC0u) == 0) << 3);
}

__device__ inline bool kind_uses_dev(uint32_t kind, uint32_t m)
{
    if (m == M_REF) return true;
    if (m == M_EP0 || m == M_SEL0) return kind == 0 || kind == 1;
    return kind != 0;
}

// Models of the endpoint-delta and selector symbols of one block, in stream
// order (crn_decomp.h:3944-4223): colour then alpha for DXT5, two alpha
// channels for DXN, one alpha channel for DXT5A.
__device__ inline uint32_t d_model0(uint32_t kind) { return kind <= 1 ? M_EP0 : M_EP1; }
__device__ inline uint32_t s_model0(uint32_t kind) { return kind <= 1 ? M_SEL0 : M_SEL1; }
__device__ inline bool two_channels(uint32_t kind) { return kind == 1 || kind == 2; }

struct alignas(16) LdsTables {
    uint32_t lut[kLutSize];
    uint32_t limit[17];
    int32_t base[17];
    uint32_t nsyms;
    uint32_t pad;
};

__device__ inline void stage_tables(LdsTables& t, const DevModel& dm, uint32_t tid, uint32_t nthreads)
{
    typedef unsigned int u4 __attribute__((ext_vector_type(4)));
    for (uint32_t i = tid; i < kLutSize / 4; i += nthreads)
        reinterpret_cast<u4*>(t.lut)[i] = reinterpret_cast<const u4*>(dm.lut)[i];
    if (tid < 17) {
        t.limit[tid] = dm.limit[tid];
        t.base[tid] = dm.base[tid];
    }
    if (tid == 0) t.nsyms = dm.num_syms;
}

// Prefix decode of the 16 bits `top` (MSB first): the direct table covers
// codes up to kLutBits bits, longer codes take the canonical walk.  A bit
// pattern that matches no code gives symbol 0 and consumes nothing
// (crn_decomp.h:3239-3244).
typedef __attribute__((address_space(1))) const uint16_t gcuint16;

__device__ inline uint32_t decode_top(const LdsTables& t, gcuint16* sorted, uint32_t top, uint32_t& len)
{
    const uint32_t e = t.lut[top >> (16 - kLutBits)];
    if (e) {
        len = e >> 16;
        return e & 0xFFFFu;
    }
    const uint32_t k = top + 1;
    uint32_t L = kLutBits + 1;
    while (L <= 16 && k > t.limit[L]) L++;
    len = 0;
    if (L > 16) return 0;
    const uint32_t idx = (uint32_t)(t.base[L] + (int32_t)(top >> (16 - L)));
    if (idx >= t.nsyms) return 0;
    len = L;
    return sorted[idx];
}

// ------------------------------------------------ stage 1: transition tables
// For every bit position q of a level stream (bits at or past the end read as
// zero, so every entry from the end of the stream on is the same; the tables
// run at least 1024 entries past it):
//   O2[q], byte f (f = 0..3): bits taken by two odd-row blocks starting at q
//          when bit 0 of f says the first carries endpoint deltas and bit 1
//          the second
//   E8[q], bits 0-10: bits taken by eight even-row pairs starting at q
//          bits 11-26: the odd-row flags of those pairs, two per pair (the
//          block below the pair's first / second block has reference code 0)
//   X[q],  bits 0-7: bits taken by one even-row pair; bits 8-9: its two flags
//          (only for a level whose even rows end in a partial group of
//          pairs, LevelJob::need_x: the walk reads X nowhere else)
// One workgroup computes kTabPerWG consecutive positions in three LDS passes:
//   1. per position: the reference group symbol and its code length, and the
//      code lengths of one block's endpoint deltas and of its selectors
//      (CW = GL | F << 8 | DL << 16 | SL << 24; two channels: both codes;
//      F = group_flags(G));
//   2. per position: the one-unit transitions from CW -- bits of an even-row
//      pair (group symbol, then per block the deltas when its reference code
//      is 0, then the selectors) and of an odd-row block without / with deltas
//      (T = pair | odd flags << 8 | odd0 << 10 in 16 bits, odd1 in 8: a pair
//      takes at most 144 bits, a block 64 (32 without deltas), a code 16);
//   3. the composed entries above from T (eight pairs reach 1008 entries on).
// Each pass runs past the tile by what the next one reads.  Every thread
// keeps several positions' lookup chains in flight at once (kTabU1..3, the
// positions tid + 512 k), so that the dependent LDS reads of one position
// overlap those of the others; the pass-1 code tables and the T array share
// LDS (the tables are dead once pass 1 ends), 50,816 B per workgroup: three
// workgroups per CU, two beside a stream walker's 40 KiB ring.  The tile of
// 6,144 positions keeps the halo that passes 1 and 2 recompute to 19 % and
// 17 % (measured on the 4096^2 pair: 4,096 positions 6.04 ms per plan of 32
// copies, 5,120 6.02, 6,144 5.75, 8,192 6.03 at four waves per SIMD).
constexpr uint32_t kTabPerWG = 6144;
constexpr uint32_t kTabT = kTabPerWG + 1024;
constexpr uint32_t kTabCW = kTabT + 160;
constexpr uint32_t kTabWords = kTabCW / 32 + 3;
constexpr uint32_t kTabThreads = 512;
constexpr uint32_t kTabU1 = 8;   // pass 1: two rounds of 8 slots cover kTabCW (7,328 = 14.3 x 512)
constexpr uint32_t kTabU2 = 7;   // pass 2: two rounds of 7 cover kTabT
constexpr uint32_t kTabU3 = 6;   // pass 3b: two rounds of 6 cover kTabPerWG
constexpr uint32_t kTabE2 = kTabPerWG + 6 * 144;   // two-pair entries an eight-pair chain reads
constexpr uint32_t kTabU3a = 7;  // pass 3a: two rounds of 7 cover kTabE2
static_assert(2 * kTabU1 * kTabThreads >= kTabCW && 2 * kTabU2 * kTabThreads == kTabT &&
                  2 * kTabU3 * kTabThreads == kTabPerWG && 2 * kTabU3a * kTabThreads >= kTabE2 &&
                  kTabE2 + 144 <= kTabT && kTabE2 * 2 <= kTabCW * 4,
              "table pass rounds");

__device__ inline uint32_t be_bytes_word(const uint8_t* src, uint32_t size, uint32_t b)
{
    uint32_t v = 0;
    for (uint32_t k = 0; k < 4; k++) {
        v <<= 8;
        if (b + k < size) v |= src[b + k];
    }
    return v;
}

// The 32 stream bits from bit r on (MSB first).
__device__ inline uint32_t peek32(const uint32_t* w, uint32_t r)
{
    const uint32_t i = r >> 5;
    const uint64_t v = ((uint64_t)w[i] << 32) | w[i + 1];
    return (uint32_t)((v << (r & 31)) >> 32);
}

// Canonical-walk constants of one model for codes longer than kLutBits, with
// limit[L] for L > kLutBits made non-decreasing (an empty length takes the
// largest limit below it: the canonical walk passes over it all the same), so
// that the code length is kLutBits + 1 plus the number of limits the window
// exceeds -- five broadcast LDS reads and no loop.
struct alignas(16) CodeLimits {
    uint32_t limit[17];
    int32_t base[17];
    uint32_t nsyms;
    uint32_t pad;
};

// Length of the code in the 16-bit window `top` whose 11-bit prefix is no
// complete code (0: no code matches, crn_decomp.h:3239-3244), and its index
// into the canonical order.
__device__ __attribute__((always_inline)) inline uint32_t long_code(const CodeLimits& t, uint32_t top, uint32_t& idx)
{
    const uint32_t k = top + 1;
    uint32_t L = kLutBits + 1;
#pragma unroll
    for (uint32_t i = kLutBits + 1; i <= 16; i++) L += k > t.limit[i] ? 1u : 0u;
    if (L > 16) return 0;
    idx = (uint32_t)(t.base[L] + (int32_t)(top >> (16 - L)));
    return idx < t.nsyms ? L : 0u;
}


typedef __attribute__((address_space(1))) uint32_t guint32;
typedef __attribute__((address_space(1))) uint16_t guint16;

constexpr uint32_t kTabRefSyms = 1024;   // reference-model symbols (low bytes) staged in LDS; the rest read from memory

struct TabCodes {
    uint32_t cw1[kLutSize];   // TexInfo::cw1
    uint16_t cw2[kLutSize];   // TexInfo::cw2
    CodeLimits lim[M_COUNT];
    uint8_t ref_sym[kTabRefSyms];
    uint32_t words[kTabWords];
};

union TabShared {
    TabCodes c;          // pass 1
    struct {
        uint16_t lo[kTabT];   // passes 2 and 3: pair bits | odd flags << 8 | bits of an odd block without deltas << 10
        uint8_t hi[kTabT];    //                 bits of an odd block with deltas
    } t;
};

// tab_job[wg] = the level job of tables workgroup wg.
__global__ void __launch_bounds__(kTabThreads) crnd_tables_kernel(const LevelJob* __restrict__ jobs,
                                                                  const uint32_t* __restrict__ tab_job,
                                                                  const TexInfo* __restrict__ texs)
{
    __shared__ TabShared s_u;
    __shared__ uint32_t s_cw[kTabCW];
    const uint32_t tid = threadIdx.x;
    const LevelJob& j = jobs[tab_job[blockIdx.x]];
    const TexInfo& tx = texs[j.tex];
    const uint32_t kind = tx.kind;
    const uint32_t P0 = (blockIdx.x - j.wg_trans) * kTabPerWG;

    {
        typedef unsigned int u4 __attribute__((ext_vector_type(4)));
        const bool two = two_channels(kind);
        for (uint32_t i = tid; i < kLutSize / 4; i += kTabThreads)
            reinterpret_cast<u4*>(s_u.c.cw1)[i] = reinterpret_cast<const u4*>(tx.cw1)[i];
        if (two)
            for (uint32_t i = tid; i < kLutSize / 8; i += kTabThreads)
                reinterpret_cast<u4*>(s_u.c.cw2)[i] = reinterpret_cast<const u4*>(tx.cw2)[i];
        if (tid < 17 * M_COUNT) {
            const uint32_t m = tid / 17, L = tid % 17;
            const DevModel& dm = tx.models[m];
            uint32_t lim = dm.limit[L];
            for (uint32_t i = kLutBits + 1; i < L; i++) lim = max(lim, dm.limit[i]);
            s_u.c.lim[m].limit[L] = lim;
            s_u.c.lim[m].base[L] = dm.base[L];
        }
        if (tid < M_COUNT) s_u.c.lim[tid].nsyms = tx.models[tid].num_syms;
        for (uint32_t i = tid; i < kTabRefSyms; i += kTabThreads)
            s_u.c.ref_sym[i] = (uint8_t)tx.sorted[M_REF * kMaxSyms + i];
    }
    for (uint32_t i = tid; i < kTabWords; i += kTabThreads)
        s_u.c.words[i] = be_bytes_word(j.src, j.src_size, P0 / 8 + 4 * i);
    __syncthreads();

    // 1. code lengths: one combined-table read per position (two for two
    // channels), the canonical walk only for codes longer than kLutBits
    {
        const uint32_t* W = s_u.c.words;
        const CodeLimits& R = s_u.c.lim[M_REF];
        const CodeLimits& D0 = s_u.c.lim[d_model0(kind)];
        const CodeLimits& S0 = s_u.c.lim[s_model0(kind)];
        const CodeLimits& D1 = s_u.c.lim[M_EP1];
        const CodeLimits& S1 = s_u.c.lim[M_SEL1];
        const bool two = two_channels(kind);
        const uint16_t* ref_sorted = tx.sorted + M_REF * kMaxSyms;
        constexpr uint32_t U = kTabU1;
#pragma unroll 1
        for (uint32_t i0 = 0; i0 < 2 * U; i0 += U) {
            uint32_t r[U], w32[U], c[U];   // w32: the 32 bits from r on; top = w32 >> 16
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                r[u] = min(tid + (i0 + u) * kTabThreads, kTabCW - 1);
                w32[u] = peek32(W, r[u]);
            }
#pragma unroll
            for (uint32_t u = 0; u < U; u++) c[u] = s_u.c.cw1[w32[u] >> (32 - kLutBits)];
            // the canonical walk only for a long group code or a prefix whose
            // windows start delta / selector codes of different lengths (at most
            // five prefixes per model: where a length's codes end off a
            // 32-window boundary)
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                if (__builtin_expect((c[u] & 0xFFu) == 0 || (c[u] & 0xFF0000u) == 0xFF0000u || (c[u] >> 24) == 0xFFu,
                                     0)) {
                    uint32_t gl = c[u] & 0xFFu, g = (c[u] >> 8) & 0xFFu, dl = (c[u] >> 16) & 0xFFu, sl = c[u] >> 24;
                    uint32_t idx = 0;
                    if (!gl) {
                        gl = long_code(R, w32[u] >> 16, idx);
                        g = group_flags(!gl ? 0u : idx < kTabRefSyms ? (uint32_t)s_u.c.ref_sym[idx]
                                                                     : (uint32_t)(ref_sorted[idx] & 0xFFu));
                    }
                    if (dl == 0xFFu) dl = long_code(D0, w32[u] >> 16, idx);
                    if (sl == 0xFFu) sl = long_code(S0, w32[u] >> 16, idx);
                    c[u] = gl | (g << 8) | (dl << 16) | (sl << 24);
                }
            }
            if (two) {
                uint32_t ta[U], tb[U], a[U], b[U];
#pragma unroll
                for (uint32_t u = 0; u < U; u++) {
                    // a first-channel code takes at most 16 bits: the second
                    // channel's window lies inside the 32 bits read at r
                    ta[u] = (w32[u] << ((c[u] >> 16) & 0xFFu)) >> 16;
                    tb[u] = (w32[u] << (c[u] >> 24)) >> 16;
                }
#pragma unroll
                for (uint32_t u = 0; u < U; u++) {
                    a[u] = s_u.c.cw2[ta[u] >> (16 - kLutBits)] & 0xFFu;
                    b[u] = s_u.c.cw2[tb[u] >> (16 - kLutBits)] >> 8;
                }
#pragma unroll
                for (uint32_t u = 0; u < U; u++) {
                    uint32_t idx = 0;
                    if (__builtin_expect(a[u] == 0xFFu, 0)) a[u] = long_code(D1, ta[u], idx);
                    if (__builtin_expect(b[u] == 0xFFu, 0)) b[u] = long_code(S1, tb[u], idx);
                    c[u] += (a[u] << 16) + (b[u] << 24);
                }
            }
#pragma unroll
            for (uint32_t u = 0; u < U; u++)
                if (tid + (i0 + u) * kTabThreads < kTabCW) s_cw[r[u]] = c[u];
        }
    }
    __syncthreads();

    // 2. one-unit transitions (branch-free: a block without deltas re-reads
    // the entry it already has)
    {
        uint16_t* T = s_u.t.lo;
        uint8_t* T8 = s_u.t.hi;
        constexpr uint32_t U = kTabU2;
#pragma unroll 1
        for (uint32_t i0 = 0; i0 < 2 * U; i0 += U) {
            uint32_t r[U], c[U], e[U], x[U], od[U];
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                r[u] = tid + (i0 + u) * kTabThreads;
                c[u] = s_cw[r[u]];
            }
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                e[u] = r[u] + (c[u] & 0xFFu);
                x[u] = s_cw[e[u]];
                od[u] = s_cw[r[u] + ((c[u] >> 16) & 0xFFu)];
            }
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                if (c[u] & 0x100u) e[u] += (x[u] >> 16) & 0xFFu;
                x[u] = s_cw[e[u]];
            }
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                e[u] += x[u] >> 24;
                x[u] = s_cw[e[u]];
            }
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                if (c[u] & 0x200u) e[u] += (x[u] >> 16) & 0xFFu;
                x[u] = s_cw[e[u]];
            }
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                e[u] += x[u] >> 24;
                const uint32_t d = (c[u] >> 16) & 0xFFu;
                T[r[u]] = (uint16_t)((e[u] - r[u]) | ((c[u] >> 2) & 0x300u) | ((c[u] >> 24) << 10));
                T8[r[u]] = (uint8_t)(d + (od[u] >> 24));
            }
        }
    }
    __syncthreads();

    // 3. composed entries (stores at 32-bit offsets from the tile's base).
    // 3a: O2 and X, and two-pair entries E2 = len (9 bits) | four odd flags
    // << 9 for every position an E8 chain reaches (kTabPerWG + 6 x 144), into
    // the CW array's space (dead after pass 2); 3b: E8 from four E2 lookups
    // (three dependent LDS reads instead of seven).
    {
        const uint16_t* T = s_u.t.lo;
        const uint8_t* T8 = s_u.t.hi;
        uint16_t* E2 = reinterpret_cast<uint16_t*>(s_cw);
        typedef __attribute__((address_space(1))) uint8_t guint8;
        guint8* O2 = (guint8*)(j.O2 + P0);
        guint8* E8 = (guint8*)(j.E8 + P0);
        guint8* X = (guint8*)(j.X + P0);
        const bool need_x = j.need_x != 0;
        {
            constexpr uint32_t U = kTabU3a;
#pragma unroll 1
            for (uint32_t i0 = 0; i0 < 2 * U; i0 += U) {
                uint32_t r[U], t0[U], o1[U], ta[U], tb[U], ta8[U], tb8[U], tp[U];
#pragma unroll
                for (uint32_t u = 0; u < U; u++) {
                    r[u] = min(tid + (i0 + u) * kTabThreads, kTabE2 - 1);
                    t0[u] = T[r[u]];
                    o1[u] = T8[r[u]];
                }
#pragma unroll
                for (uint32_t u = 0; u < U; u++) {
                    ta[u] = T[r[u] + (t0[u] >> 10)];
                    ta8[u] = T8[r[u] + (t0[u] >> 10)];
                    tb[u] = T[r[u] + o1[u]];
                    tb8[u] = T8[r[u] + o1[u]];
                    tp[u] = T[r[u] + (t0[u] & 0xFFu)];
                }
#pragma unroll
                for (uint32_t u = 0; u < U; u++) {
                    const uint32_t q = tid + (i0 + u) * kTabThreads;
                    if (q < kTabE2)
                        E2[r[u]] = (uint16_t)(((t0[u] & 0xFFu) + (tp[u] & 0xFFu)) | ((t0[u] & 0x300u) << 1) |
                                              ((tp[u] & 0x300u) << 3));
                    if (q < kTabPerWG) {
                        const uint32_t o0 = t0[u] >> 10;
                        *(guint32*)(O2 + 4 * r[u]) = (o0 + (ta[u] >> 10)) | ((o1[u] + (tb[u] >> 10)) << 8) |
                                                     ((o0 + ta8[u]) << 16) | ((o1[u] + tb8[u]) << 24);
                        if (need_x) *(guint16*)(X + 2 * r[u]) = (uint16_t)(t0[u] & 0x3FFu);
                    }
                }
            }
        }
        __syncthreads();
        {
            constexpr uint32_t U = kTabU3;
#pragma unroll 1
            for (uint32_t i0 = 0; i0 < 2 * U; i0 += U) {
                uint32_t r[U], p[U], fl[U];
#pragma unroll
                for (uint32_t u = 0; u < U; u++) {
                    r[u] = tid + (i0 + u) * kTabThreads;
                    p[u] = r[u];
                    fl[u] = 0;
                }
                // odd flags gathered at bit 9 + 4 i, shifted down once
#pragma unroll
                for (uint32_t i = 0; i < 4; i++) {
#pragma unroll
                    for (uint32_t u = 0; u < U; u++) {
                        const uint32_t e = E2[p[u]];
                        fl[u] |= (e & 0x1E00u) << (4 * i);
                        p[u] += e & 0x1FFu;
                    }
                }
#pragma unroll
                for (uint32_t u = 0; u < U; u++)
                    *(guint32*)(E8 + 4 * r[u]) = (p[u] - r[u]) | ((fl[u] >> 9) << 11);
            }
        }
    }
}

// --------------------------------------------------------- stage 2: walker
// One workgroup of two waves per level stream.  Wave 0 follows the stream
// from bit 0 and records the start bit of every group of kGroup pairs (even
// rows) or kGroup blocks (odd rows); wave 1 streams the composed tables from
// memory into an LDS ring ahead of it, so the walking wave never waits on
// memory and holds no staging registers.
//
// A lone wave issues about one instruction every four cycles, so the walking
// loop is kept short and wave-uniform (scalar registers):
//   even row: one E8 lookup per group of eight pairs (a broadcast LDS read),
//             one X lookup per pair in a row's last, partial group;
//   odd row:  one O2 lookup per two blocks, read with v_readlane from the 64
//             ring entries at `base` held one per lane in a vector register,
//             the next 64 prefetched into a second one.
// The walker tells the producer which segment it has reached at the start of
// a group; the producer never overwrites that segment or a later unread one,
// and the walker waits (bounded by the producer's progress) only when the ring
// does not yet hold kNeed entries past the group's start.
constexpr uint32_t kSeg = 1024;                       // table entries per ring segment
// The ring holds RS segments (template parameter) and the producer loads RS / 2
// at a time: 8 (80 KiB of LDS, one walker per CU, the deepest prefetch: the
// shortest walk of a lone level) or 4 (40 KiB, three walkers per CU: more
// levels at once when a plan holds hundreds of them).
constexpr uint32_t kPlanJobsSmallRing = 512;
constexpr uint32_t kGroup = 8;                        // pairs / blocks per recorded start bit
constexpr uint32_t kMaxEGroups = 64;                  // even-row groups per row (one per lane): grids up to 1024 blocks wide
constexpr uint32_t kNeed = 1152 + 128 + 64;           // a group's longest move + the two windows
// Either wave waits for the other at most this many sleeps (about a quarter
// of a second; a real wait lasts microseconds).  Past it the producer stops
// and the walker finishes its loops on whatever the ring holds and reports
// CRN2_EHIP, so a fault in the hand-off cannot hang the device.
constexpr uint32_t kSpinCap = 1u << 21;

template <uint32_t RS>
struct WalkLds {
    static constexpr uint32_t kRingPos = kSeg * RS;
    uint32_t o2[kRingPos];
    uint32_t e8[kRingPos];
    uint16_t x[kRingPos];
    uint32_t ready;       // segments [0, ready) are in the ring
    uint32_t consumed;    // the walker reads no segment below this one any more
    uint32_t done;        // the walker has finished
};

typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const u32x4 gconst_u32x4;

// Two-bit flag pairs x = f0 | f1 << 2 | f2 << 4 | f3 << 6 -> bytes f_k * 8.
__device__ inline uint32_t spread_flags(uint32_t x)
{
    return ((x & 3u) | ((x & 0xCu) << 6) | ((x & 0x30u) << 12) | ((x & 0xC0u) << 18)) << 3;
}

__device__ inline uint32_t uniform(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ inline uint32_t lds_get(const uint32_t* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }
__device__ inline void lds_put(uint32_t* p, uint32_t v) { __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); }

// Per lane and segment: 4 + 4 vectors of O2 / E8 entries, 2 of X entries.
template <uint32_t RS>
struct SegVecs {
    u32x4 v[RS / 2][10];
};

// Segments past the table reload its last one: every entry from the end of
// the stream on is the same, and the table's last segment lies past the end.
template <uint32_t RS>
__device__ inline void producer_load(const LevelJob& j, uint32_t tsegs, uint32_t s0, uint32_t lane, SegVecs<RS>& r)
{
    constexpr uint32_t kBatch = RS / 2;
#pragma unroll
    for (uint32_t b = 0; b < kBatch; b++) {
        const size_t at = (size_t)min(s0 + b, tsegs - 1) * kSeg;
        const gconst_u32x4* po = (const gconst_u32x4*)(j.O2 + at);
        const gconst_u32x4* pe = (const gconst_u32x4*)(j.E8 + at);
        const gconst_u32x4* px = (const gconst_u32x4*)(j.X + at);
#pragma unroll
        for (uint32_t i = 0; i < 4; i++) {
            r.v[b][i] = po[64 * i + lane];
            r.v[b][4 + i] = pe[64 * i + lane];
        }
        if (j.need_x) {
            r.v[b][8] = px[lane];
            r.v[b][9] = px[64 + lane];
        }
    }
}

template <uint32_t RS>
__device__ inline void producer_store(WalkLds<RS>& sh, uint32_t s0, uint32_t lane, const SegVecs<RS>& r, bool need_x)
{
    constexpr uint32_t kBatch = RS / 2, kRingSegs = RS;
#pragma unroll
    for (uint32_t b = 0; b < kBatch; b++) {
        const uint32_t at = ((s0 + b) % kRingSegs) * kSeg;
        u32x4* po = reinterpret_cast<u32x4*>(sh.o2 + at);
        u32x4* pe = reinterpret_cast<u32x4*>(sh.e8 + at);
        u32x4* px = reinterpret_cast<u32x4*>(sh.x + at);
#pragma unroll
        for (uint32_t i = 0; i < 4; i++) {
            po[64 * i + lane] = r.v[b][i];
            pe[64 * i + lane] = r.v[b][4 + i];
        }
        if (need_x) {
            px[lane] = r.v[b][8];
            px[64 + lane] = r.v[b][9];
        }
    }
}

// Wave 1: loads batch n + 1 while batch n waits for ring space; stops when
// the walker says it is done.
template <uint32_t RS>
__device__ void walk_producer(const LevelJob& j, WalkLds<RS>& sh, uint32_t lane)
{
    constexpr uint32_t kBatch = RS / 2, kRingSegs = RS;
    const uint32_t tsegs = j.tlen / kSeg;
    const bool need_x = j.need_x != 0;
    SegVecs<RS> r;
    uint32_t s = 0;
    producer_load(j, tsegs, s, lane, r);
    for (;;) {
#pragma unroll 1
        for (uint32_t spin = 0;; spin++) {
            if (uniform(lds_get(&sh.done))) return;
            if (s + kBatch <= uniform(lds_get(&sh.consumed)) + kRingSegs) break;
            if (spin >= kSpinCap) return;
            __builtin_amdgcn_s_sleep(2);
        }
        producer_store(sh, s, lane, r, need_x);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        lds_put(&sh.ready, s + kBatch);
        s += kBatch;
        producer_load(j, tsegs, s, lane, r);
    }
}

// Wave 0.
template <uint32_t RS>
__device__ void walk_stream(const LevelJob& j, WalkLds<RS>& sh, uint32_t lane, int32_t* status)
{
    constexpr uint32_t kRingPos = kSeg * RS;
    constexpr uint32_t kMask = kRingPos - 1;
    uint32_t lim = 0;                 // ring holds table entries [.., lim)
    bool stuck = false;
#pragma unroll 1
        for (uint32_t spin = 0;; spin++) {
        lim = uniform(lds_get(&sh.ready)) * kSeg;
        if (lim >= kNeed) break;
        if (spin >= kSpinCap) {
            stuck = true;
            lim = 0xFFFFFFFFu;
            break;
        }
        __builtin_amdgcn_s_sleep(1);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    uint32_t pub = 0;                 // segment last reported to the producer
    uint32_t vready = lim / kSeg;     // ring fill level read at the last report (vector register, read ahead)
    uint32_t chk = min(kSeg, lim - kNeed);   // group starts below chk need no ring check
    // Group start at bit `pos`: report a new segment to the producer, take the
    // fill level read at the previous report, read the next one ahead, and
    // wait only if the ring lacks entries the group may need.
    auto ring_check = [&](uint32_t pos) {
        const uint32_t seg = pos / kSeg;
        if (seg != pub) {
            pub = seg;
            lds_put(&sh.consumed, seg);
            lim = max(lim, uniform(vready) * kSeg);
            vready = lds_get(&sh.ready);
            __asm__ volatile("" ::: "memory");
        }
        if (pos + kNeed > lim) {
#pragma unroll 1
        for (uint32_t spin = 0;; spin++) {
                lim = uniform(lds_get(&sh.ready)) * kSeg;
                if (pos + kNeed <= lim) break;
                if (spin >= kSpinCap) {
                    stuck = true;
                    lim = 0xFFFFFFFFu;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
        }
        chk = stuck ? 0xFFFFFFFFu : min((seg + 1) * kSeg, lim - kNeed);
    };

    const uint32_t gw = j.gw, gh = j.gh, faces = j.faces;
    const uint32_t npairs = gw >> 1;
    const uint32_t gE = (npairs + kGroup - 1) / kGroup;   // even-row groups per row (at most 64)
    const uint32_t gO = (gw + kGroup - 1) / kGroup;       // odd-row groups per row (at most 128)
    guint32* posE = (guint32*)j.posE;
    guint32* posO = (guint32*)j.posO;
    uint32_t p = 0;                   // current bit
    const uint8_t* o2b = reinterpret_cast<const uint8_t*>(sh.o2 + lane);
    uint32_t rowpair = 0;
    for (uint32_t f = 0; f < faces; f++) {
        for (uint32_t yp = 0; yp < gh; yp += 2, rowpair++) {
            // even row: one E8 lookup per full group of eight pairs; lane g
            // keeps group g's start bit (vE) and odd-row flags (vflags)
            uint32_t vE = 0, vflags = 0;
            uint32_t u0 = 0, g = 0;
            for (; u0 + kGroup <= npairs; u0 += kGroup, g++) {
                if (__builtin_expect(p >= chk, 0)) ring_check(p);
                const bool mine = lane == g;
                vE = mine ? p : vE;
                const uint32_t e = uniform(sh.e8[p & kMask]);
                p += e & 0x7FFu;
                vflags = mine ? (e >> 11) : vflags;
            }
            if (u0 < npairs) {
                if (__builtin_expect(p >= chk, 0)) ring_check(p);
                vE = (lane == g) ? p : vE;
                uint32_t fw = 0;
                for (uint32_t k = 0; u0 + k < npairs; k++) {
                    const uint32_t x = uniform(sh.x[p & kMask]);
                    fw |= (x >> 8) << (2 * k);
                    p += x & 0xFFu;
                }
                vflags = (lane == g) ? fw : vflags;
            }
            if (lane < gE) posE[rowpair * gE + lane] = vE;
            // odd row: one O2 lookup per two blocks from the lane windows.  Lane
            // g holds, for odd groups 2g (vFlo) and 2g + 1 (vFhi), byte k =
            // (flags of the step's two blocks) * 8: the entry's byte offset.
            const uint32_t vFlo = spread_flags(vflags & 0xFFu), vFhi = spread_flags((vflags >> 8) & 0xFFu);
            uint32_t base = p & ~63u, o = p & 63u;
            uint32_t rpb = ((base + 64) & kMask) * 4;   // ring byte offset of the prefetched window
            uint32_t vcur = sh.o2[(base & kMask) + lane];
            uint32_t vnxt = *reinterpret_cast<const uint32_t*>(o2b + rpb);
            // Both windows in place before the loop, so that the loop's own
            // waits cover only the prefetch issued by a window advance.
            __builtin_amdgcn_s_waitcnt(0xC07F);   // lgkmcnt(0)
#define CRND_O2(FO, K)                                                                  \
    do {                                                                                \
        const uint32_t e_ = (uint32_t)__builtin_amdgcn_readlane((int)vcur, (int)o);     \
        o += (e_ >> (((FO) >> (8 * (K))) & 31u)) & 0xFFu;                               \
        if (o >= 64) {                                                                  \
            do {                                                                        \
                base += 64;                                                             \
                o -= 64;                                                                \
                vcur = vnxt;                                                            \
                rpb = (rpb + 256) & (4 * kRingPos - 1);                                 \
                vnxt = *reinterpret_cast<const uint32_t*>(o2b + rpb);                   \
            } while (o >= 64);                                                          \
        }                                                                               \
    } while (0)
#define CRND_OGROUP_START(VF, VO, H)                                                    \
    const uint32_t pos_ = base + o;                                                     \
    if (__builtin_expect(pos_ >= chk, 0)) ring_check(pos_);                                                  \
    VO = (lane == (H)) ? pos_ : VO;                                                     \
    const uint32_t fo_ = (uint32_t)__builtin_amdgcn_readlane((int)(VF), (int)(H))
            uint32_t vOa = 0, vOb = 0;   // lane h: start bits of odd groups 2h, 2h + 1
            const uint32_t n16 = gw / 16;
            uint32_t h = 0;
            for (; h < n16; h++) {
                {
                    CRND_OGROUP_START(vFlo, vOa, h);
                    CRND_O2(fo_, 0);
                    CRND_O2(fo_, 1);
                    CRND_O2(fo_, 2);
                    CRND_O2(fo_, 3);
                }
                {
                    CRND_OGROUP_START(vFhi, vOb, h);
                    CRND_O2(fo_, 0);
                    CRND_O2(fo_, 1);
                    CRND_O2(fo_, 2);
                    CRND_O2(fo_, 3);
                }
            }
            uint32_t rem = gw - 16 * n16;    // blocks left: 0..14, even
            if (rem >= kGroup) {
                CRND_OGROUP_START(vFlo, vOa, h);
                CRND_O2(fo_, 0);
                CRND_O2(fo_, 1);
                CRND_O2(fo_, 2);
                CRND_O2(fo_, 3);
                rem -= kGroup;
                if (rem) {
                    CRND_OGROUP_START(vFhi, vOb, h);
                    for (uint32_t k = 0; 2 * k < rem; k++) CRND_O2(fo_, k);
                }
            } else if (rem) {
                CRND_OGROUP_START(vFlo, vOa, h);
                for (uint32_t k = 0; 2 * k < rem; k++) CRND_O2(fo_, k);
            }
#undef CRND_OGROUP_START
#undef CRND_O2
            if (2 * lane < gO) posO[rowpair * gO + 2 * lane] = vOa;
            if (2 * lane + 1 < gO) posO[rowpair * gO + 2 * lane + 1] = vOb;
            p = base + o;
        }
    }
    lds_put(&sh.done, 1u);
    if (stuck && lane == 0) atomicExch(status, CRN2_EHIP);
}

template <uint32_t RS>
__global__ void __launch_bounds__(128) crnd_walk_kernel(const LevelJob* __restrict__ jobs,
                                                        const TexInfo* __restrict__ texs)
{
    __shared__ __align__(16) WalkLds<RS> sh;
    const LevelJob& j = jobs[blockIdx.x];
    if (threadIdx.x == 0) {
        sh.ready = 0;
        sh.consumed = 0;
        sh.done = 0;
    }
    __syncthreads();
    if (threadIdx.x >= 64) walk_producer(j, sh, threadIdx.x - 64);
    else walk_stream(j, sh, threadIdx.x, texs[j.tex].status);
}

// ------------------------------------------------------ stage 3: symbols
// One thread per recorded group decodes, from the group's start bit, the
// endpoint deltas and selector symbols of its pairs (even rows: reference group
// symbol first, kept for the odd row) or blocks (odd rows: reference codes from
// the even row's group symbols).  Record per block (uint64): bits 0-1 reference
// code, 2-15 first endpoint delta, 16-29 second endpoint delta, 30-43 first
// selector, 44-57 second selector.

// MSB-first reader starting at any bit; bytes at or past `end` read as zero
// (crn_decomp.h:3196-3207).  The stream is fetched as aligned 16-byte vectors
// kept three deep in registers (the current one and two ahead) with global
// (not flat) loads, so a refill rarely waits on memory and no LDS wait waits
// for a fetch.  A vector is loaded only if it starts before `end`; level
// streams lie in buffers padded by at least 64 bytes past their last 16-byte
// boundary.
struct WordBits {
    typedef unsigned int v4 __attribute__((ext_vector_type(4)));
    const uint8_t* vnext;   // byte address of the next vector to fetch
    const uint8_t* end;
    v4 q0, q1, q2;
    const uint8_t* wa;      // byte address of the next word to take
    uint32_t qi;            // its index in q0
    uint64_t buf;
    uint32_t cnt;

    __device__ inline v4 fetch()
    {
        typedef __attribute__((address_space(1))) const v4 gv4;
        v4 v = {0u, 0u, 0u, 0u};
        if (vnext < end) v = *(gv4*)vnext;
        vnext += 16;
        return v;
    }

    __device__ inline uint32_t next_word()
    {
        uint32_t v = qi == 0 ? q0.x : qi == 1 ? q0.y : qi == 2 ? q0.z : q0.w;
        v = __builtin_bswap32(v);
        if (wa >= end) v = 0;
        else if (end - wa < 4) v &= 0xFFFFFFFFu << (8 * (4 - (uint32_t)(end - wa)));
        wa += 4;
        if (++qi == 4) {
            qi = 0;
            q0 = q1;
            q1 = q2;
            q2 = fetch();
        }
        return v;
    }

    __device__ inline void start(const uint8_t* src, uint32_t size, uint32_t bit)
    {
        end = src + size;
        const uintptr_t ab = reinterpret_cast<uintptr_t>(src) + (bit >> 3);
        vnext = reinterpret_cast<const uint8_t*>(ab & ~(uintptr_t)15);
        wa = reinterpret_cast<const uint8_t*>(ab & ~(uintptr_t)3);
        qi = (uint32_t)(ab >> 2) & 3u;
        q0 = fetch();
        q1 = fetch();
        q2 = fetch();
        const uint32_t skip = (uint32_t)(ab & 3) * 8 + (bit & 7);
        const uint32_t w0 = next_word();
        const uint32_t w1 = next_word();
        buf = (((uint64_t)w0 << 32) | w1) << skip;
        cnt = 64 - skip;
    }

    __device__ inline uint32_t decode(const LdsTables& t, gcuint16* sorted)
    {
        if (cnt <= 32) {
            buf |= (uint64_t)next_word() << (32 - cnt);
            cnt += 32;
        }
        uint32_t len;
        const uint32_t sym = decode_top(t, sorted, (uint32_t)(buf >> 48), len);
        buf <<= len;
        cnt -= len;
        return sym;
    }
};

constexpr uint32_t kSymGroupsPerWG = 256;

struct SymCtx {
    const LdsTables* t;
    gcuint16* so;
    uint32_t dm0, sm0;
    bool two;

    __device__ inline uint64_t block(WordBits& b, uint32_t ref) const
    {
        uint32_t d0 = 0, d1 = 0, s1 = 0;
        if (ref == 0) {
            d0 = b.decode(t[dm0], so + dm0 * kMaxSyms);
            if (two) d1 = b.decode(t[M_EP1], so + M_EP1 * kMaxSyms);
        }
        const uint32_t s0 = b.decode(t[sm0], so + sm0 * kMaxSyms);
        if (two) s1 = b.decode(t[M_SEL1], so + M_SEL1 * kMaxSyms);
        return (uint64_t)ref | ((uint64_t)d0 << 2) | ((uint64_t)d1 << 16) | ((uint64_t)s0 << 30) |
               ((uint64_t)s1 << 44);
    }
};

template <bool ODD>
__global__ void __launch_bounds__(256) crnd_sym_kernel(const LevelJob* __restrict__ jobs, uint32_t njobs,
                                                       const TexInfo* __restrict__ texs)
{
    __shared__ LdsTables s_t[M_COUNT];
    const uint32_t tid = threadIdx.x;
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((ODD ? jobs[mid].wg_odd : jobs[mid].wg_sym) <= blockIdx.x) lo = mid;
        else hi = mid;
    }
    const LevelJob& j = jobs[lo];
    const TexInfo& tx = texs[j.tex];
    const uint32_t kind = tx.kind;
    for (uint32_t m = 0; m < M_COUNT; m++)
        if (kind_uses_dev(kind, m)) stage_tables(s_t[m], tx.models[m], tid, 256);
    __syncthreads();

    gcuint16* gsorted = (gcuint16*)tx.sorted;
    SymCtx sc{s_t, gsorted, d_model0(kind), s_model0(kind), two_channels(kind)};
    const uint32_t gw = j.gw, np = gw >> 1;
    const uint32_t rowpairs = j.faces * (j.gh >> 1);
    const uint32_t gpr = ODD ? (gw + kGroup - 1) / kGroup : (np + kGroup - 1) / kGroup;   // groups per row
    // odd rows: one thread per two consecutive groups of a row, decoded in one
    // run from the first group's start bit (the second starts where it ends)
    const uint32_t tpr = ODD ? (gpr + 1) / 2 : gpr;   // threads per row pair
    const uint32_t it = (blockIdx.x - (ODD ? j.wg_odd : j.wg_sym)) * kSymGroupsPerWG + tid;
    if (it >= rowpairs * tpr) return;
    const uint32_t rp = it / tpr, grp = ODD ? 2 * (it - rp * tpr) : it - rp * tpr;
    // A group's records stay in registers (the unrolled loops index them
    // statically) and are stored once at its end: vmcnt counts stores too, so
    // a store inside the loop would make the next stream refill wait for it.
    typedef __attribute__((address_space(1))) uint64_t guint64;
    typedef __attribute__((address_space(1))) uint8_t guint8;
    WordBits b;
    uint64_t rec[2 * kGroup];   // eight pairs, or two groups of eight odd-row blocks
    if (!ODD) {
        b.start(j.src, j.src_size, j.posE[it]);
        const uint32_t u0 = grp * kGroup, nu = min(np - u0, kGroup);
        uint32_t g[kGroup];
#pragma unroll
        for (uint32_t k = 0; k < kGroup; k++) {
            if (k < nu) {
                g[k] = b.decode(s_t[M_REF], gsorted + M_REF * kMaxSyms) & 0xFFu;
                rec[2 * k] = sc.block(b, g[k] & 3u);
                rec[2 * k + 1] = sc.block(b, (g[k] >> 4) & 3u);
            }
        }
        guint64* pre = (guint64*)(j.pre + (uint64_t)(2 * rp) * gw + 2 * u0);
        guint8* gv = (guint8*)(j.gvals + (uint64_t)rp * np + u0);
#pragma unroll
        for (uint32_t k = 0; k < kGroup; k++) {
            if (k < nu) {
                pre[2 * k] = rec[2 * k];
                pre[2 * k + 1] = rec[2 * k + 1];
                gv[k] = (uint8_t)g[k];
            }
        }
    } else {
        b.start(j.src, j.src_size, j.posO[rp * gpr + grp]);
        const uint32_t x0 = grp * kGroup, nx = min(gw - x0, 2 * kGroup);
        const guint8* gv = (const guint8*)(j.gvals + (uint64_t)rp * np + (x0 >> 1));
        uint32_t g[kGroup];
#pragma unroll
        for (uint32_t k = 0; k < kGroup; k++) g[k] = 2 * k < nx ? (uint32_t)gv[k] : 0u;
#pragma unroll
        for (uint32_t k = 0; k < 2 * kGroup; k++)
            if (k < nx) rec[k] = sc.block(b, (g[k >> 1] >> (2 + 4 * (k & 1u))) & 3u);
        guint64* pre = (guint64*)(j.pre + (uint64_t)(2 * rp + 1) * gw + x0);
#pragma unroll
        for (uint32_t k = 0; k < 2 * kGroup; k++)
            if (k < nx) pre[k] = rec[k];
    }
}

// ------------------------------------------------- stage 4: palette indices
// Running endpoint indices (crn_decomp.h:3944-4223): reference code 0 adds the
// decoded delta to the running index with one conditional wrap, 1 repeats it,
// 2 and 3 take the index of the block above; the running index carries across
// rows and faces, and the block above row 0 holds index 0.  Every block's
// index is therefore its parent's plus a constant (mod the palette size): the
// parent is the previous block in raster order (codes 0, 1) or the block above
// (2, 3), the root is index 0.  Pointer jumping resolves that forest:
//   crnd_index_chunk_kernel: per chunk of whole rows (at most kIdxChunk
//       blocks) in LDS, until every block points at the root or outside the
//       chunk -- which is always the chunk's row above, the last row of the
//       previous chunk; that last row is also copied to a compact array;
//   crnd_index_carry_kernel: one workgroup per level of more than one chunk
//       walks its chunks in order, resolving each chunk's last row from the
//       resolved last row of the chunk before it (held in LDS; the rows of
//       the next few chunks are loaded ahead);
//   crnd_index_final_kernel: one thread per block adds its chunk-level result
//       to the resolved row it points at, checks ranges and writes its record.
// With residues kept mod n, a block's index equals the reference's whenever no
// earlier block went out of range; the first that does (previous index +
// delta >= 2n after the one wrap, or a selector index past its palette) sees
// exact values, so the corruption check is exact too.  Records always hold
// in-range indices, so the block write never reads past a palette.
constexpr uint32_t kIdxChunk = 4096;
constexpr uint32_t kRoot = 0xFFFFFFFFu;

struct IdxSizes {
    uint32_t n0, n1, ns0, ns1;
    bool two;
};

__device__ inline IdxSizes idx_sizes(const TexInfo& tx)
{
    IdxSizes z;
    const uint32_t kind = tx.kind;
    z.two = two_channels(kind);
    z.n0 = kind <= 1 ? tx.n_ce : tx.n_ae;
    z.n1 = z.two ? tx.n_ae : 1u;
    z.ns0 = kind <= 1 ? tx.n_cs : tx.n_as;
    z.ns1 = z.two ? tx.n_as : 1u;
    return z;
}

__device__ inline uint32_t residue(uint32_t d, uint32_t n) { return n ? d % n : 0u; }

// Sum of two (index0, index1) residue pairs, 16 bits each.
__device__ inline uint32_t add_res(uint32_t a, uint32_t b, uint32_t n0, uint32_t n1)
{
    uint32_t x = (a & 0xFFFFu) + (b & 0xFFFFu);
    if (x >= n0) x -= n0;
    uint32_t y = (a >> 16) + (b >> 16);
    if (y >= n1) y -= n1;
    return x | (y << 16);
}

__device__ inline uint64_t jump(uint64_t e, uint64_t ep, const IdxSizes& z)
{
    return (ep & 0xFFFFFFFFull) | ((uint64_t)add_res((uint32_t)(e >> 32), (uint32_t)(ep >> 32), z.n0, z.n1) << 32);
}

constexpr uint32_t kIdxThreads = 1024;
constexpr uint32_t kIdxPer = kIdxChunk / kIdxThreads;   // entries per thread, in flight together

__global__ void __launch_bounds__(kIdxThreads) crnd_index_chunk_kernel(const LevelJob* __restrict__ jobs, uint32_t njobs,
                                                                       const TexInfo* __restrict__ texs)
{
    __shared__ uint64_t L[2][kIdxChunk];
    const uint32_t tid = threadIdx.x;
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].wg_idx <= blockIdx.x) lo = mid;
        else hi = mid;
    }
    const LevelJob& j = jobs[lo];
    const IdxSizes z = idx_sizes(texs[j.tex]);
    const uint32_t gw = j.gw, rows = j.faces * j.gh;
    const uint32_t c = blockIdx.x - j.wg_idx;
    const uint32_t r0 = c * j.crows, r1 = min(rows, r0 + j.crows);
    const uint32_t b0 = r0 * gw, nb = (r1 - r0) * gw;
#pragma unroll
    for (uint32_t k = 0; k < kIdxPer; k++) {
        const uint32_t i = tid + k * kIdxThreads;
        if (i >= nb) continue;
        const uint32_t b = b0 + i;
        const uint64_t q = j.pre[b];
        const uint32_t ref = (uint32_t)q & 3u;
        uint32_t par, off = 0;
        if (ref <= 1) {
            par = b ? b - 1 : kRoot;
            if (ref == 0)
                off = residue((uint32_t)(q >> 2) & 0x3FFFu, z.n0) |
                      (z.two ? residue((uint32_t)(q >> 16) & 0x3FFFu, z.n1) << 16 : 0u);
        } else {
            par = b >= gw ? b - gw : kRoot;
        }
        L[0][i] = par | ((uint64_t)off << 32);
    }
    __syncthreads();
    uint32_t cur = 0;
    for (uint32_t span = 1; span < nb; span <<= 1) {
        uint64_t e[kIdxPer], ep[kIdxPer];
#pragma unroll
        for (uint32_t k = 0; k < kIdxPer; k++) {
            const uint32_t i = min(tid + k * kIdxThreads, nb - 1);
            e[k] = L[cur][i];
        }
#pragma unroll
        for (uint32_t k = 0; k < kIdxPer; k++) {
            const uint32_t pa = (uint32_t)e[k];
            ep[k] = (pa != kRoot && pa >= b0) ? L[cur][pa - b0] : 0ull;
        }
#pragma unroll
        for (uint32_t k = 0; k < kIdxPer; k++) {
            const uint32_t i = tid + k * kIdxThreads;
            const uint32_t pa = (uint32_t)e[k];
            if (i < nb) L[cur ^ 1][i] = (pa != kRoot && pa >= b0) ? jump(e[k], ep[k], z) : e[k];
        }
        __syncthreads();
        cur ^= 1;
    }
    const uint32_t last = nb - gw;
#pragma unroll
    for (uint32_t k = 0; k < kIdxPer; k++) {
        const uint32_t i = tid + k * kIdxThreads;
        if (i >= nb) continue;
        const uint64_t e = L[cur][i];
        j.ent[b0 + i] = e;
        if (i >= last) {
            const uint32_t pa = (uint32_t)e;
            const uint32_t cp = pa == kRoot ? kRoot : (c - 1) * gw + (pa - (r0 - 1) * gw);
            j.bnd[(uint64_t)c * gw + (i - last)] = cp | (e & 0xFFFFFFFF00000000ull);
        }
    }
}

// The job holding thread t of a launch over consecutive ranges (first(job) =
// the range's first thread): one workgroup-uniform binary search for the
// workgroup's first thread (scalar loads), then a short forward scan per
// thread for workgroups that straddle jobs.
template <typename J, typename F>
__device__ inline uint32_t find_job_wg(const J* jobs, uint32_t njobs, uint64_t t, F first)
{
    const uint64_t t0 = (uint64_t)blockIdx.x * blockDim.x;
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (first(jobs[mid]) <= t0) lo = mid;
        else hi = mid;
    }
    while (lo + 1 < njobs && first(jobs[lo + 1]) <= t) lo++;
    return lo;
}

constexpr uint32_t kMaxIdxGw = 1024;       // widest padded grid a plan accepts
static_assert(2 * kGroup * kMaxEGroups <= kMaxIdxGw, "index carry row width");
constexpr uint32_t kCarryThreads = 256;
constexpr uint32_t kCarryPer = kMaxIdxGw / kCarryThreads;
constexpr uint32_t kCarryAhead = 4;        // chunks whose last rows are loaded ahead

// carry[wg] = a level job of more than one chunk.  Chunk 0's last row is
// resolved already (it points at the root); chunk c's entries point into
// chunk c - 1's last row.  Resolved rows are written back in place as
// (root, residues).
__global__ void __launch_bounds__(kCarryThreads) crnd_index_carry_kernel(const LevelJob* __restrict__ jobs,
                                                                         const uint32_t* __restrict__ carry,
                                                                         const TexInfo* __restrict__ texs)
{
    typedef __attribute__((address_space(1))) uint64_t guint64;
    __shared__ uint32_t V[2][kMaxIdxGw];
    const uint32_t tid = threadIdx.x;
    const LevelJob& j = jobs[carry[blockIdx.x]];
    const IdxSizes z = idx_sizes(texs[j.tex]);
    const uint32_t gw = j.gw, n = j.nchunks;
    guint64* bnd = (guint64*)j.bnd;
    for (uint32_t i = tid; i < gw; i += kCarryThreads) V[0][i] = (uint32_t)(bnd[i] >> 32);
    uint64_t pre[kCarryAhead][kCarryPer];
#pragma unroll
    for (uint32_t a = 0; a < kCarryAhead; a++)
#pragma unroll
        for (uint32_t k = 0; k < kCarryPer; k++) {
            const uint32_t c = 1 + a, i = tid + k * kCarryThreads;
            pre[a][k] = (c < n && i < gw) ? bnd[(uint64_t)c * gw + i] : 0ull;
        }
    __syncthreads();
    uint32_t cur = 0;
    for (uint32_t c0 = 1; c0 < n; c0 += kCarryAhead) {
#pragma unroll
        for (uint32_t a = 0; a < kCarryAhead; a++) {
            const uint32_t c = c0 + a;
            if (c < n) {
                const uint32_t cn = c + kCarryAhead;
#pragma unroll
                for (uint32_t k = 0; k < kCarryPer; k++) {
                    const uint32_t i = tid + k * kCarryThreads;
                    const uint64_t e = pre[a][k];
                    pre[a][k] = (cn < n && i < gw) ? bnd[(uint64_t)cn * gw + i] : 0ull;
                    if (i < gw) {
                        const uint32_t pa = (uint32_t)e;
                        uint32_t v = (uint32_t)(e >> 32);
                        if (pa != kRoot) v = add_res(v, V[cur][pa - (c - 1) * gw], z.n0, z.n1);
                        V[cur ^ 1u][i] = v;
                        bnd[(uint64_t)c * gw + i] = ((uint64_t)v << 32) | kRoot;
                    }
                }
            }
            __syncthreads();
            cur ^= 1u;
        }
    }
}

struct WriteJob {
    uint8_t* dst[6];        // per-face destination
    uint32_t pitch;         // bytes between block rows
    uint32_t bx, by;        // visible blocks
    uint32_t gw, gh;        // padded grid
    uint32_t faces;
    uint32_t pal;           // palette set index
};

struct WritePalettes {
    const uint32_t* ce;
    const uint32_t* cs;
    const uint16_t* ae;
    const uint16_t* as;     // 3 per entry
    uint32_t kind;
    uint32_t pad;
};

// One DXTn block from its palette indices (lo: first channel's endpoint and
// selector indices, 16 bits each; hi: the second channel's) into row `out`.
__device__ inline void write_block(const WritePalettes& P, uint8_t* out, uint32_t x, uint32_t pitch, uint32_t lo32,
                                   uint32_t hi32)
{
    if (P.kind == 0 || P.kind == 3) {
        uint2 o;
        if (P.kind == 0) {
            o.x = P.ce[lo32 & 0xFFFFu];
            o.y = P.cs[lo32 >> 16];
        } else {
            const uint16_t* s = P.as + 3 * (lo32 >> 16);
            o.x = (uint32_t)P.ae[lo32 & 0xFFFFu] | ((uint32_t)s[0] << 16);
            o.y = (uint32_t)s[1] | ((uint32_t)s[2] << 16);
        }
        if ((pitch & 7) == 0) {
            reinterpret_cast<uint2*>(out)[x] = o;
        } else {
            uint32_t* w = reinterpret_cast<uint32_t*>(out + 8 * (uint64_t)x);
            w[0] = o.x;
            w[1] = o.y;
        }
    } else {
        uint4 o;
        if (P.kind == 1) {
            const uint16_t* s = P.as + 3 * (hi32 >> 16);
            o.x = (uint32_t)P.ae[hi32 & 0xFFFFu] | ((uint32_t)s[0] << 16);
            o.y = (uint32_t)s[1] | ((uint32_t)s[2] << 16);
            o.z = P.ce[lo32 & 0xFFFFu];
            o.w = P.cs[lo32 >> 16];
        } else {
            const uint16_t* s0 = P.as + 3 * (lo32 >> 16);
            const uint16_t* s1 = P.as + 3 * (hi32 >> 16);
            o.x = (uint32_t)P.ae[lo32 & 0xFFFFu] | ((uint32_t)s0[0] << 16);
            o.y = (uint32_t)s0[1] | ((uint32_t)s0[2] << 16);
            o.z = (uint32_t)P.ae[hi32 & 0xFFFFu] | ((uint32_t)s1[0] << 16);
            o.w = (uint32_t)s1[1] | ((uint32_t)s1[2] << 16);
        }
        if ((pitch & 15) == 0) {
            reinterpret_cast<uint4*>(out)[x] = o;
        } else {
            uint32_t* w = reinterpret_cast<uint32_t*>(out + 16 * (uint64_t)x);
            w[0] = o.x;
            w[1] = o.y;
            w[2] = o.z;
            w[3] = o.w;
        }
    }
}

// Also stage 5: a visible block's DXTn bytes go straight to the caller's
// buffer (wjobs[i] belongs to jobs[i]).
__global__ void __launch_bounds__(256) crnd_index_final_kernel(const LevelJob* __restrict__ jobs, uint32_t njobs,
                                                               const TexInfo* __restrict__ texs, uint64_t total,
                                                               const WriteJob* __restrict__ wjobs,
                                                               const WritePalettes* __restrict__ pals)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const uint32_t ji = find_job_wg(jobs, njobs, t, [](const LevelJob& x) { return x.blk_first; });
    const LevelJob& j = jobs[ji];
    const TexInfo& tx = texs[j.tex];
    const IdxSizes z = idx_sizes(tx);
    const uint32_t b = (uint32_t)(t - j.blk_first);
    const uint32_t gw = j.gw;
    const uint64_t* res = j.bnd;
    auto value = [&](uint32_t blk) -> uint32_t {
        const uint64_t e = j.ent[blk];
        const uint32_t pa = (uint32_t)e;
        uint32_t v = (uint32_t)(e >> 32);
        if (pa != kRoot) {
            const uint32_t r = pa / gw;
            v = add_res(v, (uint32_t)(res[(uint64_t)(r / j.crows) * gw + (pa - r * gw)] >> 32), z.n0, z.n1);
        }
        return v;
    };
    const uint64_t q = j.pre[b];
    const uint32_t ref = (uint32_t)q & 3u;
    const uint32_t v = value(b);
    // the previous block's indices: the previous lane's v (the same job unless
    // b == 0, where they are 0); lane 0 looks them up
    uint32_t pv = (uint32_t)__shfl_up((int)v, 1);
    bool bad = z.n0 == 0 || z.n1 == 0;
    if (ref == 0) {
        if (b == 0) pv = 0u;
        else if ((threadIdx.x & 63u) == 0) pv = value(b - 1);
        bad |= (pv & 0xFFFFu) + ((uint32_t)(q >> 2) & 0x3FFFu) >= 2 * z.n0;
        if (z.two) bad |= (pv >> 16) + ((uint32_t)(q >> 16) & 0x3FFFu) >= 2 * z.n1;
    }
    uint32_t s0 = (uint32_t)(q >> 30) & 0x3FFFu, s1 = (uint32_t)(q >> 44) & 0x3FFFu;
    if (s0 >= z.ns0) {
        bad = true;
        s0 = 0;
    }
    if (z.two && s1 >= z.ns1) {
        bad = true;
        s1 = 0;
    }
    if (bad) atomicExch(tx.status, CRN2_ECORRUPT);
    const WriteJob& w = wjobs[ji];
    const uint32_t per_face = gw * j.gh;
    const uint32_t f = b / per_face, rem = b - f * per_face, y = rem / gw, x = rem - y * gw;
    if (x < w.bx && y < w.by)
        write_block(pals[w.pal], w.dst[f] + (uint64_t)y * w.pitch, x, w.pitch, (v & 0xFFFFu) | (s0 << 16),
                    (v >> 16) | (s1 << 16));
}


// ----------------------------------------------------------------- profiling
constexpr int kStages = 6;   // tables, walk, sym (even rows), sym (odd rows), index chunks + carry, index final + write
struct Profile {
    std::mutex mu;
    bool on = false;
    struct Ev { hipEvent_t e[kStages + 1]; };
    std::vector<Ev> pending;
    double ms[kStages] = {};
    uint64_t n[kStages] = {};
};
static Profile g_prof;

static void profile_collect_locked()
{
    for (auto& ev : g_prof.pending) {
        (void)hipEventSynchronize(ev.e[kStages]);
        for (int s = 0; s < kStages; s++) {
            float t = 0;
            if (hipEventElapsedTime(&t, ev.e[s], ev.e[s + 1]) == hipSuccess) {
                g_prof.ms[s] += t;
                g_prof.n[s]++;
            }
        }
        for (int s = 0; s <= kStages; s++) (void)hipEventDestroy(ev.e[s]);
    }
    g_prof.pending.clear();
}

}  // namespace

// ------------------------------------------------------------------ context
struct crn2_unpack_ctx {
    uint32_t magic = 0;
    const uint8_t* data = nullptr;
    uint32_t size = 0;
    Header h{};
    int kind = -1;
    HostModel models[M_COUNT];
    uint8_t* d_file = nullptr;   // device copy padded by 64 bytes
    DevModel* d_models = nullptr;
    uint16_t* d_sorted = nullptr;
    uint32_t* d_cw1 = nullptr;   // TexInfo::cw1
    uint16_t* d_cw2 = nullptr;   // TexInfo::cw2
    uint32_t* d_ce = nullptr;
    uint32_t* d_cs = nullptr;
    uint16_t* d_ae = nullptr;
    uint16_t* d_as = nullptr;
    int32_t* d_status = nullptr;
    uint32_t n_ce = 0, n_cs = 0, n_ae = 0, n_as = 0;
    crn2_batch* cached = nullptr;   // plan of the last crn2_unpack_levels_device / _level_device call
    void* cached_dst = nullptr;
    uint32_t cached_first = 0, cached_count = 0, cached_pitch = 0;
};

// A prepared launch over one or more contexts.
struct crn2_batch {
    uint32_t magic = 0;
    LevelJob* d_jobs = nullptr;
    TexInfo* d_texs = nullptr;
    WriteJob* d_wjobs = nullptr;
    WritePalettes* d_pals = nullptr;
    uint32_t* d_tab_job = nullptr;  // level job of every crnd_tables_kernel workgroup
    uint8_t* d_scratch = nullptr;   // transition tables, start bits, symbol records
    uint32_t njobs = 0;
    uint32_t wg_trans = 0, wg_sym = 0, wg_odd = 0, wg_idx = 0;
    uint32_t max_gw = 0;
    uint32_t* d_carry = nullptr;                // level jobs of more than one index chunk (crnd_index_carry_kernel)
    uint32_t ncarry = 0;
    uint64_t blk_total = 0;
    std::vector<int32_t*> status;   // per context
    hipGraphExec_t graph = nullptr; // the plan's launches captured once (crn2_batch_run_graph)
};

namespace {

constexpr uint32_t kCtxMagic = 0x1EF9CABDu;
constexpr uint32_t kBatchMagic = 0x7B47C2A1u;

static uint32_t level_dim(uint32_t v, uint32_t l) { return std::max<uint32_t>(1u, v >> l); }

static uint32_t bytes_per_block(uint32_t fmt)
{
    return (fmt == CRN2_FMT_DXT1 || fmt == CRN2_FMT_DXT5A) ? 8u : 16u;
}

#define HIPCHK(x)                                   \
    do {                                            \
        if ((x) != hipSuccess) return CRN2_EHIP;    \
    } while (0)

// Palette decoding on the host inside unpack_begin (crn_decomp.h:3715-3851).
static bool decode_palettes(crn2_unpack_ctx* c, std::vector<uint32_t>& ce, std::vector<uint32_t>& cs,
                            std::vector<uint16_t>& ae, std::vector<uint16_t>& as)
{
    const Header& h = c->h;
    HostBits b;
    if (h.pal_num[0]) {
        const uint32_t n = h.pal_num[0];
        if (!h.pal_size[0] || (uint64_t)h.pal_ofs[0] + h.pal_size[0] > c->size) return false;
        b.start(c->data + h.pal_ofs[0], h.pal_size[0]);
        HostModel dm[2];
        if (!receive_model(b, dm[0]) || !receive_model(b, dm[1]) || !dm[0].valid || !dm[1].valid) return false;
        ce.resize(n);
        uint32_t r0 = 0, g0 = 0, b0 = 0, r1 = 0, g1 = 0, b1 = 0;
        for (uint32_t i = 0; i < n; i++) {
            r0 = (r0 + b.decode(dm[0])) & 31;
            g0 = (g0 + b.decode(dm[1])) & 63;
            b0 = (b0 + b.decode(dm[0])) & 31;
            r1 = (r1 + b.decode(dm[0])) & 31;
            g1 = (g1 + b.decode(dm[1])) & 63;
            b1 = (b1 + b.decode(dm[0])) & 31;
            ce[i] = b0 | (g0 << 5) | (r0 << 11) | (b1 << 16) | (g1 << 21) | (r1 << 27);
        }
        const uint32_t ns = h.pal_num[1];
        cs.assign(ns ? ns : 1, 0);
        if (ns) {
            if (!h.pal_size[1] || (uint64_t)h.pal_ofs[1] + h.pal_size[1] > c->size) return false;
            b.start(c->data + h.pal_ofs[1], h.pal_size[1]);
            HostModel sm;
            if (!receive_model(b, sm) || !sm.valid) return false;
            uint32_t s = 0;
            for (uint32_t i = 0; i < ns; i++) {
                for (uint32_t k = 0; k < 32; k += 4) s ^= b.decode(sm) << k;
                // linear 2-bit selectors (0 = first colour .. 3 = second) to DXT1 codes
                cs[i] = ((s ^ (s << 1)) & 0xAAAAAAAAu) | ((s >> 1) & 0x55555555u);
            }
        }
    }
    if (h.pal_num[2]) {
        const uint32_t n = h.pal_num[2];
        if (!h.pal_size[2] || (uint64_t)h.pal_ofs[2] + h.pal_size[2] > c->size) return false;
        b.start(c->data + h.pal_ofs[2], h.pal_size[2]);
        HostModel dm;
        if (!receive_model(b, dm) || !dm.valid) return false;
        ae.resize(n);
        uint32_t a = 0, bb = 0;
        for (uint32_t i = 0; i < n; i++) {
            a = (a + b.decode(dm)) & 255;
            bb = (bb + b.decode(dm)) & 255;
            ae[i] = (uint16_t)(a | (bb << 8));
        }
        const uint32_t ns = h.pal_num[3];
        as.assign(ns ? 3 * ns : 3, 0);
        if (ns) {
            if (!h.pal_size[3] || (uint64_t)h.pal_ofs[3] + h.pal_size[3] > c->size) return false;
            b.start(c->data + h.pal_ofs[3], h.pal_size[3]);
            HostModel sm;
            if (!receive_model(b, sm) || !sm.valid) return false;
            static const uint8_t from_linear[8] = {0, 2, 3, 4, 5, 6, 7, 1};
            uint32_t l0 = 0, l1 = 0;
            for (uint32_t i = 0; i < ns; i++) {
                uint32_t r0 = 0, r1 = 0;
                for (uint32_t k = 0; k < 24; k += 6) l0 ^= b.decode(sm) << k;
                for (uint32_t k = 0; k < 24; k += 3) r0 |= (uint32_t)from_linear[(l0 >> k) & 7] << k;
                for (uint32_t k = 0; k < 24; k += 6) l1 ^= b.decode(sm) << k;
                for (uint32_t k = 0; k < 24; k += 3) r1 |= (uint32_t)from_linear[(l1 >> k) & 7] << k;
                as[3 * i] = (uint16_t)r0;
                as[3 * i + 1] = (uint16_t)((r0 >> 16) | (r1 << 8));
                as[3 * i + 2] = (uint16_t)(r1 >> 8);
            }
        }
    }
    return true;
}

static void batch_free(crn2_batch* b)
{
    if (!b) return;
    b->magic = 0;
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    void* ptrs[] = {b->d_jobs, b->d_texs, b->d_wjobs, b->d_pals, b->d_tab_job, b->d_carry, b->d_scratch};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete b;
}

static void ctx_free_device(crn2_unpack_ctx* c)
{
    batch_free(c->cached);
    c->cached = nullptr;
    void* ptrs[] = {c->d_file, c->d_models, c->d_sorted, c->d_cw1, c->d_cw2, c->d_ce, c->d_cs, c->d_ae, c->d_as, c->d_status};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
}

template <typename T>
static hipError_t dmalloc(T** p, size_t bytes)
{
    return hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(bytes, 16));
}

static crn2_unpack_ctx* begin_impl(const void* data, const void* data_dev, uint32_t size, hipStream_t stream)
{
    if (!data || size < 62) return nullptr;
    crn2_unpack_ctx* c = new (std::nothrow) crn2_unpack_ctx();
    if (!c) return nullptr;
    c->data = static_cast<const uint8_t*>(data);
    c->size = size;
    auto fail = [&]() -> crn2_unpack_ctx* {
        ctx_free_device(c);
        delete c;
        return nullptr;
    };
    if (!parse_header(c->data, size, c->h)) return fail();
    c->kind = format_kind(c->h.format);
    if (c->kind < 0) return fail();
    const Header& h = c->h;
    if (h.levels < 1 || h.levels > kMaxLevels || (h.faces != 1 && h.faces != 6) || !h.width || !h.height) return fail();
    // tables: crn_decomp.h:3662-3692
    if (!h.tables_size || (uint64_t)h.tables_ofs + h.tables_size > size) return fail();
    HostBits b;
    b.start(c->data + h.tables_ofs, h.tables_size);
    if (!receive_model(b, c->models[M_REF])) return fail();
    if (!h.pal_num[0] && !h.pal_num[2]) return fail();
    if (h.pal_num[0] && (!receive_model(b, c->models[M_EP0]) || !receive_model(b, c->models[M_SEL0]))) return fail();
    if (h.pal_num[2] && (!receive_model(b, c->models[M_EP1]) || !receive_model(b, c->models[M_SEL1]))) return fail();
    for (uint32_t m = 0; m < M_COUNT; m++)
        if (kind_uses(c->kind, m) && !c->models[m].valid) return fail();
    if ((c->kind == 0 || c->kind == 1) && (!h.pal_num[0] || !h.pal_num[1])) return fail();
    if (c->kind != 0 && (!h.pal_num[2] || !h.pal_num[3])) return fail();
    for (uint32_t l = 0; l < h.levels; l++) {
        const uint32_t next = (l + 1 < h.levels) ? h.level_ofs[l + 1] : size;
        if (h.level_ofs[l] >= next || next > size) return fail();
    }
    std::vector<uint32_t> ce, cs;
    std::vector<uint16_t> ae, as;
    if (!decode_palettes(c, ce, cs, ae, as)) return fail();
    c->n_ce = h.pal_num[0];
    c->n_cs = h.pal_num[1];
    c->n_ae = h.pal_num[2];
    c->n_as = h.pal_num[3];

    std::vector<DevModel> dms(M_COUNT);
    std::vector<uint16_t> sorted(M_COUNT * kMaxSyms, 0);
    for (uint32_t m = 0; m < M_COUNT; m++) {
        if (c->models[m].valid) {
            to_device_model(c->models[m], dms[m]);
            std::copy(c->models[m].sorted.begin(), c->models[m].sorted.end(), sorted.begin() + m * kMaxSyms);
        } else {
            memset(&dms[m], 0, sizeof(DevModel));
        }
    }
    // Per 11-bit prefix (TexInfo::cw1 / cw2): the reference code's length and
    // group flags, and per delta / selector field the length every window of
    // the prefix starts (prefix_lengths: 0xFF where the canonical walk decides).
    auto field_lengths = [&](uint32_t m, uint8_t* out) { prefix_lengths(dms[m], out); };
    std::vector<uint32_t> cw1(kLutSize);
    std::vector<uint16_t> cw2(kLutSize);
    {
        const uint32_t d0 = c->kind <= 1 ? M_EP0 : M_EP1, s0 = c->kind <= 1 ? M_SEL0 : M_SEL1;
        std::vector<uint8_t> fd0(kLutSize), fs0(kLutSize), fd1(kLutSize), fs1(kLutSize);
        field_lengths(d0, fd0.data());
        field_lengths(s0, fs0.data());
        field_lengths(M_EP1, fd1.data());
        field_lengths(M_SEL1, fs1.data());
        for (uint32_t v = 0; v < kLutSize; v++) {
            const uint32_t r = dms[M_REF].lut[v];
            cw1[v] = (r >> 16) | (group_flags(r & 0xFFu) << 8) | ((uint32_t)fd0[v] << 16) | ((uint32_t)fs0[v] << 24);
            cw2[v] = (uint16_t)(fd1[v] | (fs1[v] << 8));
        }
    }

    const size_t padded = ((size_t)size + 15) / 16 * 16 + 64;
    if (dmalloc(&c->d_file, padded) != hipSuccess || dmalloc(&c->d_models, sizeof(DevModel) * M_COUNT) != hipSuccess ||
        dmalloc(&c->d_sorted, sorted.size() * 2) != hipSuccess ||
        dmalloc(&c->d_cw1, kLutSize * 4) != hipSuccess || dmalloc(&c->d_cw2, kLutSize * 2) != hipSuccess ||
        dmalloc(&c->d_ce, std::max<size_t>(ce.size(), 1) * 4) != hipSuccess ||
        dmalloc(&c->d_cs, std::max<size_t>(cs.size(), 1) * 4) != hipSuccess ||
        dmalloc(&c->d_ae, std::max<size_t>(ae.size(), 1) * 2) != hipSuccess ||
        dmalloc(&c->d_as, std::max<size_t>(as.size(), 3) * 2) != hipSuccess || dmalloc(&c->d_status, 4) != hipSuccess)
        return fail();
    bool ok = true;
    ok &= hipMemsetAsync(c->d_file, 0, padded, stream) == hipSuccess;
    // An empty palette keeps one zero entry: records of a corrupt level still index entry 0.
    if (ce.empty()) ok &= hipMemsetAsync(c->d_ce, 0, 4, stream) == hipSuccess;
    if (cs.empty()) ok &= hipMemsetAsync(c->d_cs, 0, 4, stream) == hipSuccess;
    if (ae.empty()) ok &= hipMemsetAsync(c->d_ae, 0, 2, stream) == hipSuccess;
    if (as.empty()) ok &= hipMemsetAsync(c->d_as, 0, 6, stream) == hipSuccess;
    if (data_dev)
        ok &= hipMemcpyAsync(c->d_file, data_dev, size, hipMemcpyDeviceToDevice, stream) == hipSuccess;
    else
        ok &= hipMemcpyAsync(c->d_file, data, size, hipMemcpyHostToDevice, stream) == hipSuccess;
    ok &= hipMemcpyAsync(c->d_models, dms.data(), sizeof(DevModel) * M_COUNT, hipMemcpyHostToDevice, stream) == hipSuccess;
    ok &= hipMemcpyAsync(c->d_sorted, sorted.data(), sorted.size() * 2, hipMemcpyHostToDevice, stream) == hipSuccess;
    ok &= hipMemcpyAsync(c->d_cw1, cw1.data(), kLutSize * 4, hipMemcpyHostToDevice, stream) == hipSuccess;
    ok &= hipMemcpyAsync(c->d_cw2, cw2.data(), kLutSize * 2, hipMemcpyHostToDevice, stream) == hipSuccess;
    if (!ce.empty()) ok &= hipMemcpyAsync(c->d_ce, ce.data(), ce.size() * 4, hipMemcpyHostToDevice, stream) == hipSuccess;
    if (!cs.empty()) ok &= hipMemcpyAsync(c->d_cs, cs.data(), cs.size() * 4, hipMemcpyHostToDevice, stream) == hipSuccess;
    if (!ae.empty()) ok &= hipMemcpyAsync(c->d_ae, ae.data(), ae.size() * 2, hipMemcpyHostToDevice, stream) == hipSuccess;
    if (!as.empty()) ok &= hipMemcpyAsync(c->d_as, as.data(), as.size() * 2, hipMemcpyHostToDevice, stream) == hipSuccess;
    ok &= hipMemsetAsync(c->d_status, 0, 4, stream) == hipSuccess;
    ok &= hipStreamSynchronize(stream) == hipSuccess;
    if (!ok) return fail();
    c->magic = kCtxMagic;
    return c;
}

static bool ctx_ok(const crn2_unpack_ctx* c) { return c && c->magic == kCtxMagic; }

// One texture's part of a plan: levels [first, first + count), level source
// override (segmented files), destinations and pitch.
struct PlanPart {
    crn2_unpack_ctx* c;
    uint32_t first, count;
    const uint8_t* src_override;
    uint32_t src_size;
    std::vector<std::array<uint8_t*, 6>> dst;   // per level, per face
    std::vector<uint32_t> pitch;                // per level
};

static size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

static int build_plan(std::vector<PlanPart>& parts, crn2_batch** out)
{
    *out = nullptr;
    crn2_batch* B = new (std::nothrow) crn2_batch();
    if (!B) return CRN2_ENOMEM;
    std::vector<LevelJob> jobs;
    std::vector<TexInfo> texs;
    std::vector<WriteJob> wj;
    std::vector<WritePalettes> pals;
    size_t scratch = 0;
    struct Ofs { size_t O2, E8, X, E, O, G, pre, ent, bnd; };
    std::vector<Ofs> ofs;
    uint32_t wg_trans = 0, wg_sym = 0, wg_odd = 0, wg_idx = 0, max_gw = 0;
    uint64_t blk_total = 0;
    for (uint32_t pi = 0; pi < parts.size(); pi++) {
        PlanPart& P = parts[pi];
        crn2_unpack_ctx* c = P.c;
        const Header& h = c->h;
        TexInfo t;
        memset(&t, 0, sizeof(t));
        t.models = c->d_models;
        t.sorted = c->d_sorted;
        t.cw1 = c->d_cw1;
        t.cw2 = c->d_cw2;
        t.status = c->d_status;
        t.kind = (uint32_t)c->kind;
        t.n_ce = c->n_ce;
        t.n_cs = c->n_cs;
        t.n_ae = c->n_ae;
        t.n_as = c->n_as;
        for (uint32_t i = 0; i < P.count; i++) {
            const uint32_t l = P.first + i;
            const uint32_t bx = (level_dim(h.width, l) + 3) >> 2, by = (level_dim(h.height, l) + 3) >> 2;
            const uint32_t gw = (bx + 1) & ~1u, gh = (by + 1) & ~1u;
            LevelJob j;
            memset(&j, 0, sizeof(j));
            if (P.src_override) {
                j.src = P.src_override;
                j.src_size = P.src_size;
            } else {
                const uint32_t next = (l + 1 < h.levels) ? h.level_ofs[l + 1] : c->size;
                j.src = c->d_file + h.level_ofs[l];
                j.src_size = next - h.level_ofs[l];
            }
            if ((uint64_t)j.src_size * 8 + 4096 >= 0xFFFFFFFFull || gw > 2 * kGroup * kMaxEGroups ||
                gw > kIdxChunk || (uint64_t)gw * gh * h.faces >= 0xFFFFFFFFull) {
                batch_free(B);
                return CRN2_EUNSUPPORTED;
            }
            j.tex = (uint32_t)texs.size();
            j.gw = gw;
            j.need_x = ((gw >> 1) % kGroup) != 0 ? 1u : 0u;
            j.gh = gh;
            j.faces = h.faces;
            j.tlen = (uint32_t)align_up((size_t)j.src_size * 8 + 1024, kTabPerWG);
            j.wg_trans = wg_trans;
            wg_trans += j.tlen / kTabPerWG;
            const uint64_t rowpairs = (uint64_t)h.faces * (gh / 2);
            const uint64_t gE = rowpairs * ((gw / 2 + kGroup - 1) / kGroup), gO = rowpairs * ((gw + kGroup - 1) / kGroup);
            j.wg_sym = wg_sym;
            wg_sym += (uint32_t)((gE + kSymGroupsPerWG - 1) / kSymGroupsPerWG);
            j.wg_odd = wg_odd;
            const uint64_t tO = rowpairs * ((((gw + kGroup - 1) / kGroup) + 1) / 2);   // two odd groups per thread
            wg_odd += (uint32_t)((tO + kSymGroupsPerWG - 1) / kSymGroupsPerWG);
            const uint32_t rows = h.faces * gh;
            j.crows = kIdxChunk / gw;
            j.nchunks = (rows + j.crows - 1) / j.crows;
            j.wg_idx = wg_idx;
            wg_idx += j.nchunks;
            j.blk_first = blk_total;
            blk_total += (uint64_t)rows * gw;
            Ofs o;
            o.O2 = scratch;
            scratch = align_up(scratch + (size_t)j.tlen * 4, 256);
            o.E8 = scratch;
            scratch = align_up(scratch + (size_t)j.tlen * 4, 256);
            o.X = scratch;
            scratch = align_up(scratch + (size_t)j.tlen * 2, 256);
            o.E = scratch;
            scratch = align_up(scratch + gE * 4, 256);
            o.O = scratch;
            scratch = align_up(scratch + gO * 4, 256);
            o.G = scratch;
            scratch = align_up(scratch + rowpairs * (gw / 2), 256);
            o.pre = scratch;
            scratch = align_up(scratch + rowpairs * 2 * gw * 8, 256);
            o.ent = scratch;
            scratch = align_up(scratch + rowpairs * 2 * gw * 8, 256);
            o.bnd = scratch;
            scratch = align_up(scratch + (size_t)j.nchunks * gw * 8, 256);
            ofs.push_back(o);
            max_gw = std::max(max_gw, gw);
            jobs.push_back(j);
            WriteJob w;
            memset(&w, 0, sizeof(w));
            for (uint32_t f = 0; f < 6; f++) w.dst[f] = (f < h.faces) ? P.dst[i][f] : nullptr;
            w.pitch = P.pitch[i];
            w.bx = bx;
            w.by = by;
            w.gw = gw;
            w.gh = gh;
            w.faces = h.faces;
            w.pal = (uint32_t)pals.size();
            wj.push_back(w);
        }
        texs.push_back(t);
        WritePalettes wp{c->d_ce, c->d_cs, c->d_ae, c->d_as, (uint32_t)c->kind, 0};
        pals.push_back(wp);
        B->status.push_back(c->d_status);
    }
    B->njobs = (uint32_t)jobs.size();
    B->wg_trans = wg_trans;
    B->wg_sym = wg_sym;
    B->wg_odd = wg_odd;
    B->wg_idx = wg_idx;
    B->max_gw = max_gw;
    B->blk_total = blk_total;
    std::vector<uint32_t> carry;
    for (uint32_t i = 0; i < (uint32_t)jobs.size(); i++)
        if (jobs[i].nchunks > 1) carry.push_back(i);
    B->ncarry = (uint32_t)carry.size();
    if (dmalloc(&B->d_jobs, jobs.size() * sizeof(LevelJob)) != hipSuccess ||
        dmalloc(&B->d_texs, texs.size() * sizeof(TexInfo)) != hipSuccess ||
        dmalloc(&B->d_wjobs, wj.size() * sizeof(WriteJob)) != hipSuccess ||
        dmalloc(&B->d_pals, pals.size() * sizeof(WritePalettes)) != hipSuccess ||
        dmalloc(&B->d_tab_job, std::max<size_t>(wg_trans, 1) * sizeof(uint32_t)) != hipSuccess ||
        dmalloc(&B->d_carry, std::max<size_t>(carry.size(), 1) * sizeof(uint32_t)) != hipSuccess ||
        dmalloc(&B->d_scratch, scratch) != hipSuccess) {
        batch_free(B);
        return CRN2_ENOMEM;
    }
    for (size_t i = 0; i < jobs.size(); i++) {
        jobs[i].O2 = reinterpret_cast<uint32_t*>(B->d_scratch + ofs[i].O2);
        jobs[i].E8 = reinterpret_cast<uint32_t*>(B->d_scratch + ofs[i].E8);
        jobs[i].X = reinterpret_cast<uint16_t*>(B->d_scratch + ofs[i].X);
        jobs[i].ent = reinterpret_cast<uint64_t*>(B->d_scratch + ofs[i].ent);
        jobs[i].bnd = reinterpret_cast<uint64_t*>(B->d_scratch + ofs[i].bnd);
        jobs[i].gvals = B->d_scratch + ofs[i].G;
        jobs[i].posE = reinterpret_cast<uint32_t*>(B->d_scratch + ofs[i].E);
        jobs[i].posO = reinterpret_cast<uint32_t*>(B->d_scratch + ofs[i].O);
        jobs[i].pre = reinterpret_cast<uint64_t*>(B->d_scratch + ofs[i].pre);
    }
    bool ok = hipMemcpy(B->d_jobs, jobs.data(), jobs.size() * sizeof(LevelJob), hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(B->d_texs, texs.data(), texs.size() * sizeof(TexInfo), hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(B->d_wjobs, wj.data(), wj.size() * sizeof(WriteJob), hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(B->d_pals, pals.data(), pals.size() * sizeof(WritePalettes), hipMemcpyHostToDevice) == hipSuccess;
    {
        std::vector<uint32_t> tab_job(wg_trans);
        for (uint32_t i = 0; i < (uint32_t)jobs.size(); i++)
            std::fill(tab_job.begin() + jobs[i].wg_trans, tab_job.begin() + jobs[i].wg_trans + jobs[i].tlen / kTabPerWG, i);
        ok &= hipMemcpy(B->d_tab_job, tab_job.data(), tab_job.size() * sizeof(uint32_t), hipMemcpyHostToDevice) ==
              hipSuccess;
        if (!carry.empty())
            ok &= hipMemcpy(B->d_carry, carry.data(), carry.size() * sizeof(uint32_t), hipMemcpyHostToDevice) ==
                  hipSuccess;
    }
    if (!ok) {
        batch_free(B);
        return CRN2_EHIP;
    }
    B->magic = kBatchMagic;
    *out = B;
    return CRN2_OK;
}

// Queues the plan's launches on `stream`; with `ev`, records an event before
// each stage and after the last.
static int launch_plan(crn2_batch* B, hipStream_t stream, Profile::Ev* ev)
{
    const bool prof = ev != nullptr;
    if (prof) HIPCHK(hipEventRecord(ev->e[0], stream));
    if (B->njobs) {
        hipLaunchKernelGGL(crnd_tables_kernel, dim3(B->wg_trans), dim3(kTabThreads), 0, stream, B->d_jobs,
                           B->d_tab_job, B->d_texs);
        HIPCHK(hipGetLastError());
        if (prof) HIPCHK(hipEventRecord(ev->e[1], stream));
        if (B->njobs >= kPlanJobsSmallRing)
            hipLaunchKernelGGL(crnd_walk_kernel<4>, dim3(B->njobs), dim3(128), 0, stream, B->d_jobs, B->d_texs);
        else
            hipLaunchKernelGGL(crnd_walk_kernel<8>, dim3(B->njobs), dim3(128), 0, stream, B->d_jobs, B->d_texs);
        HIPCHK(hipGetLastError());
        if (prof) HIPCHK(hipEventRecord(ev->e[2], stream));
        hipLaunchKernelGGL(crnd_sym_kernel<false>, dim3(B->wg_sym), dim3(256), 0, stream, B->d_jobs, B->njobs,
                           B->d_texs);
        HIPCHK(hipGetLastError());
        if (prof) HIPCHK(hipEventRecord(ev->e[3], stream));
        hipLaunchKernelGGL(crnd_sym_kernel<true>, dim3(B->wg_odd), dim3(256), 0, stream, B->d_jobs, B->njobs,
                           B->d_texs);
        HIPCHK(hipGetLastError());
        if (prof) HIPCHK(hipEventRecord(ev->e[4], stream));
        hipLaunchKernelGGL(crnd_index_chunk_kernel, dim3(B->wg_idx), dim3(kIdxThreads), 0, stream, B->d_jobs, B->njobs,
                           B->d_texs);
        HIPCHK(hipGetLastError());
        if (B->ncarry) {
            hipLaunchKernelGGL(crnd_index_carry_kernel, dim3(B->ncarry), dim3(kCarryThreads), 0, stream, B->d_jobs,
                               B->d_carry, B->d_texs);
            HIPCHK(hipGetLastError());
        }
        if (prof) HIPCHK(hipEventRecord(ev->e[5], stream));
        const uint32_t fb = (uint32_t)((B->blk_total + 255) / 256);
        hipLaunchKernelGGL(crnd_index_final_kernel, dim3(fb), dim3(256), 0, stream, B->d_jobs, B->njobs, B->d_texs,
                           B->blk_total, B->d_wjobs, B->d_pals);
        HIPCHK(hipGetLastError());
    } else if (prof) {
        for (int k = 1; k < kStages; k++) HIPCHK(hipEventRecord(ev->e[k], stream));
    }
    if (prof) HIPCHK(hipEventRecord(ev->e[kStages], stream));
    return CRN2_OK;
}

static int run_plan(crn2_batch* B, hipStream_t stream)
{
    bool prof;
    {
        std::lock_guard<std::mutex> lk(g_prof.mu);
        prof = g_prof.on;
    }
    if (!prof) return launch_plan(B, stream, nullptr);
    Profile::Ev ev{};
    for (int k = 0; k <= kStages; k++) HIPCHK(hipEventCreate(&ev.e[k]));
    const int rc = launch_plan(B, stream, &ev);
    if (rc) return rc;
    std::lock_guard<std::mutex> lk(g_prof.mu);
    g_prof.pending.push_back(ev);
    return CRN2_OK;
}

// Captures the plan's launches into a graph once (on a private stream: the
// caller's may be the legacy default stream, which cannot be captured) and
// replays it on `stream`.  The launches hold no memset nodes.
static int run_plan_graph(crn2_batch* B, hipStream_t stream)
{
    if (!B->graph) {
        hipStream_t cs = nullptr;
        HIPCHK(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
        hipGraph_t g = nullptr;
        int rc = CRN2_OK;
        if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = CRN2_EHIP;
        if (!rc) rc = launch_plan(B, cs, nullptr);
        if (hipStreamEndCapture(cs, &g) != hipSuccess && !rc) rc = CRN2_EHIP;
        if (!rc && hipGraphInstantiate(&B->graph, g, nullptr, nullptr, 0) != hipSuccess) {
            B->graph = nullptr;
            rc = CRN2_EHIP;
        }
        if (g) (void)hipGraphDestroy(g);
        (void)hipStreamDestroy(cs);
        if (rc) return rc;
    }
    HIPCHK(hipGraphLaunch(B->graph, stream));
    return CRN2_OK;
}

static int check_level_args(crn2_unpack_ctx* c, uint32_t dst_size, uint32_t& row_pitch, uint32_t level)
{
    const Header& h = c->h;
    if (level >= h.levels) return CRN2_EINVAL;
    const uint32_t bx = (level_dim(h.width, level) + 3) >> 2, by = (level_dim(h.height, level) + 3) >> 2;
    const uint32_t minimal = bytes_per_block(h.format) * bx;
    if (!row_pitch) row_pitch = minimal;
    else if (row_pitch < minimal || (row_pitch & 3)) return CRN2_EINVAL;
    if ((uint64_t)dst_size < (uint64_t)row_pitch * by) return CRN2_EINVAL;
    return CRN2_OK;
}

static int read_status(int32_t* d_status, hipStream_t stream)
{
    int32_t st = 0;
    HIPCHK(hipMemcpyAsync(&st, d_status, 4, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    if (st) {
        HIPCHK(hipMemsetAsync(d_status, 0, 4, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }
    return st;
}

static PlanPart single_level_part(crn2_unpack_ctx* c, uint32_t level, uint8_t* const* dst, uint32_t pitch,
                                  const uint8_t* src_override, uint32_t src_size)
{
    PlanPart P;
    P.c = c;
    P.first = level;
    P.count = 1;
    P.src_override = src_override;
    P.src_size = src_size;
    std::array<uint8_t*, 6> d{};
    for (uint32_t f = 0; f < c->h.faces; f++) d[f] = dst[f];
    P.dst.push_back(d);
    P.pitch.push_back(pitch);
    return P;
}

static int unpack_level_host(crn2_unpack_ctx* c, const uint8_t* src_host, uint32_t src_size, void** dst,
                             uint32_t dst_size, uint32_t row_pitch, uint32_t level)
{
    if (!ctx_ok(c) || !dst || dst_size < 8 || level >= kMaxLevels) return CRN2_EINVAL;
    int rc = check_level_args(c, dst_size, row_pitch, level);
    if (rc) return rc;
    const Header& h = c->h;
    for (uint32_t f = 0; f < h.faces; f++)
        if (!dst[f]) return CRN2_EINVAL;
    const uint32_t by = (level_dim(h.height, level) + 3) >> 2;
    const uint64_t face_bytes = (uint64_t)row_pitch * by;
    hipStream_t stream = nullptr;
    uint8_t* d_out = nullptr;
    uint8_t* d_src = nullptr;
    if (dmalloc(&d_out, face_bytes * h.faces) != hipSuccess) return CRN2_ENOMEM;
    if (src_host) {
        const size_t pad = ((size_t)src_size + 15) / 16 * 16 + 64;
        if (dmalloc(&d_src, pad) != hipSuccess) rc = CRN2_ENOMEM;
        else if (hipMemsetAsync(d_src, 0, pad, stream) != hipSuccess ||
                 hipMemcpyAsync(d_src, src_host, src_size, hipMemcpyHostToDevice, stream) != hipSuccess)
            rc = CRN2_EHIP;
    }
    // the caller's buffer may hold bytes the transcoder does not write (pitch padding)
    for (uint32_t f = 0; f < h.faces && !rc; f++)
        if (hipMemcpyAsync(d_out + f * face_bytes, dst[f], face_bytes, hipMemcpyHostToDevice, stream) != hipSuccess)
            rc = CRN2_EHIP;
    crn2_batch* B = nullptr;
    if (!rc) {
        uint8_t* faces[6] = {};
        for (uint32_t f = 0; f < h.faces; f++) faces[f] = d_out + f * face_bytes;
        std::vector<PlanPart> parts{single_level_part(c, level, faces, row_pitch, d_src, src_size)};
        rc = build_plan(parts, &B);
    }
    if (!rc) rc = run_plan(B, stream);
    if (!rc) rc = read_status(c->d_status, stream);
    for (uint32_t f = 0; f < h.faces && !rc; f++)
        if (hipMemcpyAsync(dst[f], d_out + f * face_bytes, face_bytes, hipMemcpyDeviceToHost, stream) != hipSuccess)
            rc = CRN2_EHIP;
    if (hipStreamSynchronize(stream) != hipSuccess && !rc) rc = CRN2_EHIP;
    batch_free(B);
    (void)hipFree(d_out);
    if (d_src) (void)hipFree(d_src);
    return rc;
}

static PlanPart all_levels_part(crn2_unpack_ctx* c, uint8_t* base, uint32_t first, uint32_t count)
{
    PlanPart P;
    P.c = c;
    P.first = first;
    P.count = count;
    P.src_override = nullptr;
    P.src_size = 0;
    const uint32_t bpb = bytes_per_block(c->h.format);
    uint64_t acc = 0;
    for (uint32_t l = first; l < first + count; l++) {
        const uint32_t bx = (level_dim(c->h.width, l) + 3) >> 2, by = (level_dim(c->h.height, l) + 3) >> 2;
        std::array<uint8_t*, 6> d{};
        for (uint32_t f = 0; f < c->h.faces; f++) d[f] = base + acc + (uint64_t)f * bx * by * bpb;
        acc += (uint64_t)bx * by * bpb * c->h.faces;
        P.dst.push_back(d);
        P.pitch.push_back(bx * bpb);
    }
    return P;
}

}  // namespace

// ============================================================== C ABI
extern "C" {

crn2_unpack_ctx* crn2_unpack_begin(const void* data, uint32_t size)
{
    return begin_impl(data, nullptr, size, nullptr);
}

crn2_unpack_ctx* crn2_unpack_begin_device(const void* data, const void* data_dev, uint32_t size, void* stream)
{
    if (!data_dev) return nullptr;
    return begin_impl(data, data_dev, size, static_cast<hipStream_t>(stream));
}

int crn2_get_data(crn2_unpack_ctx* c, const void** data, uint32_t* size)
{
    if (!ctx_ok(c)) return CRN2_EINVAL;
    if (data) *data = c->data;
    if (size) *size = c->size;
    return CRN2_OK;
}

int crn2_unpack_level(crn2_unpack_ctx* c, void** dst, uint32_t dst_size, uint32_t row_pitch, uint32_t level)
{
    return unpack_level_host(c, nullptr, 0, dst, dst_size, row_pitch, level);
}

int crn2_unpack_level_segmented(crn2_unpack_ctx* c, const void* src, uint32_t src_size, void** dst,
                                uint32_t dst_size, uint32_t row_pitch, uint32_t level)
{
    if (!src || !src_size) return CRN2_EINVAL;
    return unpack_level_host(c, static_cast<const uint8_t*>(src), src_size, dst, dst_size, row_pitch, level);
}

int crn2_unpack_level_device(crn2_unpack_ctx* c, void** dst_dev, uint32_t dst_size, uint32_t row_pitch,
                             uint32_t level, void* stream)
{
    if (!ctx_ok(c) || !dst_dev || dst_size < 8 || level >= kMaxLevels) return CRN2_EINVAL;
    int rc = check_level_args(c, dst_size, row_pitch, level);
    if (rc) return rc;
    for (uint32_t f = 0; f < c->h.faces; f++)
        if (!dst_dev[f]) return CRN2_EINVAL;
    uint8_t* faces[6] = {};
    for (uint32_t f = 0; f < c->h.faces; f++) faces[f] = static_cast<uint8_t*>(dst_dev[f]);
    std::vector<PlanPart> parts{single_level_part(c, level, faces, row_pitch, nullptr, 0)};
    crn2_batch* B = nullptr;
    rc = build_plan(parts, &B);
    if (rc) return rc;
    rc = run_plan(B, static_cast<hipStream_t>(stream));
    // the job tables must outlive the queued kernels: keep the plan on the context
    batch_free(c->cached);
    c->cached = B;
    c->cached_dst = nullptr;
    return rc;
}

uint64_t crn2_levels_size(crn2_unpack_ctx* c, uint32_t first, uint32_t count)
{
    if (!ctx_ok(c) || first + count > c->h.levels) return 0;
    uint64_t s = 0;
    for (uint32_t l = first; l < first + count; l++) {
        const uint32_t bx = (level_dim(c->h.width, l) + 3) >> 2, by = (level_dim(c->h.height, l) + 3) >> 2;
        s += (uint64_t)bx * by * bytes_per_block(c->h.format) * c->h.faces;
    }
    return s;
}

int crn2_unpack_levels_device(crn2_unpack_ctx* c, void* dst_dev, uint64_t dst_size, uint32_t first, uint32_t count,
                              void* stream)
{
    if (!ctx_ok(c) || !dst_dev || !count || first + count > c->h.levels) return CRN2_EINVAL;
    if (dst_size < crn2_levels_size(c, first, count)) return CRN2_EINVAL;
    if (!(c->cached && c->cached_dst == dst_dev && c->cached_first == first && c->cached_count == count)) {
        std::vector<PlanPart> parts{all_levels_part(c, static_cast<uint8_t*>(dst_dev), first, count)};
        crn2_batch* B = nullptr;
        const int rc = build_plan(parts, &B);
        if (rc) return rc;
        batch_free(c->cached);
        c->cached = B;
        c->cached_dst = dst_dev;
        c->cached_first = first;
        c->cached_count = count;
    }
    return run_plan(c->cached, static_cast<hipStream_t>(stream));
}

int crn2_batch_create(crn2_unpack_ctx* const* ctxs, uint32_t n, void* const* dst_dev, const uint64_t* dst_size,
                      crn2_batch** out)
{
    if (!ctxs || !n || !dst_dev || !dst_size || !out) return CRN2_EINVAL;
    std::vector<PlanPart> parts;
    for (uint32_t i = 0; i < n; i++) {
        crn2_unpack_ctx* c = ctxs[i];
        if (!ctx_ok(c) || !dst_dev[i] || dst_size[i] < crn2_levels_size(c, 0, c->h.levels)) return CRN2_EINVAL;
        parts.push_back(all_levels_part(c, static_cast<uint8_t*>(dst_dev[i]), 0, c->h.levels));
    }
    return build_plan(parts, out);
}

int crn2_batch_run(crn2_batch* b, void* stream)
{
    if (!b || b->magic != kBatchMagic) return CRN2_EINVAL;
    return run_plan(b, static_cast<hipStream_t>(stream));
}

int crn2_batch_run_graph(crn2_batch* b, void* stream)
{
    if (!b || b->magic != kBatchMagic) return CRN2_EINVAL;
    return run_plan_graph(b, static_cast<hipStream_t>(stream));
}

int crn2_batch_status(crn2_batch* b, void* stream)
{
    if (!b || b->magic != kBatchMagic) return CRN2_EINVAL;
    int worst = CRN2_OK;
    for (int32_t* s : b->status) {
        const int st = read_status(s, static_cast<hipStream_t>(stream));
        if (st) worst = st;
    }
    return worst;
}

int crn2_batch_destroy(crn2_batch* b)
{
    if (!b || b->magic != kBatchMagic) return CRN2_EINVAL;
    batch_free(b);
    return CRN2_OK;
}

int crn2_unpack_end(crn2_unpack_ctx* c)
{
    if (!ctx_ok(c)) return CRN2_EINVAL;
    c->magic = 0;
    (void)hipDeviceSynchronize();
    ctx_free_device(c);
    delete c;
    return CRN2_OK;
}

void crn2_profile_enable(int on)
{
    std::lock_guard<std::mutex> lk(g_prof.mu);
    g_prof.on = on != 0;
}

void crn2_profile_reset(void)
{
    std::lock_guard<std::mutex> lk(g_prof.mu);
    profile_collect_locked();
    for (int s = 0; s < kStages; s++) {
        g_prof.ms[s] = 0;
        g_prof.n[s] = 0;
    }
}

int crn2_profile_read(int kernel, double* total_ms, uint64_t* launches)
{
    if (kernel < 0 || kernel >= kStages || !total_ms || !launches) return CRN2_EINVAL;
    std::lock_guard<std::mutex> lk(g_prof.mu);
    profile_collect_locked();
    *total_ms = g_prof.ms[kernel];
    *launches = g_prof.n[kernel];
    return CRN2_OK;
}

}  // extern "C"
