// crnd_gpu.hip -- CRN->DXTn transcoder on gfx950.
//
// Reference: inc/crn_decomp.h (crnd::crn_unpacker).  A level of a .crn file is
// one prefix-coded stream (unpack_dxt1 and friends, crn_decomp.h:3944-4223):
// which code decodes the next symbol depends on reference codes decoded
// earlier, and the odd row of each row pair takes its references from the
// even row above.  A level decodes in one pass from its first bit; work
// splits across levels and textures, not inside a stream.
//
// Device pipeline for a batch of textures ("plan", built once, run many times):
//   1. crnd_parse_kernel: one workgroup (one wave) per texture, one lane per
//      level stream.  The texture's prefix-code lookup tables, canonical
//      symbol orders and the above-row block buffers of all its levels
//      (crn_decomp.h:3652-3660) are staged in LDS; the compressed bits are
//      read in 16-byte chunks prefetched two chunks ahead.  Each lane writes
//      one packed palette-index record per block of the even-padded grid.
//   2. crnd_write_kernel: one thread per visible block gathers the palette
//      entries named by its record and writes the 8- or 16-byte DXTn block;
//      consecutive threads write consecutive blocks of a row (coalesced).
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
constexpr uint32_t kNoLds = 0xFFFFFFFFu;

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

struct ParseJob {
    const uint8_t* src;   // level bytes (device); readable up to the next 16-byte boundary past src + src_size + 32
    uint32_t src_size;
    uint32_t gw, gh, faces;
    uint32_t bb_ofs;      // entries into the group's LDS block-buffer region
    void* rec;            // faces * gh * gw records (uint32 for kinds 0/3, uint64 for kinds 1/2)
};

struct ParseGroup {
    const DevModel* models;      // M_COUNT models
    const uint16_t* sorted;      // M_COUNT x kMaxSyms canonical symbol orders (global copy)
    int32_t* status;             // set to CRN2_ECORRUPT on an out-of-range palette index
    uint32_t kind, n_ce, n_cs, n_ae, n_as;
    uint32_t first_job, njobs;
    uint32_t lut_ofs[M_COUNT];   // LDS byte offsets of the tables (kNoLds when unused)
    uint32_t sorted_ofs[M_COUNT];// LDS byte offsets of canonical orders (kNoLds: read the global copy)
    uint32_t sorted_n[M_COUNT];
    uint32_t bb_ofs;             // LDS byte offset of the block buffers
};

struct LdsTables {
    uint32_t lut[kLutSize];
    uint32_t limit[17];
    int32_t base[17];
    uint32_t nsyms;
    uint32_t pad;
};

// ---------------------------------------------------------------- bit reader
// MSB-first reader over 16-byte chunks: `cur` is being consumed word by word,
// `nxt` is already in flight.  Bytes at or past `end` read as zero
// (crn_decomp.h:3196-3207); chunks at or past the end are not loaded.
struct DevBits {
    const uint4* cp;      // next chunk to load
    const uint8_t* end;
    const uint8_t* qaddr; // address of `cur`
    uint4 cur, nxt;
    int qi;               // next word of `cur`
    uint64_t buf;         // valid bits MSB-aligned
    int cnt;

    __device__ inline uint4 load_chunk(const uint4* p) const
    {
        if (reinterpret_cast<const uint8_t*>(p) < end) return *p;
        return make_uint4(0, 0, 0, 0);
    }

    __device__ inline uint32_t word(int i) const
    {
        const uint32_t w = (i == 0) ? cur.x : (i == 1) ? cur.y : (i == 2) ? cur.z : cur.w;
        const uint8_t* wa = qaddr + 4 * i;
        uint32_t v = __builtin_bswap32(w);
        if (wa + 4 > end) v = (wa >= end) ? 0u : (v & (0xFFFFFFFFu << (8 * (4 - (int)(end - wa)))));
        return v;
    }

    __device__ inline void refill()
    {
        if (qi == 4) {
            cur = nxt;
            qaddr += 16;
            nxt = load_chunk(cp);
            cp++;
            qi = 0;
        }
        buf |= (uint64_t)word(qi) << (32 - cnt);
        cnt += 32;
        qi++;
    }

    __device__ inline void start(const uint8_t* p, uint32_t size)
    {
        const uintptr_t a = reinterpret_cast<uintptr_t>(p);
        const uint4* c0 = reinterpret_cast<const uint4*>(a & ~(uintptr_t)15);
        end = p + size;
        qaddr = reinterpret_cast<const uint8_t*>(c0);
        cur = load_chunk(c0);
        nxt = load_chunk(c0 + 1);
        cp = c0 + 2;
        qi = (int)((a & 15) >> 2);
        buf = 0;
        cnt = 0;
        refill();
        const int skip = (int)(a & 3) * 8;
        buf <<= skip;
        cnt -= skip;
        if (cnt <= 32) refill();
    }

    __device__ inline uint32_t decode(const LdsTables& t, const uint16_t* sorted)
    {
        if (cnt <= 32) refill();
        const uint32_t top = (uint32_t)(buf >> 48);
        const uint32_t e = t.lut[top >> (16 - kLutBits)];
        uint32_t sym, len;
        if (e) {
            sym = e & 0xFFFFu;
            len = e >> 16;
        } else {
            // code longer than the table or no code at all: canonical walk
            const uint32_t k = top + 1;
            len = kLutBits + 1;
            while (len <= 16 && k > t.limit[len]) len++;
            if (len > 16) return 0;   // no code: symbol 0, nothing consumed (crn_decomp.h:3239-3244)
            const uint32_t idx = (uint32_t)(t.base[len] + (int32_t)(top >> (16 - len)));
            if (idx >= t.nsyms) return 0;
            sym = sorted[idx];
        }
        buf <<= len;
        cnt -= (int)len;
        return sym;
    }
};

__global__ void __launch_bounds__(64) crnd_parse_kernel(const ParseGroup* __restrict__ groups,
                                                        const ParseJob* __restrict__ jobs)
{
    extern __shared__ __align__(16) uint8_t lds[];
    const ParseGroup& g = groups[blockIdx.x];
    const int kind = (int)g.kind;

    // ---- stage tables, canonical orders; zero the block buffers
    for (uint32_t m = 0; m < M_COUNT; m++) {
        if (g.lut_ofs[m] == kNoLds) continue;
        LdsTables* t = reinterpret_cast<LdsTables*>(lds + g.lut_ofs[m]);
        const DevModel& dm = g.models[m];
        for (uint32_t i = threadIdx.x; i < kLutSize; i += blockDim.x) t->lut[i] = dm.lut[i];
        if (threadIdx.x < 17) {
            t->limit[threadIdx.x] = dm.limit[threadIdx.x];
            t->base[threadIdx.x] = dm.base[threadIdx.x];
        }
        if (threadIdx.x == 0) t->nsyms = dm.num_syms;
        if (g.sorted_ofs[m] != kNoLds) {
            uint16_t* s = reinterpret_cast<uint16_t*>(lds + g.sorted_ofs[m]);
            const uint16_t* src = g.sorted + m * kMaxSyms;
            for (uint32_t i = threadIdx.x; i < g.sorted_n[m]; i += blockDim.x) s[i] = src[i];
        }
    }
    uint32_t bb_total = 0;
    for (uint32_t k = 0; k < g.njobs; k++) bb_total = max(bb_total, jobs[g.first_job + k].bb_ofs + jobs[g.first_job + k].gw);
    uint32_t* bbase = reinterpret_cast<uint32_t*>(lds + g.bb_ofs);
    for (uint32_t i = threadIdx.x; i < bb_total; i += blockDim.x) bbase[i] = 0;
    __syncthreads();
    if (threadIdx.x >= g.njobs) return;
    const ParseJob j = jobs[g.first_job + threadIdx.x];

    const uint32_t slot_ep0 = (kind == 0 || kind == 1) ? M_EP0 : M_EP1;
    const uint32_t slot_sel0 = (kind == 0 || kind == 1) ? M_SEL0 : M_SEL1;
    const LdsTables& T_ref = *reinterpret_cast<const LdsTables*>(lds + g.lut_ofs[M_REF]);
    const LdsTables& T_ep0 = *reinterpret_cast<const LdsTables*>(lds + g.lut_ofs[slot_ep0]);
    const LdsTables& T_sel0 = *reinterpret_cast<const LdsTables*>(lds + g.lut_ofs[slot_sel0]);
    const LdsTables& T_ep1 = *reinterpret_cast<const LdsTables*>(lds + g.lut_ofs[kind == 0 ? M_REF : M_EP1]);
    const LdsTables& T_sel1 = *reinterpret_cast<const LdsTables*>(lds + g.lut_ofs[kind == 0 ? M_REF : M_SEL1]);
    auto sorted_of = [&](uint32_t m) -> const uint16_t* {
        return g.sorted_ofs[m] != kNoLds ? reinterpret_cast<const uint16_t*>(lds + g.sorted_ofs[m])
                                         : g.sorted + m * kMaxSyms;
    };
    const uint16_t* S_ref = sorted_of(M_REF);
    const uint16_t* S_ep0 = sorted_of(slot_ep0);
    const uint16_t* S_sel0 = sorted_of(slot_sel0);
    const uint16_t* S_ep1 = sorted_of(kind == 0 ? M_REF : M_EP1);
    const uint16_t* S_sel1 = sorted_of(kind == 0 ? M_REF : M_SEL1);

    DevBits bits;
    bits.start(j.src, j.src_size);
    const uint32_t gw = j.gw, gh = j.gh;
    uint32_t* bb = bbase + j.bb_ofs;
    // i0: colour endpoint (kinds 0, 1) or alpha0 endpoint (kinds 2, 3);
    // i1: alpha endpoint (kind 1) or alpha1 endpoint (kind 2)
    uint32_t i0 = 0, i1 = 0;
    uint32_t group = 0;
    bool bad = false;
    uint64_t blk = 0;
    const uint32_t n0 = (kind == 0 || kind == 1) ? g.n_ce : g.n_ae;
    const uint32_t n1 = (kind == 1 || kind == 2) ? g.n_ae : 1u;
    const uint32_t ns0 = (kind == 0 || kind == 1) ? g.n_cs : g.n_as;
    const uint32_t ns1 = g.n_as;
    uint32_t* rec32 = static_cast<uint32_t*>(j.rec);
    uint64_t* rec64 = static_cast<uint64_t*>(j.rec);
    const bool two = (kind == 1 || kind == 2);
    for (uint32_t f = 0; f < j.faces; f++) {
        for (uint32_t y = 0; y < gh; y++) {
            for (uint32_t x = 0; x < gw; x++, blk++) {
                const uint32_t cur = bb[x];
                uint32_t ref, below;
                if (!(y & 1)) {
                    if (!(x & 1)) group = bits.decode(T_ref, S_ref) & 0xFFu;
                    ref = group & 3u;
                    below = (group >> 2) & 3u;
                    group >>= 4;
                } else {
                    ref = cur & 3u;
                    below = ref;
                }
                if (ref == 0) {
                    i0 += bits.decode(T_ep0, S_ep0);
                    if (i0 >= n0) i0 -= n0;
                    if (two) {
                        i1 += bits.decode(T_ep1, S_ep1);
                        if (i1 >= n1) i1 -= n1;
                    }
                } else if (ref >= 2) {
                    i0 = (cur >> 2) & 0x3FFFu;
                    i1 = (cur >> 16) & 0x3FFFu;
                }
                bb[x] = below | (i0 << 2) | (i1 << 16);
                const uint32_t s0 = bits.decode(T_sel0, S_sel0);
                bad |= (i0 >= n0) | (s0 >= ns0);
                if (!two) {
                    rec32[blk] = i0 | (s0 << 16);
                } else {
                    const uint32_t s1 = bits.decode(T_sel1, S_sel1);
                    bad |= (i1 >= n1) | (s1 >= ns1);
                    rec64[blk] = (uint64_t)(i0 | (s0 << 16)) | ((uint64_t)(i1 | (s1 << 16)) << 32);
                }
            }
        }
    }
    if (bad) atomicExch(g.status, CRN2_ECORRUPT);
}

struct WriteJob {
    const void* rec;        // records of the level (padded grid, faces stacked)
    uint8_t* dst[6];        // per-face destination
    uint32_t pitch;         // bytes between block rows
    uint32_t bx, by;        // visible blocks
    uint32_t gw, gh;        // padded grid
    uint32_t faces;
    uint32_t pal;           // palette set index
    uint64_t first_block;   // prefix of visible blocks over the jobs of the launch
};

struct WritePalettes {
    const uint32_t* ce;
    const uint32_t* cs;
    const uint16_t* ae;
    const uint16_t* as;     // 3 per entry
    uint32_t kind;
    uint32_t pad;
};

__global__ void __launch_bounds__(256) crnd_write_kernel(const WriteJob* __restrict__ jobs, uint32_t njobs,
                                                         const WritePalettes* __restrict__ pals, uint64_t total)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].first_block <= t) lo = mid;
        else hi = mid;
    }
    const WriteJob& j = jobs[lo];
    const WritePalettes& P = pals[j.pal];
    uint64_t v = t - j.first_block;
    const uint32_t per_face = j.bx * j.by;
    const uint32_t f = (uint32_t)(v / per_face);
    v -= (uint64_t)f * per_face;
    const uint32_t y = (uint32_t)(v / j.bx), x = (uint32_t)(v % j.bx);
    const uint64_t r = (uint64_t)(f * j.gh + y) * j.gw + x;
    uint8_t* out = j.dst[f] + (uint64_t)y * j.pitch;
    if (P.kind == 0 || P.kind == 3) {
        const uint32_t rec = static_cast<const uint32_t*>(j.rec)[r];
        uint2 o;
        if (P.kind == 0) {
            o.x = P.ce[rec & 0xFFFFu];
            o.y = P.cs[rec >> 16];
        } else {
            const uint16_t* s = P.as + 3 * (rec >> 16);
            o.x = (uint32_t)P.ae[rec & 0xFFFFu] | ((uint32_t)s[0] << 16);
            o.y = (uint32_t)s[1] | ((uint32_t)s[2] << 16);
        }
        if ((j.pitch & 7) == 0) {
            reinterpret_cast<uint2*>(out)[x] = o;
        } else {
            uint32_t* w = reinterpret_cast<uint32_t*>(out + 8 * (uint64_t)x);
            w[0] = o.x;
            w[1] = o.y;
        }
    } else {
        const uint64_t rec = static_cast<const uint64_t*>(j.rec)[r];
        const uint32_t lo32 = (uint32_t)rec, hi32 = (uint32_t)(rec >> 32);
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
        if ((j.pitch & 15) == 0) {
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

// ----------------------------------------------------------------- profiling
struct Profile {
    std::mutex mu;
    bool on = false;
    struct Ev { hipEvent_t a, b, c; };
    std::vector<Ev> pending;
    double ms[2] = {0, 0};
    uint64_t n[2] = {0, 0};
};
static Profile g_prof;

static void profile_collect_locked()
{
    for (auto& e : g_prof.pending) {
        float t0 = 0, t1 = 0;
        (void)hipEventSynchronize(e.c);
        if (hipEventElapsedTime(&t0, e.a, e.b) == hipSuccess) {
            g_prof.ms[0] += t0;
            g_prof.n[0]++;
        }
        if (hipEventElapsedTime(&t1, e.b, e.c) == hipSuccess) {
            g_prof.ms[1] += t1;
            g_prof.n[1]++;
        }
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
        (void)hipEventDestroy(e.c);
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
    uint32_t* d_ce = nullptr;
    uint32_t* d_cs = nullptr;
    uint16_t* d_ae = nullptr;
    uint16_t* d_as = nullptr;
    int32_t* d_status = nullptr;
    void* d_rec = nullptr;       // records of every level
    uint64_t rec_ofs[kMaxLevels + 1] = {};
    uint32_t n_ce = 0, n_cs = 0, n_ae = 0, n_as = 0;
    crn2_batch* cached = nullptr;   // plan of the last crn2_unpack_levels_device / _level_device call
    void* cached_dst = nullptr;
    uint32_t cached_first = 0, cached_count = 0, cached_pitch = 0;
};

// A prepared launch over one or more contexts.
struct crn2_batch {
    uint32_t magic = 0;
    ParseGroup* d_groups = nullptr;
    ParseJob* d_jobs = nullptr;
    WriteJob* d_wjobs = nullptr;
    WritePalettes* d_pals = nullptr;
    uint32_t ngroups = 0, nwjobs = 0;
    uint64_t total_blocks = 0;
    size_t lds_bytes = 0;
    std::vector<int32_t*> status;   // per context
};

namespace {

constexpr uint32_t kCtxMagic = 0x1EF9CABDu;
constexpr uint32_t kBatchMagic = 0x7B47C2A1u;

static uint32_t level_dim(uint32_t v, uint32_t l) { return std::max<uint32_t>(1u, v >> l); }

static uint32_t bytes_per_block(uint32_t fmt)
{
    return (fmt == CRN2_FMT_DXT1 || fmt == CRN2_FMT_DXT5A) ? 8u : 16u;
}

static size_t max_lds_bytes()
{
    static size_t v = 0;
    if (!v) {
        int dev = 0;
        hipDeviceProp_t p;
        if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&p, dev) == hipSuccess)
            v = std::max<size_t>(p.sharedMemPerBlock, 65536);
        else
            v = 65536;
    }
    return v;
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
    void* ptrs[] = {b->d_groups, b->d_jobs, b->d_wjobs, b->d_pals};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    delete b;
}

static void ctx_free_device(crn2_unpack_ctx* c)
{
    batch_free(c->cached);
    c->cached = nullptr;
    void* ptrs[] = {c->d_file, c->d_models, c->d_sorted, c->d_ce, c->d_cs, c->d_ae, c->d_as, c->d_status, c->d_rec};
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
    uint64_t rec = 0;
    const uint32_t rec_bytes = (c->kind == 1 || c->kind == 2) ? 8 : 4;
    for (uint32_t l = 0; l < h.levels; l++) {
        const uint32_t bx = (level_dim(h.width, l) + 3) >> 2, by = (level_dim(h.height, l) + 3) >> 2;
        const uint32_t gw = (bx + 1) & ~1u, gh = (by + 1) & ~1u;
        c->rec_ofs[l] = rec;
        rec += (uint64_t)gw * gh * h.faces * rec_bytes;
    }
    c->rec_ofs[h.levels] = rec;

    const size_t padded = ((size_t)size + 15) / 16 * 16 + 64;
    if (dmalloc(&c->d_file, padded) != hipSuccess || dmalloc(&c->d_models, sizeof(DevModel) * M_COUNT) != hipSuccess ||
        dmalloc(&c->d_sorted, sorted.size() * 2) != hipSuccess || dmalloc(&c->d_ce, ce.size() * 4) != hipSuccess ||
        dmalloc(&c->d_cs, cs.size() * 4) != hipSuccess || dmalloc(&c->d_ae, ae.size() * 2) != hipSuccess ||
        dmalloc(&c->d_as, as.size() * 2) != hipSuccess || dmalloc(&c->d_status, 4) != hipSuccess ||
        dmalloc(&c->d_rec, rec) != hipSuccess)
        return fail();
    bool ok = true;
    ok &= hipMemsetAsync(c->d_file, 0, padded, stream) == hipSuccess;
    if (data_dev)
        ok &= hipMemcpyAsync(c->d_file, data_dev, size, hipMemcpyDeviceToDevice, stream) == hipSuccess;
    else
        ok &= hipMemcpyAsync(c->d_file, data, size, hipMemcpyHostToDevice, stream) == hipSuccess;
    ok &= hipMemcpyAsync(c->d_models, dms.data(), sizeof(DevModel) * M_COUNT, hipMemcpyHostToDevice, stream) == hipSuccess;
    ok &= hipMemcpyAsync(c->d_sorted, sorted.data(), sorted.size() * 2, hipMemcpyHostToDevice, stream) == hipSuccess;
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

static int build_plan(std::vector<PlanPart>& parts, crn2_batch** out)
{
    *out = nullptr;
    crn2_batch* B = new (std::nothrow) crn2_batch();
    if (!B) return CRN2_ENOMEM;
    std::vector<ParseGroup> groups;
    std::vector<ParseJob> jobs;
    std::vector<WriteJob> wj;
    std::vector<WritePalettes> pals;
    uint64_t total = 0;
    size_t lds_need = 0;
    for (uint32_t pi = 0; pi < parts.size(); pi++) {
        PlanPart& P = parts[pi];
        crn2_unpack_ctx* c = P.c;
        const Header& h = c->h;
        ParseGroup g;
        memset(&g, 0, sizeof(g));
        g.models = c->d_models;
        g.sorted = c->d_sorted;
        g.status = c->d_status;
        g.kind = (uint32_t)c->kind;
        g.n_ce = c->n_ce;
        g.n_cs = c->n_cs;
        g.n_ae = c->n_ae;
        g.n_as = c->n_as;
        g.first_job = (uint32_t)jobs.size();
        g.njobs = P.count;
        // LDS layout: tables, then block buffers, then canonical orders while they fit
        size_t ofs = 0;
        for (uint32_t m = 0; m < M_COUNT; m++) {
            g.lut_ofs[m] = kNoLds;
            g.sorted_ofs[m] = kNoLds;
            g.sorted_n[m] = 0;
            if (!kind_uses(c->kind, m)) continue;
            g.lut_ofs[m] = (uint32_t)ofs;
            ofs += sizeof(LdsTables);
        }
        uint32_t bb_entries = 0;
        g.bb_ofs = (uint32_t)ofs;
        const uint32_t first_job = (uint32_t)jobs.size();
        for (uint32_t i = 0; i < P.count; i++) {
            const uint32_t l = P.first + i;
            const uint32_t bx = (level_dim(h.width, l) + 3) >> 2, by = (level_dim(h.height, l) + 3) >> 2;
            const uint32_t gw = (bx + 1) & ~1u, gh = (by + 1) & ~1u;
            ParseJob j;
            memset(&j, 0, sizeof(j));
            if (P.src_override) {
                j.src = P.src_override;
                j.src_size = P.src_size;
            } else {
                const uint32_t next = (l + 1 < h.levels) ? h.level_ofs[l + 1] : c->size;
                j.src = c->d_file + h.level_ofs[l];
                j.src_size = next - h.level_ofs[l];
            }
            j.gw = gw;
            j.gh = gh;
            j.faces = h.faces;
            j.bb_ofs = bb_entries;
            bb_entries += gw;
            j.rec = static_cast<uint8_t*>(c->d_rec) + c->rec_ofs[l];
            jobs.push_back(j);
            WriteJob w;
            memset(&w, 0, sizeof(w));
            w.rec = j.rec;
            for (uint32_t f = 0; f < 6; f++) w.dst[f] = (f < h.faces) ? P.dst[i][f] : nullptr;
            w.pitch = P.pitch[i];
            w.bx = bx;
            w.by = by;
            w.gw = gw;
            w.gh = gh;
            w.faces = h.faces;
            w.pal = (uint32_t)pals.size();
            w.first_block = total;
            total += (uint64_t)bx * by * h.faces;
            wj.push_back(w);
        }
        (void)first_job;
        ofs += (size_t)bb_entries * 4;
        ofs = (ofs + 15) & ~(size_t)15;
        for (uint32_t m = 0; m < M_COUNT; m++) {
            if (!kind_uses(c->kind, m)) continue;
            const uint32_t n = (uint32_t)c->models[m].sorted.size();
            const size_t need = ((size_t)n * 2 + 15) & ~(size_t)15;
            if (ofs + need <= max_lds_bytes()) {
                g.sorted_ofs[m] = (uint32_t)ofs;
                g.sorted_n[m] = n;
                ofs += need;
            }
        }
        if (ofs > max_lds_bytes()) {
            batch_free(B);
            return CRN2_EUNSUPPORTED;
        }
        lds_need = std::max(lds_need, ofs);
        groups.push_back(g);
        WritePalettes wp{c->d_ce, c->d_cs, c->d_ae, c->d_as, (uint32_t)c->kind, 0};
        pals.push_back(wp);
        B->status.push_back(c->d_status);
        if (P.count > 64) {
            batch_free(B);
            return CRN2_EINVAL;
        }
    }
    B->ngroups = (uint32_t)groups.size();
    B->nwjobs = (uint32_t)wj.size();
    B->total_blocks = total;
    B->lds_bytes = lds_need;
    if (dmalloc(&B->d_groups, groups.size() * sizeof(ParseGroup)) != hipSuccess ||
        dmalloc(&B->d_jobs, jobs.size() * sizeof(ParseJob)) != hipSuccess ||
        dmalloc(&B->d_wjobs, wj.size() * sizeof(WriteJob)) != hipSuccess ||
        dmalloc(&B->d_pals, pals.size() * sizeof(WritePalettes)) != hipSuccess) {
        batch_free(B);
        return CRN2_ENOMEM;
    }
    bool ok = hipMemcpy(B->d_groups, groups.data(), groups.size() * sizeof(ParseGroup), hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(B->d_jobs, jobs.data(), jobs.size() * sizeof(ParseJob), hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(B->d_wjobs, wj.data(), wj.size() * sizeof(WriteJob), hipMemcpyHostToDevice) == hipSuccess;
    ok &= hipMemcpy(B->d_pals, pals.data(), pals.size() * sizeof(WritePalettes), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        batch_free(B);
        return CRN2_EHIP;
    }
    if (B->lds_bytes > 65536 &&
        hipFuncSetAttribute(reinterpret_cast<const void*>(&crnd_parse_kernel),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)B->lds_bytes) != hipSuccess) {
        batch_free(B);
        return CRN2_EHIP;
    }
    B->magic = kBatchMagic;
    *out = B;
    return CRN2_OK;
}

static int run_plan(crn2_batch* B, hipStream_t stream)
{
    Profile::Ev ev{};
    bool prof;
    {
        std::lock_guard<std::mutex> lk(g_prof.mu);
        prof = g_prof.on;
    }
    if (prof) {
        HIPCHK(hipEventCreate(&ev.a));
        HIPCHK(hipEventCreate(&ev.b));
        HIPCHK(hipEventCreate(&ev.c));
        HIPCHK(hipEventRecord(ev.a, stream));
    }
    hipLaunchKernelGGL(crnd_parse_kernel, dim3(B->ngroups), dim3(64), B->lds_bytes, stream, B->d_groups, B->d_jobs);
    HIPCHK(hipGetLastError());
    if (prof) HIPCHK(hipEventRecord(ev.b, stream));
    if (B->total_blocks) {
        const uint32_t tb = 256;
        const uint64_t nb = (B->total_blocks + tb - 1) / tb;
        hipLaunchKernelGGL(crnd_write_kernel, dim3((uint32_t)nb), dim3(tb), 0, stream, B->d_wjobs, B->nwjobs, B->d_pals,
                           B->total_blocks);
        HIPCHK(hipGetLastError());
    }
    if (prof) {
        HIPCHK(hipEventRecord(ev.c, stream));
        std::lock_guard<std::mutex> lk(g_prof.mu);
        g_prof.pending.push_back(ev);
    }
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
    g_prof.ms[0] = g_prof.ms[1] = 0;
    g_prof.n[0] = g_prof.n[1] = 0;
}

int crn2_profile_read(int kernel, double* total_ms, uint64_t* launches)
{
    if (kernel < 0 || kernel > 1 || !total_ms || !launches) return CRN2_EINVAL;
    std::lock_guard<std::mutex> lk(g_prof.mu);
    profile_collect_locked();
    *total_ms = g_prof.ms[kernel];
    *launches = g_prof.n[kernel];
    return CRN2_OK;
}

}  // extern "C"
