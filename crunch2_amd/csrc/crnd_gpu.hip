// crnd_gpu.hip -- CRN->DXTn transcoder on gfx950.
//
// Reference: inc/crn_decomp.h (crnd::crn_unpacker).  The level streams of a
// .crn file are strictly sequential prefix-coded streams (unpack_dxt1 and
// friends, crn_decomp.h:3944-4223): the code used for each symbol depends on
// the reference codes already decoded, and the odd row of every row pair takes
// its reference codes from the even row above it.  A level therefore decodes
// in one pass from its first bit; the work splits across levels, faces' level
// streams and textures, not inside a stream.
//
// Device pipeline per batch of level streams:
//   1. crnd_parse_kernel: one lane walks one level stream (all faces) with the
//      prefix-code lookup tables of its texture staged in LDS, and writes one
//      packed index record per block of the even-padded block grid.  The
//      above-row block buffer (crn_decomp.h:3652-3660) lives in a per-stream
//      scratch array and is prefetched one block ahead.
//   2. crnd_write_kernel: one thread per visible block gathers the palette
//      entries named by the record and writes the 8- or 16-byte DXTn block;
//      consecutive threads write consecutive blocks of a row (coalesced).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <vector>

#include "crn2_internal.h"
#include "crnd_host.h"

using namespace crn2;

namespace {

enum : uint32_t { M_REF = 0, M_EP0 = 1, M_SEL0 = 2, M_EP1 = 3, M_SEL1 = 4, M_COUNT = 5 };

// formats: kind 0 = DXT1 (colour), 1 = DXT5 family (alpha + colour),
//          2 = DXN (two alpha), 3 = DXT5A (one alpha)
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

struct ParseJob {
    const uint8_t* src;   // level bytes (device); src - (src & 3) is 4-byte aligned and readable
    uint32_t src_size;
    uint32_t gw, gh, faces;
    uint32_t* bb;         // gw entries of block-buffer scratch
    void* rec;            // faces * gh * gw records (uint32 for kinds 0/3, uint64 for kinds 1/2)
};

struct ParseGroup {
    const DevModel* models;   // M_COUNT models
    const uint16_t* sorted;   // M_COUNT x kMaxSyms canonical symbol order
    int32_t* status;          // set to CRN2_ECORRUPT on an out-of-range index
    uint32_t kind, n_ce, n_cs, n_ae, n_as;
    uint32_t first_job, njobs;
};

// ---------------------------------------------------------------- bit reader
struct DevBits {
    const uint32_t* wp;   // next aligned word to load
    const uint8_t* end;   // first byte past the stream
    uint64_t buf;         // valid bits MSB-aligned
    int cnt;

    __device__ inline void refill()
    {
        const uint8_t* wb = reinterpret_cast<const uint8_t*>(wp);
        uint32_t w = 0;
        if (wb + 4 <= end) {
            w = __builtin_bswap32(*wp);
        } else if (wb < end) {
            // bytes at or past the end read as zero (crn_decomp.h:3196-3207); the
            // aligned word holding the last byte lies inside the padded copy
            w = __builtin_bswap32(*wp) & (0xFFFFFFFFu << (8 * (4 - (int)(end - wb))));
        }
        buf |= (uint64_t)w << (32 - cnt);
        cnt += 32;
        wp++;
    }

    __device__ inline void start(const uint8_t* p, uint32_t size)
    {
        const uintptr_t a = reinterpret_cast<uintptr_t>(p);
        wp = reinterpret_cast<const uint32_t*>(a & ~(uintptr_t)3);
        end = p + size;
        buf = 0;
        cnt = 0;
        const int skip = (int)(a & 3) * 8;
        // load the first word but mark the bytes before p as consumed
        refill();
        buf <<= skip;
        cnt -= skip;
        if (cnt <= 32) refill();
    }

    __device__ inline uint32_t decode(const uint32_t* lut, const uint32_t* limit, const int32_t* base,
                                      const uint16_t* sorted, uint32_t nsyms)
    {
        if (cnt <= 32) refill();
        const uint32_t top = (uint32_t)(buf >> 48);
        uint32_t t = lut[top >> (16 - kLutBits)];
        uint32_t sym, len;
        if (t) {
            sym = t & 0xFFFFu;
            len = t >> 16;
        } else {
            // code longer than the table (or no code): canonical walk
            const uint32_t k = top + 1;
            len = kLutBits + 1;
            while (len <= 16 && k > limit[len]) len++;
            if (len > 16) return 0;
            const uint32_t idx = (uint32_t)(base[len] + (int32_t)(top >> (16 - len)));
            if (idx >= nsyms) return 0;
            sym = sorted[idx];
        }
        buf <<= len;
        cnt -= (int)len;
        return sym;
    }
};

struct LdsModel {
    uint32_t lut[kLutSize];
    uint32_t limit[17];
    int32_t base[17];
    uint32_t nsyms;
};

// Models used per kind: kind 0 {ref, ep0, sel0}; 1 {ref, ep0, sel0, ep1, sel1};
// 2 and 3 {ref, ep1, sel1}.  LDS slots: 0 = ref, 1 = colour ep, 2 = colour
// sel, 3 = alpha ep, 4 = alpha sel (slots 1,2 unused by kinds 2/3).
__global__ void __launch_bounds__(64) crnd_parse_kernel(const ParseGroup* groups, const ParseJob* jobs)
{
    extern __shared__ LdsModel lds[];
    const ParseGroup g = groups[blockIdx.x];
    const int kind = (int)g.kind;
    // stage the models this format uses
    const bool use_color = (kind == 0 || kind == 1);
    const bool use_alpha = (kind != 0);
    for (int m = 0; m < M_COUNT; m++) {
        const bool need = (m == M_REF) || ((m == M_EP0 || m == M_SEL0) && use_color) || ((m == M_EP1 || m == M_SEL1) && use_alpha);
        if (!need) continue;
        const DevModel& dm = g.models[m];
        for (uint32_t i = threadIdx.x; i < kLutSize; i += blockDim.x) lds[m].lut[i] = dm.lut[i];
        if (threadIdx.x < 17) {
            lds[m].limit[threadIdx.x] = dm.limit[threadIdx.x];
            lds[m].base[threadIdx.x] = dm.base[threadIdx.x];
        }
        if (threadIdx.x == 0) lds[m].nsyms = dm.num_syms;
    }
    __syncthreads();
    if (threadIdx.x >= g.njobs) return;
    const ParseJob j = jobs[g.first_job + threadIdx.x];

#define DEC(M) bits.decode(lds[M].lut, lds[M].limit, lds[M].base, g.sorted + (M) * kMaxSyms, lds[M].nsyms)

    DevBits bits;
    bits.start(j.src, j.src_size);
    const uint32_t gw = j.gw, gh = j.gh;
    uint32_t* __restrict__ bb = j.bb;
    uint32_t i0 = 0, i1 = 0;   // running endpoint indices: colour/alpha0 (kinds 0,1,3) ; alpha0/alpha1 (kind 2)
    uint32_t group = 0;
    bool bad = false;
    uint64_t blk = 0;
    // the block buffer starts zeroed for the first face of the level
    for (uint32_t x = 0; x < gw; x++) bb[x] = 0;
    const uint32_t nA = (kind == 1) ? g.n_ae : 0;
    const uint32_t n0 = (kind == 0 || kind == 1) ? g.n_ce : g.n_ae;
    const uint32_t n1 = (kind == 1 || kind == 2) ? g.n_ae : 1;
    (void)nA;
    for (uint32_t f = 0; f < j.faces; f++) {
        for (uint32_t y = 0; y < gh; y++) {
            uint32_t cur = bb[0];
            for (uint32_t x = 0; x < gw; x++, blk++) {
                const uint32_t nxt = (x + 1 < gw) ? bb[x + 1] : 0u;   // prefetch for the next block
                uint32_t ref, below;
                if (!(y & 1)) {
                    if (!(x & 1)) group = DEC(M_REF) & 0xFFu;
                    ref = group & 3u;
                    below = (group >> 2) & 3u;
                    group >>= 4;
                } else {
                    ref = cur & 3u;
                    below = ref;
                }
                if (ref == 0) {
                    if (kind == 0 || kind == 1) {
                        i0 += DEC(M_EP0);
                        if (i0 >= n0) i0 -= n0;
                    } else {
                        i0 += DEC(M_EP1);
                        if (i0 >= n0) i0 -= n0;
                    }
                    if (kind == 1 || kind == 2) {
                        i1 += DEC(M_EP1);
                        if (i1 >= n1) i1 -= n1;
                    }
                } else if (ref >= 2) {
                    i0 = (cur >> 2) & 0x3FFFu;
                    i1 = (cur >> 16) & 0x3FFFu;
                }
                // buffer entry: reference for the row below (even rows) or kept (odd rows) + indices
                bb[x] = below | (i0 << 2) | (i1 << 16);
                if (kind == 0) {
                    const uint32_t s0 = DEC(M_SEL0);
                    bad |= (i0 >= g.n_ce) | (s0 >= g.n_cs);
                    static_cast<uint32_t*>(j.rec)[blk] = i0 | (s0 << 16);
                } else if (kind == 3) {
                    const uint32_t s0 = DEC(M_SEL1);
                    bad |= (i0 >= g.n_ae) | (s0 >= g.n_as);
                    static_cast<uint32_t*>(j.rec)[blk] = i0 | (s0 << 16);
                } else if (kind == 1) {
                    const uint32_t s0 = DEC(M_SEL0);
                    const uint32_t s1 = DEC(M_SEL1);
                    bad |= (i0 >= g.n_ce) | (s0 >= g.n_cs) | (i1 >= g.n_ae) | (s1 >= g.n_as);
                    static_cast<uint64_t*>(j.rec)[blk] =
                        (uint64_t)(i0 | (s0 << 16)) | ((uint64_t)(i1 | (s1 << 16)) << 32);
                } else {
                    const uint32_t s0 = DEC(M_SEL1);
                    const uint32_t s1 = DEC(M_SEL1);
                    bad |= (i0 >= g.n_ae) | (s0 >= g.n_as) | (i1 >= g.n_ae) | (s1 >= g.n_as);
                    static_cast<uint64_t*>(j.rec)[blk] =
                        (uint64_t)(i0 | (s0 << 16)) | ((uint64_t)(i1 | (s1 << 16)) << 32);
                }
                cur = nxt;
            }
        }
    }
#undef DEC
    if (bad) atomicExch(g.status, CRN2_ECORRUPT);
}

struct WriteJob {
    const void* rec;        // records of the level (padded grid, faces stacked)
    uint8_t* dst[6];        // per-face destination
    uint32_t pitch;         // bytes between block rows
    uint32_t bx, by;        // visible blocks
    uint32_t gw, gh;        // padded grid
    uint32_t faces;
    uint64_t first_block;   // prefix of visible blocks over the jobs of the launch
};

struct WritePalettes {
    const uint32_t* ce;
    const uint32_t* cs;
    const uint16_t* ae;
    const uint16_t* as;     // 3 per entry
    uint32_t kind;
};

__global__ void __launch_bounds__(256) crnd_write_kernel(const WriteJob* jobs, uint32_t njobs, const WritePalettes* pals,
                                                         const uint32_t* job_pal, uint64_t total)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    // locate the job (few jobs per launch: linear scan from a binary search would do; levels are <= 16 per texture)
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].first_block <= t) lo = mid;
        else hi = mid;
    }
    const WriteJob& j = jobs[lo];
    const WritePalettes& P = pals[job_pal[lo]];
    uint64_t v = t - j.first_block;
    const uint32_t per_face = j.bx * j.by;
    const uint32_t f = (uint32_t)(v / per_face);
    v -= (uint64_t)f * per_face;
    const uint32_t y = (uint32_t)(v / j.bx), x = (uint32_t)(v % j.bx);
    const uint64_t r = (uint64_t)(f * j.gh + y) * j.gw + x;
    uint8_t* out = j.dst[f] + (uint64_t)y * j.pitch;
    if (P.kind == 0) {
        const uint32_t rec = static_cast<const uint32_t*>(j.rec)[r];
        uint2 o;
        o.x = P.ce[rec & 0xFFFFu];
        o.y = P.cs[rec >> 16];
        reinterpret_cast<uint2*>(out)[x] = o;
    } else if (P.kind == 3) {
        const uint32_t rec = static_cast<const uint32_t*>(j.rec)[r];
        const uint16_t* s = P.as + 3 * (rec >> 16);
        uint2 o;
        o.x = (uint32_t)P.ae[rec & 0xFFFFu] | ((uint32_t)s[0] << 16);
        o.y = (uint32_t)s[1] | ((uint32_t)s[2] << 16);
        reinterpret_cast<uint2*>(out)[x] = o;
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
        // the row pitch is only guaranteed to be a multiple of 4
        if ((j.pitch & 15) == 0) {
            reinterpret_cast<uint4*>(out)[x] = o;
        } else {
            uint32_t* w = reinterpret_cast<uint32_t*>(out + 16 * (uint64_t)x);
            w[0] = o.x; w[1] = o.y; w[2] = o.z; w[3] = o.w;
        }
    }
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
    // device
    uint8_t* d_file = nullptr;   // file copy padded to a multiple of 4 (+16)
    DevModel* d_models = nullptr;
    uint16_t* d_sorted = nullptr;
    uint32_t* d_ce = nullptr;
    uint32_t* d_cs = nullptr;
    uint16_t* d_ae = nullptr;
    uint16_t* d_as = nullptr;
    int32_t* d_status = nullptr;
    // scratch: records for every level and a block buffer per level
    void* d_rec = nullptr;
    uint32_t* d_bb = nullptr;
    uint64_t rec_ofs[kMaxLevels + 1] = {};
    uint64_t bb_ofs[kMaxLevels + 1] = {};
    ParseGroup* d_group = nullptr;
    ParseJob* d_jobs = nullptr;
    WriteJob* d_wjobs = nullptr;
    WritePalettes* d_pal = nullptr;
    uint32_t* d_jobpal = nullptr;
    uint32_t n_ce = 0, n_cs = 0, n_ae = 0, n_as = 0;
};

namespace {

constexpr uint32_t kCtxMagic = 0x1EF9CABDu;

static uint32_t level_dim(uint32_t v, uint32_t l) { return std::max<uint32_t>(1u, v >> l); }

static uint32_t bytes_per_block(uint32_t fmt)
{
    return (fmt == CRN2_FMT_DXT1 || fmt == CRN2_FMT_DXT5A) ? 8u : 16u;
}

#define HIPCHK(x)                                   \
    do {                                            \
        if ((x) != hipSuccess) return CRN2_EHIP;    \
    } while (0)

// Palette decoding on the host during unpack_begin (crn_decomp.h:3715-3851).
static bool decode_palettes(crn2_unpack_ctx* c, std::vector<uint32_t>& ce, std::vector<uint32_t>& cs,
                            std::vector<uint16_t>& ae, std::vector<uint16_t>& as)
{
    const Header& h = c->h;
    HostBits b;
    if (h.pal_num[0]) {
        const uint32_t n = h.pal_num[0];
        if (!h.pal_size[0] || h.pal_ofs[0] + h.pal_size[0] > c->size) return false;
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
            if (!h.pal_size[1] || h.pal_ofs[1] + h.pal_size[1] > c->size) return false;
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
        if (!h.pal_size[2] || h.pal_ofs[2] + h.pal_size[2] > c->size) return false;
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
            if (!h.pal_size[3] || h.pal_ofs[3] + h.pal_size[3] > c->size) return false;
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

static void ctx_free_device(crn2_unpack_ctx* c)
{
    void* ptrs[] = {c->d_file, c->d_models, c->d_sorted, c->d_ce, c->d_cs, c->d_ae, c->d_as, c->d_status,
                    c->d_rec, c->d_bb, c->d_group, c->d_jobs, c->d_wjobs, c->d_pal, c->d_jobpal};
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
        for (auto& m : c->models) m = HostModel();
        delete c;
        return nullptr;
    };
    if (!parse_header(c->data, size, c->h)) return fail();
    c->kind = format_kind(c->h.format);
    if (c->kind < 0) return fail();
    const Header& h = c->h;
    if (h.levels < 1 || h.levels > kMaxLevels || (h.faces != 1 && h.faces != 6) || !h.width || !h.height) return fail();
    // tables: crn_decomp.h:3662-3692
    if (!h.tables_size || h.tables_ofs + h.tables_size > size) return fail();
    HostBits b;
    b.start(c->data + h.tables_ofs, h.tables_size);
    if (!receive_model(b, c->models[M_REF])) return fail();
    if (!h.pal_num[0] && !h.pal_num[2]) return fail();
    if (h.pal_num[0] && (!receive_model(b, c->models[M_EP0]) || !receive_model(b, c->models[M_SEL0]))) return fail();
    if (h.pal_num[2] && (!receive_model(b, c->models[M_EP1]) || !receive_model(b, c->models[M_SEL1]))) return fail();
    // the models the format walks with must exist
    if (!c->models[M_REF].valid) return fail();
    if ((c->kind == 0 || c->kind == 1) && (!c->models[M_EP0].valid || !c->models[M_SEL0].valid)) return fail();
    if (c->kind != 0 && (!c->models[M_EP1].valid || !c->models[M_SEL1].valid)) return fail();
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

    // device state
    std::vector<DevModel> dms(M_COUNT);
    std::vector<uint16_t> sorted(M_COUNT * kMaxSyms, 0);
    for (int m = 0; m < M_COUNT; m++) {
        if (c->models[m].valid) {
            to_device_model(c->models[m], dms[m]);
            std::copy(c->models[m].sorted.begin(), c->models[m].sorted.end(), sorted.begin() + m * kMaxSyms);
        } else {
            memset(&dms[m], 0, sizeof(DevModel));
        }
    }
    // scratch sizes
    uint64_t rec = 0, bbn = 0;
    const uint32_t rec_bytes = (c->kind == 1 || c->kind == 2) ? 8 : 4;
    for (uint32_t l = 0; l < h.levels; l++) {
        const uint32_t bx = (level_dim(h.width, l) + 3) >> 2, by = (level_dim(h.height, l) + 3) >> 2;
        const uint32_t gw = (bx + 1) & ~1u, gh = (by + 1) & ~1u;
        c->rec_ofs[l] = rec;
        c->bb_ofs[l] = bbn;
        rec += (uint64_t)gw * gh * h.faces * rec_bytes;
        bbn += gw;
    }
    c->rec_ofs[h.levels] = rec;
    c->bb_ofs[h.levels] = bbn;

    const size_t padded = ((size_t)size + 3) / 4 * 4 + 16;
    // the device copy is padded so that aligned word reads never leave it
    if (dmalloc(&c->d_file, padded) != hipSuccess) return fail();
    if (dmalloc(&c->d_models, sizeof(DevModel) * M_COUNT) != hipSuccess ||
        dmalloc(&c->d_sorted, sorted.size() * 2) != hipSuccess || dmalloc(&c->d_ce, ce.size() * 4) != hipSuccess ||
        dmalloc(&c->d_cs, cs.size() * 4) != hipSuccess || dmalloc(&c->d_ae, ae.size() * 2) != hipSuccess ||
        dmalloc(&c->d_as, as.size() * 2) != hipSuccess || dmalloc(&c->d_status, 4) != hipSuccess ||
        dmalloc(&c->d_rec, rec) != hipSuccess || dmalloc(&c->d_bb, bbn * 4) != hipSuccess ||
        dmalloc(&c->d_group, sizeof(ParseGroup)) != hipSuccess ||
        dmalloc(&c->d_jobs, sizeof(ParseJob) * kMaxLevels) != hipSuccess ||
        dmalloc(&c->d_wjobs, sizeof(WriteJob) * kMaxLevels) != hipSuccess ||
        dmalloc(&c->d_pal, sizeof(WritePalettes)) != hipSuccess ||
        dmalloc(&c->d_jobpal, 4 * kMaxLevels) != hipSuccess)
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
    WritePalettes P{c->d_ce, c->d_cs, c->d_ae, c->d_as, (uint32_t)c->kind};
    ok &= hipMemcpyAsync(c->d_pal, &P, sizeof(P), hipMemcpyHostToDevice, stream) == hipSuccess;
    ok &= hipStreamSynchronize(stream) == hipSuccess;
    if (!ok) return fail();
    c->magic = kCtxMagic;
    return c;
}

static bool ctx_ok(const crn2_unpack_ctx* c) { return c && c->magic == kCtxMagic; }

// Queues parse + write for levels [first, first + count).  dst(l, f) gives the
// destination of face f of level l, pitch(l) its row pitch.
template <typename DstFn, typename PitchFn>
static int launch_levels(crn2_unpack_ctx* c, uint32_t first, uint32_t count, const uint8_t* level_src_override,
                         uint32_t override_size, DstFn dst, PitchFn pitch, hipStream_t stream)
{
    const Header& h = c->h;
    ParseJob jobs[kMaxLevels];
    WriteJob wj[kMaxLevels];
    uint32_t jp[kMaxLevels] = {0};
    uint64_t total = 0;
    for (uint32_t i = 0; i < count; i++) {
        const uint32_t l = first + i;
        const uint32_t bx = (level_dim(h.width, l) + 3) >> 2, by = (level_dim(h.height, l) + 3) >> 2;
        const uint32_t gw = (bx + 1) & ~1u, gh = (by + 1) & ~1u;
        const uint32_t ofs = h.level_ofs[l];
        const uint32_t next = (l + 1 < h.levels) ? h.level_ofs[l + 1] : c->size;
        ParseJob& j = jobs[i];
        if (level_src_override) {
            j.src = level_src_override;
            j.src_size = override_size;
        } else {
            j.src = c->d_file + ofs;
            j.src_size = next - ofs;
        }
        j.gw = gw;
        j.gh = gh;
        j.faces = h.faces;
        j.bb = c->d_bb + c->bb_ofs[l];
        j.rec = static_cast<uint8_t*>(c->d_rec) + c->rec_ofs[l];
        WriteJob& w = wj[i];
        w.rec = j.rec;
        for (uint32_t f = 0; f < 6; f++) w.dst[f] = (f < h.faces) ? dst(l, f) : nullptr;
        w.pitch = pitch(l);
        w.bx = bx;
        w.by = by;
        w.gw = gw;
        w.gh = gh;
        w.faces = h.faces;
        w.first_block = total;
        total += (uint64_t)bx * by * h.faces;
    }
    ParseGroup g;
    g.models = c->d_models;
    g.sorted = c->d_sorted;
    g.status = c->d_status;
    g.kind = (uint32_t)c->kind;
    g.n_ce = c->n_ce;
    g.n_cs = c->n_cs;
    g.n_ae = c->n_ae;
    g.n_as = c->n_as;
    g.first_job = 0;
    g.njobs = count;
    HIPCHK(hipMemcpyAsync(c->d_group, &g, sizeof(g), hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(c->d_jobs, jobs, sizeof(ParseJob) * count, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(c->d_wjobs, wj, sizeof(WriteJob) * count, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync(c->d_jobpal, jp, 4 * count, hipMemcpyHostToDevice, stream));
    const size_t lds = sizeof(LdsModel) * M_COUNT;
    hipLaunchKernelGGL(crnd_parse_kernel, dim3(1), dim3(64), lds, stream, c->d_group, c->d_jobs);
    HIPCHK(hipGetLastError());
    if (total) {
        const uint32_t tb = 256;
        const uint64_t nb = (total + tb - 1) / tb;
        hipLaunchKernelGGL(crnd_write_kernel, dim3((uint32_t)nb), dim3(tb), 0, stream, c->d_wjobs, count, c->d_pal,
                           c->d_jobpal, total);
        HIPCHK(hipGetLastError());
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

static int read_status(crn2_unpack_ctx* c, hipStream_t stream)
{
    int32_t st = 0;
    HIPCHK(hipMemcpyAsync(&st, c->d_status, 4, hipMemcpyDeviceToHost, stream));
    HIPCHK(hipStreamSynchronize(stream));
    if (st) {
        HIPCHK(hipMemsetAsync(c->d_status, 0, 4, stream));
        HIPCHK(hipStreamSynchronize(stream));
    }
    return st;
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
    rc = CRN2_OK;
    if (src_host) {
        const size_t pad = ((size_t)src_size + 3) / 4 * 4 + 16;
        if (dmalloc(&d_src, pad) != hipSuccess) rc = CRN2_ENOMEM;
        else if (hipMemsetAsync(d_src, 0, pad, stream) != hipSuccess ||
                 hipMemcpyAsync(d_src, src_host, src_size, hipMemcpyHostToDevice, stream) != hipSuccess)
            rc = CRN2_EHIP;
    }
    // the caller's buffer may hold bytes the transcoder does not write (pitch padding)
    for (uint32_t f = 0; f < h.faces && !rc; f++)
        if (hipMemcpyAsync(d_out + f * face_bytes, dst[f], face_bytes, hipMemcpyHostToDevice, stream) != hipSuccess)
            rc = CRN2_EHIP;
    if (!rc)
        rc = launch_levels(
            c, level, 1, d_src, src_size, [&](uint32_t, uint32_t f) { return d_out + f * face_bytes; },
            [&](uint32_t) { return row_pitch; }, stream);
    if (!rc) rc = read_status(c, stream);
    for (uint32_t f = 0; f < h.faces && !rc; f++)
        if (hipMemcpyAsync(dst[f], d_out + f * face_bytes, face_bytes, hipMemcpyDeviceToHost, stream) != hipSuccess)
            rc = CRN2_EHIP;
    if (hipStreamSynchronize(stream) != hipSuccess && !rc) rc = CRN2_EHIP;
    (void)hipFree(d_out);
    if (d_src) (void)hipFree(d_src);
    return rc;
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
    return launch_levels(
        c, level, 1, nullptr, 0, [&](uint32_t, uint32_t f) { return static_cast<uint8_t*>(dst_dev[f]); },
        [&](uint32_t) { return row_pitch; }, static_cast<hipStream_t>(stream));
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
    uint64_t lofs[kMaxLevels] = {0};
    uint64_t acc = 0;
    for (uint32_t l = first; l < first + count; l++) {
        lofs[l] = acc;
        acc += crn2_levels_size(c, l, 1);
    }
    const uint32_t bpb = bytes_per_block(c->h.format);
    uint8_t* base = static_cast<uint8_t*>(dst_dev);
    return launch_levels(
        c, first, count, nullptr, 0,
        [&](uint32_t l, uint32_t f) {
            const uint32_t bx = (level_dim(c->h.width, l) + 3) >> 2, by = (level_dim(c->h.height, l) + 3) >> 2;
            return base + lofs[l] + (uint64_t)f * bx * by * bpb;
        },
        [&](uint32_t l) { return ((level_dim(c->h.width, l) + 3) >> 2) * bpb; }, static_cast<hipStream_t>(stream));
}

int crn2_unpack_end(crn2_unpack_ctx* c)
{
    if (!ctx_ok(c)) return CRN2_EINVAL;
    c->magic = 0;
    ctx_free_device(c);
    for (auto& m : c->models) m = HostModel();
    delete c;
    return CRN2_OK;
}

}  // extern "C"
