// mip_gpu.hip -- mip generation on the GPU: mipmapped_texture::
// generate_mipmaps (crnlib/crn_mipmapped_texture.cpp:2140-2220) with
// image_utils::resample (crnlib/crn_image_utils.cpp:478-865) and crnlib's
// separable Resampler (crnlib/crn_resampler.cpp), bit-exact.
//
// Every level is resampled from level 0 in two passes.  A pass is a set of
// sequential dot products: output o of line k is
//     s = 0;  s += in[k][p] * w  for (p, w) in list(o), in list order,
// (the reference's Y pass starts from the first product, which differs from
// 0 + it only in the sign of a zero) with separate float multiplies and adds
// (-ffp-contract=off), so every sum rounds exactly as the reference's.  A
// list is stored as runs of contributors: a run of one texel repeated (the
// clamped edge of a wide filter: most of the contributors of the deepest
// levels) loads it once; a run of consecutive texels walks them.
// The order of the passes is the reference's choice per level: X then Y
// (always, for the multithreaded resampler) or, where the single-thread
// Resampler delays X, Y then X.  Y then X is run as X then Y on the
// transposed image -- the same sums in the same order.
//
// Layout in HBM (one plane per component c):
//   level-0 planes ... bytes, lx[c][x][y] (column-major: the lanes of a wave
//                      are 64 consecutive rows, so every contributor column is
//                      one coalesced 64-byte load) and/or ly[c][y][x]; every
//                      level re-reads them, so bytes rather than floats (the
//                      sRGB table sits in LDS) cut the dominant traffic 4x;
//   pass-1 output .... inter[c][line][o], o along the first pass's axis;
//   pass 2 ........... lanes = 64 consecutive o, one wave per component, the
//                      contributor list of the block's output line in scalar
//                      registers; clamp to [0, 1], (int)(255 x + .5) or the
//                      linear-to-sRGB table, optional renorm_normal_map, one
//                      RGBA8 texel per lane.
// Both passes cover every level in one launch, the deepest levels first:
// their sums are the longest serial chains (level l has about 5.4 * 2^l
// contributors per output with the default Kaiser filter) and start while the
// wide levels fill the rest of the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <new>
#include <vector>

#include "crn2_internal.h"
#include "mip_plan.h"

namespace {

// pass-1 outputs per lane: a lane's outputs are contiguous in its line of
// the intermediate, so 16 of them leave as four 16-byte stores (lanes are
// lines, a line apart); a level whose outputs have long lists takes one per
// lane (shorter serial chains, more waves)
constexpr uint32_t kOutPerLane = 16;
constexpr uint32_t kLongList = 256;

struct Pass1Level {
    uint32_t unit0;   // first wave of this level
    uint32_t nlb;     // blocks of 64 lines
    uint32_t per;     // outputs per lane (1 or kOutPerLane)
    uint32_t runs;    // groups of `per` outputs
    uint32_t lines;   // lines (rows for X first, columns for Y first)
    uint32_t d;       // outputs per line
    uint32_t s;       // source samples per line
    uint32_t list;    // index of the list's offs[0]
    uint32_t layout;  // 0 = lx, 1 = ly
    uint64_t inter;   // float offset of this level's pass-1 output
};

struct Pass2Level {
    uint32_t block0;  // first block of this level
    uint32_t nxb;     // blocks of 64 pass-1 outputs
    uint32_t d2;      // output lines of pass 2
    uint32_t lines;   // contributors' range (pass-1 lines)
    uint32_t d;       // pass-1 outputs per line = lanes
    uint32_t list;
    uint32_t os_lane, os_o;  // texel strides of the output: per lane, per pass-2 line
    uint64_t inter;
    uint64_t out;  // texel offset of this level in the chain
};

// contributors pix[k] .. pix[k] + (n - 1) * step with weights wt[w .. w + n)
struct Run {
    uint32_t pix, n, step, w;
};

struct Tables {
    const Pass1Level* p1;
    const Pass2Level* p2;
    const uint32_t* sofs;  // runs of output o of a list: sofs[list + o] .. sofs[list + o + 1]
    const Run* runs;
    const float* wt;
    const float* s2l;
    const uint8_t* l2s;
};

__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float unif(float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); }

// Source samples: float (the intermediate), or level-0 bytes converted as
// resample_single_thread does (crn_image_utils.cpp:575-590): colour through
// srgb_to_linear when sRGB (the 256 floats staged in LDS), alpha and linear
// colour v * (1/255).
enum Src { kFloat, kLinear8, kTable8 };

template <Src S>
__device__ __forceinline__ float sample(const void* base, uint32_t i, const float* tab)
{
    if constexpr (S == kFloat) {
        return static_cast<const float*>(base)[i];
    } else {
        const uint32_t v = static_cast<const uint8_t*>(base)[i];
        if constexpr (S == kTable8) return tab[v];
        else return __fmul_rn((float)v, 1.0f / 255.0f);
    }
}

// s + src(p * stride + lane) * w over the runs [rb, re) of one output, in
// order.  Lists, runs and weights are wave-uniform (scalar loads); each
// contributor of a consecutive run is one coalesced load of 64 samples, 16 in
// flight.  A batch past the end of a run repeats its last texel with weight
// 0: s + v * 0 == s (up to the sign of a zero sum), so the sums are the
// reference's.  The weights of a batch are read as 16 consecutive floats
// (the table buffer continues past the last list).
template <Src S>
__device__ __forceinline__ float dot_runs(const Tables& t, uint32_t rb, uint32_t re, const void* base,
                                          uint32_t lane, uint32_t stride, float s, const float* tab)
{
    for (uint32_t g = rb; g < re; g++) {
        const Run r = t.runs[g];
        const uint32_t n = uni(r.n), p = uni(r.pix);
        const float* wp = t.wt + uni(r.w);
        if (uni(r.step) == 0) {
            const float v = sample<S>(base, p * stride + lane, tab);
            uint32_t k = 0;
            for (; k + 16 <= n; k += 16) {
#pragma unroll
                for (int i = 0; i < 16; i++) s = __fadd_rn(s, __fmul_rn(v, unif(wp[k + i])));
            }
            for (; k < n; k++) s = __fadd_rn(s, __fmul_rn(v, unif(wp[k])));
        } else {
            for (uint32_t k = 0; k < n; k += 16) {
                float v[16];
#pragma unroll
                for (int i = 0; i < 16; i++) v[i] = sample<S>(base, (p + min(k + i, n - 1)) * stride + lane, tab);
#pragma unroll
                for (int i = 0; i < 16; i++) {
                    const float w = unif(wp[k + i]);
                    s = __fadd_rn(s, __fmul_rn(v[i], k + i < n ? w : 0.0f));
                }
            }
        }
    }
    return s;
}

// ------------------------------------------------------------ level-0 planes

// RGBA8 -> one byte plane per component, column-major (lx) and/or row-major
// (ly); 32x32 tiles, the column-major planes written from the transposed tile.
__global__ void __launch_bounds__(256) mip_planes_kernel(const uint32_t* __restrict__ rgba, uint32_t w, uint32_t h,
                                                         uint8_t* __restrict__ lx, uint8_t* __restrict__ ly,
                                                         uint32_t ncomp)
{
    __shared__ uint32_t tile[32][33];
    const uint32_t tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    const uint32_t x0 = blockIdx.x * 32, y0 = blockIdx.y * 32;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t y = y0 + ty + 8 * i, x = x0 + tx;
        const bool in = x < w && y < h;
        const uint32_t p = in ? rgba[(size_t)y * w + x] : 0;
        tile[ty + 8 * i][tx] = p;
        if (ly && in)
            for (uint32_t c = 0; c < ncomp; c++) ly[((size_t)c * h + y) * w + x] = (uint8_t)(p >> (8 * c));
    }
    if (!lx) return;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const uint32_t x = x0 + ty + 8 * i, y = y0 + tx;
        if (x < w && y < h) {
            const uint32_t p = tile[tx][ty + 8 * i];
            for (uint32_t c = 0; c < ncomp; c++) lx[((size_t)c * w + x) * h + y] = (uint8_t)(p >> (8 * c));
        }
    }
}

// ------------------------------------------------------------------- pass 1

// One wave per (level, component, 64 lines, 1 or kOutPerLane outputs); lanes
// are lines.  Results wait in LDS (each lane's own) and leave as rows.
__global__ void __launch_bounds__(256) mip_pass1_kernel(Tables t, const uint8_t* __restrict__ lx,
                                                        const uint8_t* __restrict__ ly, float* __restrict__ inter,
                                                        uint32_t nlevels, uint32_t units, uint32_t ncomp, int srgb)
{
    __shared__ float stage[4][kOutPerLane][64];
    __shared__ float tab[256];
    tab[threadIdx.x] = t.s2l[threadIdx.x];
    __syncthreads();
    const uint32_t u = uni(blockIdx.x * 4 + (threadIdx.x >> 6));
    if (u >= units) return;
    uint32_t l = 0;
    while (l + 1 < nlevels && t.p1[l + 1].unit0 <= u) l++;
    const Pass1Level L = t.p1[l];
    uint32_t rem = u - L.unit0;
    const uint32_t run = rem % L.runs;
    rem /= L.runs;
    const uint32_t lb = rem % L.nlb, c = rem / L.nlb;
    const uint32_t line = lb * 64 + (threadIdx.x & 63);
    const bool valid = line < L.lines;
    const uint32_t rl = valid ? line : L.lines - 1;
    const uint8_t* base = (L.layout ? ly : lx) + (size_t)c * L.s * L.lines;
    float* out = inter + L.inter + ((size_t)c * L.lines + rl) * L.d + run * L.per;
    const uint32_t nout = min(L.per, L.d - run * L.per);
    const uint32_t w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const bool table = srgb && c < 3;
    for (uint32_t r = 0; r < nout; r++) {
        const uint32_t o = run * L.per + r;
        const uint32_t rb = uni(t.sofs[L.list + o]), re = uni(t.sofs[L.list + o + 1]);
        stage[w][r][lane] = table ? dot_runs<kTable8>(t, rb, re, base, rl, L.lines, 0.0f, tab)
                                  : dot_runs<kLinear8>(t, rb, re, base, rl, L.lines, 0.0f, tab);
    }
    if (!valid) return;
    // each lane reads back only what it wrote
    if (nout == kOutPerLane && !(reinterpret_cast<uintptr_t>(out) & 15)) {
#pragma unroll
        for (uint32_t q = 0; q < kOutPerLane; q += 4)
            *reinterpret_cast<float4*>(out + q) =
                make_float4(stage[w][q][lane], stage[w][q + 1][lane], stage[w][q + 2][lane], stage[w][q + 3][lane]);
    } else {
        for (uint32_t r = 0; r < nout; r++) out[r] = stage[w][r][lane];
    }
}

// ------------------------------------------------------------------- pass 2

__device__ __forceinline__ float clamp01(float x)
{
    if (x < 0.0f) x = 0.0f;
    else if (x > 1.0f) x = 1.0f;
    return x;
}

// image_utils::renorm_normal_map for one texel (crn_image_utils.cpp:369-420)
__device__ void renorm(uint32_t* c)
{
    if (c[0] == 128 && c[1] == 128 && c[2] == 128) return;
    float v[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        float f = (float)c[i];
        f = __fmul_rn(f, 1.0f / 255.0f);
        f = __fmul_rn(f, 2.0f);
        f = f - 1.0f;
        if (f < -1.0f) f = -1.0f;
        else if (f > 1.0f) f = 1.0f;
        v[i] = f;
    }
    float sum = __fmul_rn(v[0], v[0]);
    sum = __fadd_rn(sum, __fmul_rn(v[1], v[1]));
    sum = __fadd_rn(sum, __fmul_rn(v[2], v[2]));
    const float len = sqrtf(sum);
    if (len < .077f) {
        c[0] = c[1] = c[2] = 128;
    } else if (fabsf(len - 1.0f) > .077f) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const float n = v[i] / len;
            float f = floorf(__fadd_rn(__fmul_rn(__fmul_rn(n + 1.0f, .5f), 255.0f), .5f));
            f = f < 0.0f ? 0.0f : (f > 255.0f ? 255.0f : f);
            c[i] = (uint32_t)f;
        }
        if (c[0] == 128 && c[1] == 128) c[2] = c[2] < 128 ? 0 : 255;
    }
}

// One block per (level, output line of pass 2, 64 lanes), one wave per
// component; texels packed by wave 0.
__global__ void __launch_bounds__(256) mip_pass2_kernel(Tables t, const float* __restrict__ inter,
                                                        uint32_t* __restrict__ out, uint32_t nlevels,
                                                        uint32_t ncomp, int srgb, int renormalize)
{
    __shared__ uint32_t px[4][64];
    const uint32_t b = blockIdx.x;
    uint32_t l = 0;
    while (l + 1 < nlevels && t.p2[l + 1].block0 <= b) l++;
    const Pass2Level L = t.p2[l];
    const uint32_t rem = b - L.block0;
    const uint32_t xb = rem % L.nxb, o2 = rem / L.nxb;
    const uint32_t c = uni(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const uint32_t x = xb * 64 + lane;
    const bool valid = x < L.d;
    const uint32_t rx = valid ? x : L.d - 1;
    uint32_t val = 255;
    if (c < ncomp) {
        const float* base = inter + L.inter + (size_t)c * L.lines * L.d;
        float s = dot_runs<kFloat>(t, uni(t.sofs[L.list + o2]), uni(t.sofs[L.list + o2 + 1]), base, rx, L.d, 0.0f,
                                   nullptr);
        s = clamp01(s);
        if (!srgb || c == 3) {
            const int v = (int)__fadd_rn(__fmul_rn(255.0f, s), .5f);
            val = (uint32_t)min(255, max(0, v));
        } else {
            const int j = (int)__fadd_rn(__fmul_rn(8192.0f, s), .5f);
            val = t.l2s[min(8191, max(0, j))];
        }
    }
    px[c][lane] = val;
    __syncthreads();
    if (c == 0 && valid) {
        uint32_t q[4] = {px[0][lane], px[1][lane], px[2][lane], px[3][lane]};
        if (renormalize) renorm(q);
        out[L.out + (size_t)o2 * L.os_o + (size_t)x * L.os_lane] = q[0] | q[1] << 8 | q[2] << 16 | q[3] << 24;
    }
}

bool hip_ok(hipError_t e) { return e == hipSuccess; }

}  // namespace

struct crn2_mip_plan {
    uint32_t magic;
    uint32_t width, height, levels, ncomp;
    crn2_mip_params params;
    std::vector<uint64_t> level_ofs;  // bytes
    std::vector<uint8_t> y_first;
    uint64_t chain_bytes;
    // device
    void* tables = nullptr;
    void* scratch = nullptr;
    Tables t{};
    uint8_t *lx = nullptr, *ly = nullptr;  // level-0 byte planes
    float* inter = nullptr;
    uint32_t units1 = 0, blocks2 = 0;
    // stage timing (HIP events on the run stream): planes, pass 1, pass 2
    bool prof = false;
    hipEvent_t ev[4] = {};
    double ms[3] = {};
    uint64_t runs = 0;
};

namespace {

constexpr uint32_t kPlanMagic = 0x4d495032u;

bool plan_ok(const crn2_mip_plan* p) { return p && p->magic == kPlanMagic; }

void plan_free(crn2_mip_plan* p)
{
    if (p->tables) (void)hipFree(p->tables);
    if (p->scratch) (void)hipFree(p->scratch);
    for (hipEvent_t& e : p->ev)
        if (e) (void)hipEventDestroy(e);
    p->magic = 0;
    delete p;
}

size_t align_up(size_t v, size_t a) { return (v + a - 1) / a * a; }

}  // namespace

extern "C" {

uint32_t crn2_mip_level_count(uint32_t width, uint32_t height, uint32_t min_mip_size, uint32_t max_levels)
{
    if (!width || !height) return 0;
    return crn2mip::level_count(width, height, min_mip_size, max_levels);
}

int crn2_mip_contributors(uint32_t src, uint32_t dst, int wrap, uint32_t filter, float filter_scale, uint32_t* offs,
                          uint32_t* pix, float* weight, uint32_t cap)
{
    if (!src || !dst || src > 16384 || dst > 16384 || (cap && (!pix || !weight))) return CRN2_EINVAL;
    crn2mip::List L;
    if (!crn2mip::make_list(src, dst, wrap != 0, filter, filter_scale, L)) return CRN2_EINVAL;
    if (offs) std::memcpy(offs, L.offs.data(), L.offs.size() * sizeof(uint32_t));
    const size_t n = std::min<size_t>(cap, L.pix.size());
    if (n) {
        std::memcpy(pix, L.pix.data(), n * sizeof(uint32_t));
        std::memcpy(weight, L.wt.data(), n * sizeof(float));
    }
    return (int)L.pix.size();
}

int crn2_mip_gamma_tables(float gamma, float* srgb_to_linear, uint8_t* linear_to_srgb)
{
    if (!srgb_to_linear || !linear_to_srgb || !(gamma > 0.0f)) return CRN2_EINVAL;
    crn2mip::gamma_tables(gamma, srgb_to_linear, linear_to_srgb);
    return CRN2_OK;
}

int crn2_mip_plan_create(uint32_t width, uint32_t height, const crn2_mip_params* params, crn2_mip_plan** out)
{
    if (!out || !params) return CRN2_EINVAL;
    *out = nullptr;
    if (params->flags & ~31u) return CRN2_EINVAL;
    std::vector<crn2mip::Level> lv;
    if (!crn2mip::plan_levels(width, height, *params, lv)) return CRN2_EINVAL;
    crn2_mip_plan* p = new (std::nothrow) crn2_mip_plan;
    if (!p) return CRN2_ENOMEM;
    p->magic = kPlanMagic;
    p->width = width;
    p->height = height;
    p->params = *params;
    p->levels = (uint32_t)lv.size() + 1;
    p->ncomp = (params->flags & CRN2_MIP_ALPHA) ? 4 : 3;
    p->level_ofs.assign(1, 0);
    p->y_first.assign(1, 0);
    uint64_t ofs = (uint64_t)width * height * 4;
    for (const auto& v : lv) {
        p->level_ofs.push_back(ofs);
        p->y_first.push_back(v.xfirst ? 0 : 1);
        ofs += (uint64_t)v.width * v.height * 4;
    }
    p->chain_bytes = ofs;
    if (lv.empty()) {
        *out = p;
        return CRN2_OK;
    }
    const uint32_t nl = (uint32_t)lv.size();
    const uint32_t nc = p->ncomp;
    // host image of the device tables
    std::vector<Pass1Level> p1(nl);
    std::vector<Pass2Level> p2(nl);
    std::vector<uint32_t> sofs;
    std::vector<Run> runs;
    std::vector<float> wt;
    bool need_lx = false, need_ly = false;
    uint64_t inter_floats = 0;
    uint32_t units = 0, blocks = 0;
    // a list as runs: one texel repeated (step 0) or consecutive texels (step 1)
    auto add_list = [&](const crn2mip::List& L) {
        const uint32_t at = (uint32_t)sofs.size(), base = (uint32_t)wt.size();
        wt.insert(wt.end(), L.wt.begin(), L.wt.end());
        for (size_t o = 0; o + 1 < L.offs.size(); o++) {
            sofs.push_back((uint32_t)runs.size());
            for (uint32_t k = L.offs[o], end = L.offs[o + 1]; k < end;) {
                const uint32_t p = L.pix[k];
                uint32_t n = 1, step = 1;
                if (k + 1 < end && L.pix[k + 1] == p) {
                    step = 0;
                    while (k + n < end && L.pix[k + n] == p) n++;
                } else {
                    while (k + n < end && L.pix[k + n] == p + n) n++;
                }
                runs.push_back(Run{p, n, step, base + k});
                k += n;
            }
        }
        sofs.push_back((uint32_t)runs.size());
        return at;
    };
    // deepest level first in both launches
    for (uint32_t i = 0; i < nl; i++) {
        const crn2mip::Level& v = lv[nl - 1 - i];
        const uint32_t level = nl - i;  // 1-based chain index
        Pass1Level& a = p1[i];
        Pass2Level& b = p2[i];
        const crn2mip::List& first = v.xfirst ? v.lx : v.ly;
        const crn2mip::List& second = v.xfirst ? v.ly : v.lx;
        a.lines = v.xfirst ? height : width;
        a.s = v.xfirst ? width : height;
        a.d = v.xfirst ? v.width : v.height;
        a.layout = v.xfirst ? 0 : 1;
        (v.xfirst ? need_lx : need_ly) = true;
        a.nlb = (a.lines + 63) / 64;
        a.per = first.ops() / a.d >= kLongList ? 1 : kOutPerLane;
        a.runs = (a.d + a.per - 1) / a.per;
        a.unit0 = units;
        units += nc * a.nlb * a.runs;
        a.list = add_list(first);
        a.inter = inter_floats;
        inter_floats += (uint64_t)nc * a.lines * a.d;
        b.lines = a.lines;
        b.d = a.d;
        b.d2 = v.xfirst ? v.height : v.width;
        b.nxb = (b.d + 63) / 64;
        b.block0 = blocks;
        blocks += b.d2 * b.nxb;
        b.list = add_list(second);
        b.inter = a.inter;
        b.out = p->level_ofs[level] / 4;
        b.os_lane = v.xfirst ? 1 : v.width;
        b.os_o = v.xfirst ? v.width : 1;
    }
    p->units1 = units;
    p->blocks2 = blocks;
    float s2l[256];
    std::vector<uint8_t> l2s(8192);
    crn2mip::gamma_tables(params->source_gamma, s2l, l2s.data());
    // one device buffer: level descriptors, lists, gamma tables
    size_t at = 0;
    const size_t o_p1 = at;
    at = align_up(at + nl * sizeof(Pass1Level), 256);
    const size_t o_p2 = at;
    at = align_up(at + nl * sizeof(Pass2Level), 256);
    const size_t o_sofs = at;
    at = align_up(at + sofs.size() * 4, 256);
    const size_t o_runs = at;
    at = align_up(at + runs.size() * sizeof(Run), 256);
    const size_t o_wt = at;
    at = align_up(at + wt.size() * 4 + 64, 256);  // batches read up to 15 weights past a list
    const size_t o_s2l = at;
    at = align_up(at + sizeof(s2l), 256);
    const size_t o_l2s = at;
    at = align_up(at + l2s.size(), 256);
    std::vector<uint8_t> host(at, 0);
    std::memcpy(host.data() + o_p1, p1.data(), nl * sizeof(Pass1Level));
    std::memcpy(host.data() + o_p2, p2.data(), nl * sizeof(Pass2Level));
    std::memcpy(host.data() + o_sofs, sofs.data(), sofs.size() * 4);
    std::memcpy(host.data() + o_runs, runs.data(), runs.size() * sizeof(Run));
    std::memcpy(host.data() + o_wt, wt.data(), wt.size() * 4);
    std::memcpy(host.data() + o_s2l, s2l, sizeof(s2l));
    std::memcpy(host.data() + o_l2s, l2s.data(), l2s.size());
    const uint64_t plane = align_up((uint64_t)nc * width * height, 256);  // bytes
    const uint64_t scratch_bytes = (need_lx ? plane : 0) + (need_ly ? plane : 0) + inter_floats * sizeof(float);
    int rc = CRN2_OK;
    if (!hip_ok(hipMalloc(&p->tables, at)) || !hip_ok(hipMalloc(&p->scratch, scratch_bytes)))
        rc = CRN2_ENOMEM;
    else if (!hip_ok(hipMemcpy(p->tables, host.data(), at, hipMemcpyHostToDevice)))
        rc = CRN2_EHIP;
    if (rc != CRN2_OK) {
        plan_free(p);
        return rc;
    }
    const uint8_t* d = static_cast<const uint8_t*>(p->tables);
    p->t.p1 = reinterpret_cast<const Pass1Level*>(d + o_p1);
    p->t.p2 = reinterpret_cast<const Pass2Level*>(d + o_p2);
    p->t.sofs = reinterpret_cast<const uint32_t*>(d + o_sofs);
    p->t.runs = reinterpret_cast<const Run*>(d + o_runs);
    p->t.wt = reinterpret_cast<const float*>(d + o_wt);
    p->t.s2l = reinterpret_cast<const float*>(d + o_s2l);
    p->t.l2s = d + o_l2s;
    uint8_t* f = static_cast<uint8_t*>(p->scratch);
    if (need_lx) { p->lx = f; f += plane; }
    if (need_ly) { p->ly = f; f += plane; }
    p->inter = reinterpret_cast<float*>(f);
    *out = p;
    return CRN2_OK;
}

int crn2_mip_plan_info(const crn2_mip_plan* p, uint32_t* levels, uint64_t* chain_bytes, uint64_t* level_offsets,
                       uint8_t* y_first)
{
    if (!plan_ok(p)) return CRN2_EINVAL;
    if (levels) *levels = p->levels;
    if (chain_bytes) *chain_bytes = p->chain_bytes;
    if (level_offsets) std::memcpy(level_offsets, p->level_ofs.data(), p->levels * sizeof(uint64_t));
    if (y_first) std::memcpy(y_first, p->y_first.data(), p->levels);
    return CRN2_OK;
}

int crn2_mip_plan_run(crn2_mip_plan* p, const void* rgba, void* out, uint64_t out_size, void* stream)
{
    if (!plan_ok(p) || !rgba || !out || out_size < p->chain_bytes) return CRN2_EINVAL;
    if ((reinterpret_cast<uintptr_t>(rgba) | reinterpret_cast<uintptr_t>(out)) & 3) return CRN2_EINVAL;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const uint64_t l0 = (uint64_t)p->width * p->height * 4;
    if (out != rgba && !hip_ok(hipMemcpyAsync(out, rgba, l0, hipMemcpyDeviceToDevice, st))) return CRN2_EHIP;
    if (p->levels == 1) return CRN2_OK;
    const int srgb = (p->params.flags & CRN2_MIP_SRGB) != 0;
    if (p->prof) (void)hipEventRecord(p->ev[0], st);
    hipLaunchKernelGGL(mip_planes_kernel, dim3((p->width + 31) / 32, (p->height + 31) / 32), dim3(256), 0, st,
                       static_cast<const uint32_t*>(rgba), p->width, p->height, p->lx, p->ly, p->ncomp);
    if (p->prof) (void)hipEventRecord(p->ev[1], st);
    hipLaunchKernelGGL(mip_pass1_kernel, dim3((p->units1 + 3) / 4), dim3(256), 0, st, p->t, p->lx, p->ly, p->inter,
                       p->levels - 1, p->units1, p->ncomp, srgb);
    if (p->prof) (void)hipEventRecord(p->ev[2], st);
    hipLaunchKernelGGL(mip_pass2_kernel, dim3(p->blocks2), dim3(256), 0, st, p->t, p->inter,
                       static_cast<uint32_t*>(out), p->levels - 1, p->ncomp, srgb,
                       (p->params.flags & CRN2_MIP_RENORMALIZE) ? 1 : 0);
    if (p->prof) {
        (void)hipEventRecord(p->ev[3], st);
        if (hipEventSynchronize(p->ev[3]) == hipSuccess) {
            for (int s = 0; s < 3; s++) {
                float ms = 0;
                if (hipEventElapsedTime(&ms, p->ev[s], p->ev[s + 1]) == hipSuccess) p->ms[s] += ms;
            }
            p->runs++;
        }
    }
    return hip_ok(hipGetLastError()) ? CRN2_OK : CRN2_EHIP;
}

int crn2_mip_plan_destroy(crn2_mip_plan* p)
{
    if (!plan_ok(p)) return CRN2_EINVAL;
    (void)hipDeviceSynchronize();
    plan_free(p);
    return CRN2_OK;
}

/* Stage timing of a plan's runs: on != 0 records HIP events around the three
 * launches of every run (and waits for them); ms[0..2] = planes, pass 1,
 * pass 2 totals since it was switched on. */
int crn2_mip_plan_profile(crn2_mip_plan* p, int on, double* ms, uint64_t* runs)
{
    if (!plan_ok(p)) return CRN2_EINVAL;
    if (on && !p->prof) {
        for (hipEvent_t& e : p->ev)
            if (!e && hipEventCreate(&e) != hipSuccess) return CRN2_EHIP;
        p->ms[0] = p->ms[1] = p->ms[2] = 0;
        p->runs = 0;
    }
    p->prof = on != 0;
    if (ms) std::memcpy(ms, p->ms, sizeof(p->ms));
    if (runs) *runs = p->runs;
    return CRN2_OK;
}

int crn2_generate_mipmaps_device(const void* rgba, uint32_t width, uint32_t height, const crn2_mip_params* params,
                                 void* out, uint64_t out_size, void* stream)
{
    crn2_mip_plan* p = nullptr;
    int rc = crn2_mip_plan_create(width, height, params, &p);
    if (rc != CRN2_OK) return rc;
    rc = crn2_mip_plan_run(p, rgba, out, out_size, stream);
    if (hipStreamSynchronize(static_cast<hipStream_t>(stream)) != hipSuccess && rc == CRN2_OK) rc = CRN2_EHIP;
    plan_free(p);
    return rc;
}

int crn2_generate_mipmaps(const void* rgba, uint32_t width, uint32_t height, const crn2_mip_params* params,
                          void* out, uint64_t out_size)
{
    if (!rgba || !out) return CRN2_EINVAL;
    crn2_mip_plan* p = nullptr;
    int rc = crn2_mip_plan_create(width, height, params, &p);
    if (rc != CRN2_OK) return rc;
    if (out_size < p->chain_bytes) {
        plan_free(p);
        return CRN2_EINVAL;
    }
    const uint64_t l0 = (uint64_t)width * height * 4;
    void *d_in = nullptr, *d_out = nullptr;
    if (!hip_ok(hipMalloc(&d_in, l0)) || !hip_ok(hipMalloc(&d_out, p->chain_bytes))) rc = CRN2_ENOMEM;
    if (rc == CRN2_OK && !hip_ok(hipMemcpy(d_in, rgba, l0, hipMemcpyHostToDevice))) rc = CRN2_EHIP;
    if (rc == CRN2_OK) rc = crn2_mip_plan_run(p, d_in, d_out, p->chain_bytes, nullptr);
    if (rc == CRN2_OK && !hip_ok(hipMemcpy(out, d_out, p->chain_bytes, hipMemcpyDeviceToHost))) rc = CRN2_EHIP;
    if (d_in) (void)hipFree(d_in);
    if (d_out) (void)hipFree(d_out);
    plan_free(p);
    return rc;
}

}  // extern "C"
