// dxt5a_gpu.hip -- DXT5 alpha-block endpoint search (crnlib's
// dxt5_endpoint_optimizer, crnlib/crn_dxt5a.cpp:40-262) on gfx950.
//
// One lane searches one block (16 values of one channel).  The search is pure
// integer work and is reproduced step for step: unique values in first-
// appearance order, all unique pairs, then (quality >= better, error > 0) the
// +-16 / +-8 probe box whose centre follows every improvement.
//
// The reference skips box pairs already visited in either order
// (crn_dxt5a.cpp:131-137).  Skipping them cannot change the result: the
// 8-value and 6-value palettes of (l, h) and (h, l) hold the same values, so
// both orders have the same per-value minima and the same total, and a pair
// that was evaluated once has an error no smaller than the current best,
// which only decreases; re-evaluating it never replaces the best (strict <).
// The lane therefore evaluates every box pair and keeps no visited set.
//
// Block gather for whole images follows crnlib/crn_dxt_image.cpp:330-344
// (4x4 texels with edge clamp) and :1491-1523 (component index, quality and
// both-block-types flag); the block is written as DXT5 alpha bytes
// {first, second, 48 bits of 3-bit selectors, texel 0 in the low bits}.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "autotune.h"
#include "crn2_internal.h"
#include "encode_set.h"

namespace {

__device__ __constant__ uint8_t k_six_invert[8] = {1, 0, 5, 4, 3, 2, 6, 7};
__device__ __constant__ uint8_t k_eight_invert[8] = {1, 0, 7, 6, 5, 4, 3, 2};

struct E2Result {
    uint64_t err;
    uint32_t first, second, type, reordered;
    uint64_t sel48;   // 3-bit selectors of texels 0..15
};

struct E2 {
    uint32_t nu;
    uint32_t uval[16];
    uint32_t uw[16];
    uint32_t map16[16];   // texel -> unique index
    uint64_t best_err;
    uint32_t best_first, best_second, best_type;
    uint64_t best_sel;    // 3 bits per unique value
    int use_both;

    __device__ inline void evaluate(uint32_t lo, uint32_t hi)
    {
        for (uint32_t type = 0; type < (use_both ? 2u : 1u); type++) {
            uint32_t pal[8];
            pal[0] = lo;
            pal[1] = hi;
            if (!type) {
                pal[2] = (lo * 6 + hi) / 7;
                pal[3] = (lo * 5 + hi * 2) / 7;
                pal[4] = (lo * 4 + hi * 3) / 7;
                pal[5] = (lo * 3 + hi * 4) / 7;
                pal[6] = (lo * 2 + hi * 5) / 7;
                pal[7] = (lo + hi * 6) / 7;
            } else {
                pal[2] = (lo * 4 + hi) / 5;
                pal[3] = (lo * 3 + hi * 2) / 5;
                pal[4] = (lo * 2 + hi * 3) / 5;
                pal[5] = (lo + hi * 4) / 5;
                pal[6] = 0;
                pal[7] = 255;
            }
            uint64_t err = 0, sel = 0;
            bool over = false;
            for (uint32_t i = 0; i < nu; i++) {
                const int val = (int)uval[i];
                const int w = (int)uw[i];
                uint32_t best = 0xFFFFFFFFu, bsel = 0;
                for (uint32_t j = 0; j < 8; j++) {
                    int e = val - (int)pal[j];
                    e = e * e * w;
                    if ((uint32_t)e < best) {
                        best = (uint32_t)e;
                        bsel = j;
                        if (!best) break;
                    }
                }
                sel |= (uint64_t)bsel << (3 * i);
                err += best;
                if (err > best_err) {
                    over = true;
                    break;
                }
            }
            if (!over && err < best_err) {
                best_err = err;
                best_first = lo;
                best_second = hi;
                best_type = type;
                best_sel = sel;
                if (!err) break;
            }
        }
    }
};

__device__ inline E2Result dxt5a_search(const uint8_t* v, int quality, int use_both)
{
    E2 s;
    s.nu = 0;
    for (int i = 0; i < 16; i++) {
        uint32_t k = 16;
        for (uint32_t u = 0; u < s.nu; u++)
            if (s.uval[u] == v[i]) { k = u; break; }
        if (k == 16) {
            k = s.nu++;
            s.uval[k] = v[i];
            s.uw[k] = 0;
        }
        s.uw[k]++;
        s.map16[i] = k;
    }
    E2Result r;
    if (s.nu == 1) {
        r.err = 0;
        r.first = r.second = s.uval[0];
        r.type = 0;
        r.reordered = 0;
        r.sel48 = 0;
        return r;
    }
    s.use_both = use_both;
    s.best_err = ~0ull;
    s.best_first = s.best_second = s.best_type = 0;
    s.best_sel = 0;
    for (uint32_t i = 0; i + 1 < s.nu; i++)
        for (uint32_t j = i + 1; j < s.nu; j++) s.evaluate(s.uval[i], s.uval[j]);
    if (quality >= 3 && s.best_err) {
        const int probe = (quality == 4) ? 16 : 8;
        for (int dl = -probe; dl <= probe; dl++) {
            const int l = (int)s.best_first + dl;
            if (l < 0) continue;
            if (l > 255) break;
            for (int dh = -probe; dh <= probe; dh++) {
                const int h = (int)s.best_second + dh;
                if (h < 0) continue;
                if (h > 255) break;
                s.evaluate((uint32_t)l, (uint32_t)h);
            }
        }
    }
    uint32_t f = s.best_first, sc = s.best_second;
    r.reordered = 0;
    uint64_t usel = s.best_sel;
    if (f == sc) {
        usel = 0;
    } else if (s.best_type) {
        if (f > sc) {
            const uint32_t t = f; f = sc; sc = t;
            r.reordered = 1;
            uint64_t o = 0;
            for (uint32_t i = 0; i < s.nu; i++) o |= (uint64_t)k_six_invert[(usel >> (3 * i)) & 7] << (3 * i);
            usel = o;
        }
    } else if (f <= sc) {
        const uint32_t t = f; f = sc; sc = t;
        r.reordered = 1;
        uint64_t o = 0;
        for (uint32_t i = 0; i < s.nu; i++) o |= (uint64_t)k_eight_invert[(usel >> (3 * i)) & 7] << (3 * i);
        usel = o;
    }
    uint64_t sel48 = 0;
    for (int i = 0; i < 16; i++) sel48 |= ((usel >> (3 * s.map16[i])) & 7) << (3 * i);
    r.err = s.best_err;
    r.first = f;
    r.second = sc;
    r.type = s.best_type;
    r.sel48 = sel48;
    return r;
}

// values: 16 bytes per block; outputs per block: sel48 (u64), err (u64),
// meta = first | second << 8 | type << 16 | reordered << 24
__global__ void __launch_bounds__(256) dxt5a_blocks_kernel(const uint8_t* __restrict__ values, uint32_t nblocks,
                                                           int quality, int use_both, uint64_t* __restrict__ sel,
                                                           uint64_t* __restrict__ err, uint32_t* __restrict__ meta)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblocks) return;
    const uint4 raw = reinterpret_cast<const uint4*>(values)[b];
    uint8_t v[16];
    const uint32_t w[4] = {raw.x, raw.y, raw.z, raw.w};
    for (int i = 0; i < 16; i++) v[i] = (uint8_t)(w[i >> 2] >> (8 * (i & 3)));
    const E2Result r = dxt5a_search(v, quality, use_both);
    sel[b] = r.sel48;
    err[b] = r.err;
    meta[b] = r.first | (r.second << 8) | (r.type << 16) | (r.reordered << 24);
}

// Whole image: RGBA8 texels, one channel, DXT5 alpha blocks out (8 bytes each).
__global__ void __launch_bounds__(256) dxt5a_image_kernel(const uint8_t* __restrict__ rgba, uint32_t width,
                                                          uint32_t height, uint32_t comp, uint32_t blocks_x,
                                                          uint32_t blocks_y, int quality, int use_both,
                                                          uint64_t* __restrict__ out, uint32_t out_stride,
                                                          uint64_t* __restrict__ err)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= blocks_x * blocks_y) return;
    const uint32_t bx = b % blocks_x, by = b / blocks_x;
    uint8_t v[16];
    for (int y = 0; y < 4; y++)
        for (int x = 0; x < 4; x++) {
            const uint32_t sx = min(bx * 4 + x, width - 1), sy = min(by * 4 + y, height - 1);
            v[y * 4 + x] = rgba[4 * ((size_t)sy * width + sx) + comp];
        }
    const E2Result r = dxt5a_search(v, quality, use_both);
    out[(size_t)b * out_stride] = (uint64_t)r.first | ((uint64_t)r.second << 8) | (r.sel48 << 16);
    if (err) err[b] = r.err;
}

// Texture-set encode, alpha elements (DXT5 alpha, DXN channels, DXT5A) of
// many level images in one flat order, grid-stride as dxt1_set_kernel.
__global__ void __launch_bounds__(256) dxt5a_set_kernel(const crn2::SetJob* __restrict__ jobs, uint32_t njobs,
                                                        uint64_t begin, uint64_t end, int quality)
{
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = begin + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < end; i += step) {
        const crn2::SetJob j = jobs[crn2::set_job_of(jobs, njobs, i)];
        const uint32_t b = (uint32_t)(i - j.first);
        const uint32_t bx = b % j.blocks_x, by = b / j.blocks_x;
        uint8_t v[16];
        for (int y = 0; y < 4; y++)
            for (int x = 0; x < 4; x++) {
                const uint32_t sx = min(bx * 4 + x, j.width - 1), sy = min(by * 4 + y, j.height - 1);
                v[y * 4 + x] = (uint8_t)(j.rgba[(size_t)sy * j.width + sx] >> (8 * j.comp));
            }
        const E2Result r = dxt5a_search(v, quality, (int)j.flags);
        j.out[(size_t)b * j.stride] = (uint64_t)r.first | ((uint64_t)r.second << 8) | (r.sel48 << 16);
    }
}

}  // namespace

extern "C" int crn2_dxt5a_blocks_device(const void* values, uint32_t nblocks, int quality, int use_both,
                                        void* selectors48, void* errors, void* meta, void* stream)
{
    if (!values || !selectors48 || !errors || !meta || quality < 0 || quality > 4) return CRN2_EINVAL;
    if (reinterpret_cast<uintptr_t>(values) & 15) return CRN2_EINVAL;
    if (!nblocks) return CRN2_OK;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // launch shapes (threads per workgroup, one block per lane; identical results): 256, 64, 128
    auto launch = [&](int shape) -> hipError_t {
        const uint32_t tb = shape == 1 ? 64 : shape == 2 ? 128 : 256, nb = (nblocks + tb - 1) / tb;
        hipLaunchKernelGGL(dxt5a_blocks_kernel, dim3(nb), dim3(tb), 0, s, static_cast<const uint8_t*>(values),
                           nblocks, quality, use_both, static_cast<uint64_t*>(selectors48),
                           static_cast<uint64_t*>(errors), static_cast<uint32_t*>(meta));
        return hipGetLastError();
    };
    const int shape = crn2::tuner().pick(crn2::TUNE_DXT5A, nblocks, 3, 0, s, launch);
    return launch(shape) == hipSuccess ? CRN2_OK : CRN2_EHIP;
}

extern "C" int crn2_encode_dxt5a_image_device(const void* rgba, uint32_t width, uint32_t height, uint32_t comp,
                                              int quality, int use_both, void* blocks, uint32_t block_stride_u64,
                                              void* errors, void* stream)
{
    if (!rgba || !blocks || !width || !height || comp > 3 || quality < 0 || quality > 4 || !block_stride_u64)
        return CRN2_EINVAL;
    const uint32_t bx = (width + 3) / 4, by = (height + 3) / 4;
    const uint32_t n = bx * by, tb = 256, nb = (n + tb - 1) / tb;
    hipLaunchKernelGGL(dxt5a_image_kernel, dim3(nb), dim3(tb), 0, static_cast<hipStream_t>(stream),
                       static_cast<const uint8_t*>(rgba), width, height, comp, bx, by, quality, use_both,
                       static_cast<uint64_t*>(blocks), block_stride_u64, static_cast<uint64_t*>(errors));
    return hipGetLastError() == hipSuccess ? CRN2_OK : CRN2_EHIP;
}

extern "C" int crn2_dxt5a_set_launch(const crn2::SetJob* jobs, uint32_t njobs, uint64_t begin, uint64_t end,
                                     int quality, uint32_t grid, uint32_t tb, hipStream_t stream)
{
    if (!jobs || !njobs || !grid || (tb != 64 && tb != 128 && tb != 256) || quality < 0 || quality > 4)
        return CRN2_EINVAL;
    if (begin >= end) return CRN2_OK;
    hipLaunchKernelGGL(dxt5a_set_kernel, dim3(grid), dim3(tb), 0, stream, jobs, njobs, begin, end, quality);
    return hipGetLastError() == hipSuccess ? CRN2_OK : CRN2_EHIP;
}
