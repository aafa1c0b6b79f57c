// hc_gpu.hip -- dxt_hc colour selector assignment on gfx950 INT8 matrix cores.
//
// Reference: dxt_hc::create_color_selector_codebook_task
// (crnlib/crn_dxt_hc.cpp:1306-1360).  For block b and codebook selector s the
// reference error is
//     err(b, s) = sum_p E2_b[p][sel_s(p)],  E2_b[p][c] = color_distance(pixel p, cluster colour c)
// (color_distance: crn_color.h:728-745, weights 8/25/1 or 1/1/1, alpha
// ignored).  Written over the 64 (pixel, colour) pairs k = 4p + c this is a
// product of an n_blocks x 64 matrix E2 with a 64 x n_selectors one-hot matrix
// H (H[4p + c][s] = 1 iff selector s gives pixel p code c, 16 ones a column):
//     ERR = E2 * H.
// E2 entries are below 2^22 (2^18 uniform), so E2 is split into 7-bit limbs
// (4 perceptual, 3 uniform), each limb times H runs on
// v_mfma_i32_32x32x32_i8 with exact INT32 accumulation (a limb sum is at most
// 16 * 127), and the limbs recombine as sum_l acc_l << 7l (below 2^26): the
// error matrix is exact.  The selector with the smallest error, lowest index
// on ties (the reference's strict '<' scan), is kept per block; the error
// matrix is never stored.
//
// Kernel 1 (crn2_hc_sel_gemm_kernel): a wave owns 32 blocks and keeps their
// limb fragments in registers for the whole codebook; H fragments are built in
// registers from the selector words (staged in LDS) and never touch memory.
// Kernel 2 (crn2_hc_sel_scatter_kernel): adds each block's E2 into the 16 x 4
// error totals of its chosen selector (integer atomics: order-independent,
// so exact) and marks the selector used.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "autotune.h"
#include "crn2_internal.h"

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef int v16i __attribute__((ext_vector_type(16)));

constexpr uint32_t kRowsPerWave = 32;
constexpr uint32_t kSelChunk = 8192;   // selectors staged in LDS at a time (32 KiB)

__device__ inline uint32_t color_distance(bool perceptual, uint32_t a, uint32_t b)
{
    const int dr = (int)(a & 255u) - (int)(b & 255u);
    const int dg = (int)((a >> 8) & 255u) - (int)((b >> 8) & 255u);
    const int db = (int)((a >> 16) & 255u) - (int)((b >> 16) & 255u);
    if (perceptual) return (uint32_t)(8 * dr * dr + 25 * dg * dg + db * db);
    return (uint32_t)(dr * dr + dg * dg + db * db);
}

struct SelArgs {
    const uint32_t* blocks;          // n_blocks x 16 RGBA8 words
    const uint32_t* endpoint_index;  // n_blocks (or 2 n_blocks with subblocks)
    const uint32_t* cluster_colors;  // n_clusters x 4 RGBA8 words
    const uint32_t* selectors;       // n_selectors
    uint32_t* best_index;
    uint32_t* best_error;            // may be null
    uint32_t* details;               // n_selectors x 64, may be null
    uint8_t* used;                   // n_selectors, may be null
    uint32_t n_blocks, n_selectors;
    int subblocks, perceptual;
};

// E2 of pixel p of block `row` against its cluster's 4 colours.
__device__ inline void e2_pixel(const SelArgs& a, uint32_t row, uint32_t p, uint32_t (&e)[4])
{
    const uint32_t px = a.blocks[(size_t)row * 16 + p];
    const uint32_t cl = a.subblocks ? a.endpoint_index[2 * (size_t)row + (p >> 3)] : a.endpoint_index[row];
    const uint32_t* cc = a.cluster_colors + (size_t)cl * 4;
    for (uint32_t c = 0; c < 4; c++) e[c] = color_distance(a.perceptual != 0, px, cc[c]);
}

template <int NL, int WAVES>
__global__ void __launch_bounds__(64 * WAVES) crn2_hc_sel_gemm_kernel(SelArgs a)
{
    constexpr uint32_t kRowsPerWG = kRowsPerWave * WAVES;
    __shared__ uint32_t s_sel[kSelChunk];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t g = lane >> 5;                 // K half of the MFMA operand layout
    const uint32_t col = lane & 31u;              // column (selector) of B / C owned by the lane
    const uint32_t row0 = blockIdx.x * kRowsPerWG + wave * kRowsPerWave;
    const uint32_t arow = row0 + col;             // A row owned by the lane

    // A fragments: fragment t covers k = 32 t + 16 g + j, j = 0..15, i.e. pixels
    // p = 8t + 4g + j/4, colours c = j%4.  The same (t, g, j) -> k map is used
    // for B, so the product does not depend on how the instruction orders K.
    v4i A[NL][2];
    for (uint32_t t = 0; t < 2; t++) {
        uint32_t w[NL][4];
        for (uint32_t q = 0; q < 4; q++) {
            uint32_t e[4] = {0, 0, 0, 0};
            if (arow < a.n_blocks) e2_pixel(a, arow, 8 * t + 4 * g + q, e);
            for (int l = 0; l < NL; l++) {
                uint32_t v = 0;
                for (uint32_t c = 0; c < 4; c++) v |= ((e[c] >> (7 * l)) & 127u) << (8 * c);
                w[l][q] = v;
            }
        }
        for (int l = 0; l < NL; l++) A[l][t] = v4i{(int)w[l][0], (int)w[l][1], (int)w[l][2], (int)w[l][3]};
    }

    uint32_t best[16], bt[16];
    for (int i = 0; i < 16; i++) {
        best[i] = 0xFFFFFFFFu;
        bt[i] = 0;
    }
    for (uint32_t c0 = 0; c0 < a.n_selectors; c0 += kSelChunk) {
        const uint32_t cn = min(kSelChunk, a.n_selectors - c0);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < cn; i += blockDim.x) s_sel[i] = a.selectors[c0 + i];
        __syncthreads();
        for (uint32_t t0 = 0; t0 < cn; t0 += 32) {
            const bool valid = t0 + col < cn;
            const uint32_t sel = valid ? s_sel[t0 + col] : 0u;
            v4i B[2];
            for (uint32_t t = 0; t < 2; t++) {
                uint32_t w[4];
                for (uint32_t q = 0; q < 4; q++) {
                    const uint32_t p = 8 * t + 4 * g + q;
                    w[q] = 1u << (8 * ((sel >> (2 * p)) & 3u));
                }
                B[t] = v4i{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
            }
            v16i acc[NL];
            for (int l = 0; l < NL; l++) {
                acc[l] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[l][0], B[0], v16i{0}, 0, 0, 0);
                acc[l] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[l][1], B[1], acc[l], 0, 0, 0);
            }
            const uint32_t tile = (c0 + t0) >> 5;
            for (int i = 0; i < 16; i++) {
                uint32_t e = (uint32_t)acc[0][i];
                for (int l = 1; l < NL; l++) e += (uint32_t)acc[l][i] << (7 * l);
                if (valid && e < best[i]) {
                    best[i] = e;
                    bt[i] = tile;
                }
            }
        }
    }
    // C layout of the 32x32 product: register i of lane l holds column l%32,
    // row 8(i/4) + 4(l/32) + i%4.  Reduce each row over its 32 columns:
    // smallest error, then smallest selector index.
    for (int i = 0; i < 16; i++) {
        uint32_t e = best[i];
        uint32_t idx = bt[i] * 32 + col;
        for (int off = 1; off < 32; off <<= 1) {
            const uint32_t oe = (uint32_t)__shfl_xor((int)e, off);
            const uint32_t oi = (uint32_t)__shfl_xor((int)idx, off);
            if (oe < e || (oe == e && oi < idx)) {
                e = oe;
                idx = oi;
            }
        }
        const uint32_t r = row0 + 8 * (i >> 2) + 4 * g + (i & 3);
        if (col == 0 && r < a.n_blocks) {
            a.best_index[r] = idx;
            if (a.best_error) a.best_error[r] = e;
        }
    }
}

__global__ void __launch_bounds__(256) crn2_hc_sel_scatter_kernel(SelArgs a)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t b = t >> 4;
    const uint32_t p = (uint32_t)(t & 15u);
    if (b >= a.n_blocks) return;
    uint32_t e[4];
    e2_pixel(a, (uint32_t)b, p, e);
    const uint32_t s = a.best_index[b];
    if (a.details)
        for (uint32_t c = 0; c < 4; c++) atomicAdd(&a.details[(size_t)s * 64 + 4 * p + c], e[c]);
    if (a.used && p == 0) a.used[s] = 1;
}

// ----------------------------------------------------------------------------
// Alpha selector assignment (dxt_hc::create_alpha_selector_codebook_task,
// crnlib/crn_dxt_hc.cpp:1516-1586): E3[p][v] = (a_p - value_v)^2 over the 8
// values of the block's cluster, error(b, s) = sum_p E3[p][sel_s(p)] with 3-bit
// selectors.  As a product: E3 (n x 128, k = 8p + v) times a one-hot H
// (128 x m).  E3 < 2^16 splits into three 7-bit limbs; a limb sum is at most
// 16 * 127, the recombined error below 2^21: exact.

struct AlphaSelArgs {
    const uint8_t* block_vals;       // n_blocks x 16 values of the channel
    const uint32_t* endpoint_index;  // n_blocks
    const uint8_t* values;           // n_clusters x 8 (the assignment's palette)
    const uint8_t* scatter_values;   // n_clusters x 8 (the totals' palette)
    const uint64_t* selectors;       // n_selectors, texel p at bits 3p..3p+2
    uint32_t* best_index;
    uint32_t* best_error;            // may be null
    uint32_t* details;               // n_selectors x 128, may be null
    uint8_t* used;                   // may be null
    uint32_t n_blocks, n_selectors;
};

constexpr uint32_t kAlphaSelChunk = 4096;   // u64 selectors staged in LDS at a time (32 KiB)

template <int WAVES>
__global__ void __launch_bounds__(64 * WAVES) crn2_hc_alpha_sel_gemm_kernel(AlphaSelArgs a)
{
    constexpr int NL = 3;
    constexpr uint32_t kRowsPerWG = kRowsPerWave * WAVES;
    __shared__ uint64_t s_sel[kAlphaSelChunk];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t g = lane >> 5, col = lane & 31u;
    const uint32_t row0 = blockIdx.x * kRowsPerWG + wave * kRowsPerWave;
    const uint32_t arow = row0 + col;

    // Fragment t covers k = 32 t + 16 g + j (j = 0..15): texel p = 4t + 2g + j/8,
    // value v = j % 8; operand word q holds j = 4q..4q+3, i.e. texel
    // 4t + 2g + q/2 and values 4(q%2)..4(q%2)+3.
    v4i A[NL][4];
    {
        uint32_t vals[8];
        const uint8_t* bv = nullptr;
        if (arow < a.n_blocks) {
            const uint8_t* cv = a.values + (size_t)a.endpoint_index[arow] * 8;
            for (int v = 0; v < 8; v++) vals[v] = cv[v];
            bv = a.block_vals + (size_t)arow * 16;
        }
        for (uint32_t t = 0; t < 4; t++) {
            uint32_t w[NL][4];
            for (uint32_t q = 0; q < 4; q++) {
                const uint32_t p = 4 * t + 2 * g + (q >> 1);
                uint32_t e[4] = {0, 0, 0, 0};
                if (bv) {
                    const int x = bv[p];
                    for (uint32_t c = 0; c < 4; c++) {
                        const int d = x - (int)vals[4 * (q & 1) + c];
                        e[c] = (uint32_t)(d * d);
                    }
                }
                for (int l = 0; l < NL; l++) {
                    uint32_t v = 0;
                    for (uint32_t c = 0; c < 4; c++) v |= ((e[c] >> (7 * l)) & 127u) << (8 * c);
                    w[l][q] = v;
                }
            }
            for (int l = 0; l < NL; l++) A[l][t] = v4i{(int)w[l][0], (int)w[l][1], (int)w[l][2], (int)w[l][3]};
        }
    }

    uint32_t best[16], bt[16];
    for (int i = 0; i < 16; i++) {
        best[i] = 0xFFFFFFFFu;
        bt[i] = 0;
    }
    for (uint32_t c0 = 0; c0 < a.n_selectors; c0 += kAlphaSelChunk) {
        const uint32_t cn = min(kAlphaSelChunk, a.n_selectors - c0);
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < cn; i += blockDim.x) s_sel[i] = a.selectors[c0 + i];
        __syncthreads();
        for (uint32_t t0 = 0; t0 < cn; t0 += 32) {
            const bool valid = t0 + col < cn;
            const uint64_t sel = valid ? s_sel[t0 + col] : 0ull;
            v16i acc[NL];
            for (int l = 0; l < NL; l++) acc[l] = v16i{0};
            for (uint32_t t = 0; t < 4; t++) {
                uint32_t w[4];
                for (uint32_t q = 0; q < 4; q++) {
                    const uint32_t p = 4 * t + 2 * g + (q >> 1);
                    const uint32_t s3 = (uint32_t)(sel >> (3 * p)) & 7u;
                    w[q] = (s3 >> 2) == (q & 1) ? 1u << (8 * (s3 & 3u)) : 0u;
                }
                const v4i B = v4i{(int)w[0], (int)w[1], (int)w[2], (int)w[3]};
                for (int l = 0; l < NL; l++) acc[l] = __builtin_amdgcn_mfma_i32_32x32x32_i8(A[l][t], B, acc[l], 0, 0, 0);
            }
            const uint32_t tile = (c0 + t0) >> 5;
            for (int i = 0; i < 16; i++) {
                const uint32_t e = (uint32_t)acc[0][i] + ((uint32_t)acc[1][i] << 7) + ((uint32_t)acc[2][i] << 14);
                if (valid && e < best[i]) {
                    best[i] = e;
                    bt[i] = tile;
                }
            }
        }
    }
    for (int i = 0; i < 16; i++) {
        uint32_t e = best[i];
        uint32_t idx = bt[i] * 32 + col;
        for (int off = 1; off < 32; off <<= 1) {
            const uint32_t oe = (uint32_t)__shfl_xor((int)e, off);
            const uint32_t oi = (uint32_t)__shfl_xor((int)idx, off);
            if (oe < e || (oe == e && oi < idx)) {
                e = oe;
                idx = oi;
            }
        }
        const uint32_t r = row0 + 8 * (i >> 2) + 4 * g + (i & 3);
        if (col == 0 && r < a.n_blocks) {
            a.best_index[r] = idx;
            if (a.best_error) a.best_error[r] = e;
        }
    }
}

// Adds each block's E3 against its cluster's scatter palette into the totals
// of its chosen selector (integer atomics: exact, order-independent).
__global__ void __launch_bounds__(256) crn2_hc_alpha_sel_scatter_kernel(AlphaSelArgs a)
{
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t b = t >> 4;
    const uint32_t p = (uint32_t)(t & 15u);
    if (b >= a.n_blocks) return;
    const int x = a.block_vals[b * 16 + p];
    const uint8_t* cv = a.scatter_values + (size_t)a.endpoint_index[b] * 8;
    const uint32_t s = a.best_index[b];
    if (a.details)
        for (uint32_t v = 0; v < 8; v++) {
            const int d = x - (int)cv[v];
            atomicAdd(&a.details[(size_t)s * 128 + 8 * p + v], (uint32_t)(d * d));
        }
    if (a.used && p == 0) a.used[s] = 1;
}

// Per codebook entry and texel, the value with the smallest accumulated error,
// chosen by the reference's pairwise tournament (its tie order is not "lowest
// value first": crn_dxt_hc.cpp:1493-1501 for 2-bit, 1706-1717 for 3-bit).
__global__ void __launch_bounds__(256) crn2_hc_sel_refine_kernel(const uint32_t* __restrict__ details, uint32_t n,
                                                                 int bits, void* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (bits == 2) {
        const uint32_t* e = details + (size_t)i * 64;
        uint32_t sel = 0;
        for (uint32_t p = 0; p < 16; p++, e += 4) {
            const uint32_t s03 = e[3] < e[0] ? 3 : 0;
            const uint32_t s12 = e[2] < e[1] ? 2 : 1;
            sel |= (e[s12] < e[s03] ? s12 : s03) << (2 * p);
        }
        static_cast<uint32_t*>(out)[i] = sel;
    } else {
        const uint32_t* e = details + (size_t)i * 128;
        uint64_t sel = 0;
        for (uint32_t p = 0; p < 16; p++, e += 8) {
            const uint32_t s07 = e[7] < e[0] ? 7 : 0;
            const uint32_t s12 = e[2] < e[1] ? 2 : 1;
            const uint32_t s34 = e[4] < e[3] ? 4 : 3;
            const uint32_t s56 = e[6] < e[5] ? 6 : 5;
            const uint32_t s02 = e[s12] < e[s07] ? s12 : s07;
            const uint32_t s36 = e[s56] < e[s34] ? s56 : s34;
            sel |= (uint64_t)(e[s36] < e[s02] ? s36 : s02) << (3 * p);
        }
        static_cast<uint64_t*>(out)[i] = sel;
    }
}

}  // namespace

extern "C" int crn2_hc_color_selectors_device(const void* blocks, const uint32_t* endpoint_index,
                                              const void* cluster_colors, uint32_t n_blocks, int subblocks,
                                              int perceptual, const uint32_t* selectors, uint32_t n_selectors,
                                              uint32_t* best_index, uint32_t* best_error, uint32_t* details,
                                              uint8_t* used, void* stream)
{
    if (!n_blocks) return CRN2_OK;
    if (!blocks || !endpoint_index || !cluster_colors || !selectors || !n_selectors || !best_index)
        return CRN2_EINVAL;
    SelArgs a;
    a.blocks = static_cast<const uint32_t*>(blocks);
    a.endpoint_index = endpoint_index;
    a.cluster_colors = static_cast<const uint32_t*>(cluster_colors);
    a.selectors = selectors;
    a.best_index = best_index;
    a.best_error = best_error;
    a.details = details;
    a.used = used;
    a.n_blocks = n_blocks;
    a.n_selectors = n_selectors;
    a.subblocks = subblocks ? 1 : 0;
    a.perceptual = perceptual ? 1 : 0;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // launch shapes (waves per workgroup; identical results): 4, 2, 8
    auto launch = [&](int shape) -> hipError_t {
        const int waves = shape == 1 ? 2 : shape == 2 ? 8 : 4;
        const uint32_t rows = kRowsPerWave * waves;
        const uint32_t wgs = (n_blocks + rows - 1) / rows;
        if (a.perceptual) {
            if (waves == 2) hipLaunchKernelGGL((crn2_hc_sel_gemm_kernel<4, 2>), dim3(wgs), dim3(128), 0, s, a);
            else if (waves == 8) hipLaunchKernelGGL((crn2_hc_sel_gemm_kernel<4, 8>), dim3(wgs), dim3(512), 0, s, a);
            else hipLaunchKernelGGL((crn2_hc_sel_gemm_kernel<4, 4>), dim3(wgs), dim3(256), 0, s, a);
        } else {
            if (waves == 2) hipLaunchKernelGGL((crn2_hc_sel_gemm_kernel<3, 2>), dim3(wgs), dim3(128), 0, s, a);
            else if (waves == 8) hipLaunchKernelGGL((crn2_hc_sel_gemm_kernel<3, 8>), dim3(wgs), dim3(512), 0, s, a);
            else hipLaunchKernelGGL((crn2_hc_sel_gemm_kernel<3, 4>), dim3(wgs), dim3(256), 0, s, a);
        }
        return hipGetLastError();
    };
    const int shape = crn2::tuner().pick(crn2::TUNE_HC_SEL, (uint64_t)n_blocks * n_selectors, 3, 0, s, launch);
    if (launch(shape) != hipSuccess) return CRN2_EHIP;
    if (details || used) {
        const uint64_t n = (uint64_t)n_blocks * 16;
        hipLaunchKernelGGL(crn2_hc_sel_scatter_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, a);
        if (hipGetLastError() != hipSuccess) return CRN2_EHIP;
    }
    return CRN2_OK;
}

extern "C" int crn2_hc_alpha_selectors_device(const void* block_vals, const uint32_t* endpoint_index,
                                              const void* values, const void* scatter_values, uint32_t n_blocks,
                                              const uint64_t* selectors, uint32_t n_selectors, uint32_t* best_index,
                                              uint32_t* best_error, uint32_t* details, uint8_t* used, void* stream)
{
    if (!n_blocks) return CRN2_OK;
    if (!block_vals || !endpoint_index || !values || !selectors || !n_selectors || !best_index) return CRN2_EINVAL;
    AlphaSelArgs a;
    a.block_vals = static_cast<const uint8_t*>(block_vals);
    a.endpoint_index = endpoint_index;
    a.values = static_cast<const uint8_t*>(values);
    a.scatter_values = static_cast<const uint8_t*>(scatter_values ? scatter_values : values);
    a.selectors = selectors;
    a.best_index = best_index;
    a.best_error = best_error;
    a.details = details;
    a.used = used;
    a.n_blocks = n_blocks;
    a.n_selectors = n_selectors;
    hipStream_t s = static_cast<hipStream_t>(stream);
    const uint32_t wgs = (n_blocks + kRowsPerWave * 4 - 1) / (kRowsPerWave * 4);
    hipLaunchKernelGGL(crn2_hc_alpha_sel_gemm_kernel<4>, dim3(wgs), dim3(256), 0, s, a);
    if (hipGetLastError() != hipSuccess) return CRN2_EHIP;
    if (details || used) {
        const uint64_t n = (uint64_t)n_blocks * 16;
        hipLaunchKernelGGL(crn2_hc_alpha_sel_scatter_kernel, dim3((uint32_t)((n + 255) / 256)), dim3(256), 0, s, a);
        if (hipGetLastError() != hipSuccess) return CRN2_EHIP;
    }
    return CRN2_OK;
}

extern "C" int crn2_hc_selector_refine_device(const uint32_t* details, uint32_t n_selectors, int bits, void* selectors,
                                              void* stream)
{
    if (!n_selectors) return CRN2_OK;
    if (!details || !selectors || (bits != 2 && bits != 3)) return CRN2_EINVAL;
    hipLaunchKernelGGL(crn2_hc_sel_refine_kernel, dim3((n_selectors + 255) / 256), dim3(256), 0,
                       static_cast<hipStream_t>(stream), details, n_selectors, bits, selectors);
    return hipGetLastError() == hipSuccess ? CRN2_OK : CRN2_EHIP;
}

namespace crn2 {
Autotuner& tuner()
{
    static Autotuner t;
    return t;
}
}  // namespace crn2

extern "C" int crn2_autotune_mode(int mode)
{
    if (mode < 0 || mode > 1) return CRN2_EINVAL;
    std::lock_guard<std::mutex> lk(crn2::tuner().mu);
    crn2::tuner().mode = mode;
    if (!mode) {
        crn2::tuner().best.clear();
        crn2::tuner().best_ms.clear();
    }
    return CRN2_OK;
}

extern "C" int crn2_autotune_choice(int kernel, uint64_t work, int* shape, float* ms)
{
    if (kernel < 0 || kernel >= crn2::TUNE_COUNT || !shape) return CRN2_EINVAL;
    std::lock_guard<std::mutex> lk(crn2::tuner().mu);
    const auto key = std::make_pair(kernel, crn2::Autotuner::size_class(work));
    auto it = crn2::tuner().best.find(key);
    if (it == crn2::tuner().best.end()) return CRN2_EINVAL;
    *shape = it->second;
    if (ms) *ms = crn2::tuner().best_ms[key];
    return CRN2_OK;
}
