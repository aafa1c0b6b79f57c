// hc_codebook_gpu.hip -- dxt_hc's selector codebooks on gfx950:
// create_color_selector_codebook (crnlib/crn_dxt_hc.cpp:1379-1504) and
// create_alpha_selector_codebook (:1588-1720).
//
// Steps, all on the device:
//  1. merge: the blocks' (selector << w_bits | weight) words sorted ascending
//     (radix sort), equal selectors merged with their weights summed and held
//     at 2^32 - 1 (the reference's saturating sum, :1440-1447);
//  2. each distinct selector becomes a 16-float vector, texel t ->
//     (code + 0.5) / 2^bits (:1431-1438);
//  3. the tree clusterizer's codebook of those vectors (crn2_tree_codebook_device);
//  4. each codebook vector back to codes, texel t -> (uint)(v * 2^bits)
//     (:1455-1463);
//  5. every block assigned its best codebook entry on INT8 matrix cores, with
//     the per-entry error totals (crn2_hc_color_selectors_device /
//     crn2_hc_alpha_selectors_device);
//  6. every entry replaced by the per-texel tournament argmin of its totals
//     (crn2_hc_selector_refine_device).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdint>

#include "crn2_internal.h"

extern "C" int crn2_tree_codebook_device(const float* vectors, const uint32_t* weights, uint32_t n, uint32_t dim,
                                         uint32_t max_splits, float* codebook, uint32_t* codebook_size,
                                         uint32_t* vector_leaf, void* stream);
extern "C" int crn2_hc_color_selectors_device(const void* blocks, const uint32_t* endpoint_index,
                                              const void* cluster_colors, uint32_t n_blocks, int subblocks,
                                              int perceptual, const uint32_t* selectors, uint32_t n_selectors,
                                              uint32_t* best_index, uint32_t* best_error, uint32_t* details,
                                              uint8_t* used, void* stream);
extern "C" int crn2_hc_alpha_selectors_device(const void* block_vals, const uint32_t* endpoint_index,
                                              const void* values, const void* scatter_values, uint32_t n_blocks,
                                              const uint64_t* selectors, uint32_t n_selectors, uint32_t* best_index,
                                              uint32_t* best_error, uint32_t* details, uint8_t* used, void* stream);
extern "C" int crn2_hc_selector_refine_device(const uint32_t* details, uint32_t n_selectors, int bits, void* selectors,
                                              void* stream);

namespace {

struct SatAdd {
    __host__ __device__ uint32_t operator()(uint32_t a, uint32_t b) const
    {
        return a > 0xFFFFFFFFu - b ? 0xFFFFFFFFu : a + b;
    }
};

// word = selector << wbits | weight  ->  (selector, weight)
__global__ void __launch_bounds__(256) split_words_kernel(const uint64_t* __restrict__ words, uint32_t n, int wbits,
                                                         uint64_t* __restrict__ sel, uint32_t* __restrict__ w)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint64_t x = words[i];
    sel[i] = x >> wbits;
    w[i] = (uint32_t)(x & ((1ull << wbits) - 1));
}

// Distinct selector (texel 0 in the top bits) -> vector, texel t at element t.
__global__ void __launch_bounds__(256) sel_vectors_kernel(const uint64_t* __restrict__ sel, const uint32_t* nsel,
                                                         int bits, float* __restrict__ vec)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *nsel) return;
    uint64_t s = sel[i];
    const uint32_t mask = (1u << bits) - 1u;
    const float scale = bits == 2 ? 0.25f : 0.125f;
    for (int p = 0; p < 16; p++, s >>= bits) vec[(size_t)i * 16 + 15 - p] = ((float)(s & mask) + 0.5f) * scale;
}

// Codebook vector -> codes, texel t at bits t * bits.
__global__ void __launch_bounds__(256) quantize_kernel(const float* __restrict__ cb, uint32_t n, int bits,
                                                      void* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* v = cb + (size_t)i * 16;
    if (bits == 2) {
        uint32_t s = 0;
        for (int j = 0; j < 16; j++) s |= (uint32_t)(v[j] * 4.0f) << (2 * j);
        static_cast<uint32_t*>(out)[i] = s;
    } else {
        uint64_t s = 0;
        for (int j = 0; j < 16; j++) s |= (uint64_t)(v[j] * 8.0f) << (3 * j);
        static_cast<uint64_t*>(out)[i] = s;
    }
}

struct Scratch {
    hipStream_t st;
    void* p[12] = {};
    int n = 0;
    template <class T>
    T* get(size_t count)
    {
        void* q = nullptr;
        if (n >= 12 || hipMallocAsync(&q, count ? count * sizeof(T) : 1, st) != hipSuccess) return nullptr;
        p[n++] = q;
        return static_cast<T*>(q);
    }
    ~Scratch()
    {
        for (int i = 0; i < n; i++) (void)hipFreeAsync(p[i], st);
        (void)hipStreamSynchronize(st);
    }
};

// Steps 1-4: merged selector vectors -> quantized codebook (bits 2: u32, 3: u64)
// in `codebook`, its size in *k.
int build_codebook(Scratch& sc, const uint64_t* words, uint32_t n, int bits, uint32_t max_size, void* codebook,
                   uint32_t* k)
{
    hipStream_t st = sc.st;
    const int wbits = bits == 2 ? 32 : 16;
    uint64_t* sorted = sc.get<uint64_t>(n);
    uint64_t* sel = sc.get<uint64_t>(n);
    uint32_t* w = sc.get<uint32_t>(n);
    uint64_t* usel = sc.get<uint64_t>(n);
    uint32_t* uw = sc.get<uint32_t>(n);
    uint32_t* nrun = sc.get<uint32_t>(1);
    float* vec = sc.get<float>((size_t)n * 16);
    float* cb = sc.get<float>((size_t)max_size * 16);
    if (!sorted || !sel || !w || !usel || !uw || !nrun || !vec || !cb) return CRN2_ENOMEM;
    size_t t1 = 0, t2 = 0;
    if (hipcub::DeviceRadixSort::SortKeys(nullptr, t1, words, sorted, (int)n, 0, 64, st) != hipSuccess ||
        hipcub::DeviceReduce::ReduceByKey(nullptr, t2, sel, usel, w, uw, nrun, SatAdd(), (int)n, st) != hipSuccess)
        return CRN2_EHIP;
    uint8_t* temp = sc.get<uint8_t>(t1 > t2 ? t1 : t2);
    if (!temp) return CRN2_ENOMEM;
    if (hipcub::DeviceRadixSort::SortKeys(temp, t1, words, sorted, (int)n, 0, 64, st) != hipSuccess)
        return CRN2_EHIP;
    hipLaunchKernelGGL(split_words_kernel, dim3((n + 255) / 256), dim3(256), 0, st, sorted, n, wbits, sel, w);
    size_t t3 = t2;
    if (hipcub::DeviceReduce::ReduceByKey(temp, t3, sel, usel, w, uw, nrun, SatAdd(), (int)n, st) != hipSuccess)
        return CRN2_EHIP;
    hipLaunchKernelGGL(sel_vectors_kernel, dim3((n + 255) / 256), dim3(256), 0, st, usel, nrun, bits, vec);
    uint32_t nu = 0;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&nu, nrun, sizeof(nu), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return CRN2_EHIP;
    int rc = crn2_tree_codebook_device(vec, uw, nu, 16, max_size, cb, k, nullptr, st);
    if (rc != CRN2_OK) return rc;
    hipLaunchKernelGGL(quantize_kernel, dim3((*k + 255) / 256), dim3(256), 0, st, cb, *k, bits, codebook);
    return hipGetLastError() == hipSuccess ? CRN2_OK : CRN2_EHIP;
}

}  // namespace

extern "C" int crn2_hc_color_selector_codebook_device(const uint64_t* block_selectors, uint32_t n_blocks,
                                                      uint32_t codebook_size, const void* blocks,
                                                      const uint32_t* endpoint_index, const void* cluster_colors,
                                                      int perceptual, uint32_t* selectors, uint32_t* n_selectors,
                                                      uint32_t* selector_index, uint8_t* used, void* stream)
{
    if (!block_selectors || !n_blocks || !codebook_size || !blocks || !endpoint_index || !cluster_colors ||
        !selectors || !n_selectors || !selector_index || !used)
        return CRN2_EINVAL;
    Scratch sc;
    sc.st = static_cast<hipStream_t>(stream);
    uint32_t k = 0;
    int rc = build_codebook(sc, block_selectors, n_blocks, 2, codebook_size, selectors, &k);
    if (rc != CRN2_OK) return rc;
    uint32_t* details = sc.get<uint32_t>((size_t)k * 64);
    if (!details) return CRN2_ENOMEM;
    if (hipMemsetAsync(details, 0, (size_t)k * 64 * sizeof(uint32_t), sc.st) != hipSuccess ||
        hipMemsetAsync(used, 0, k, sc.st) != hipSuccess)
        return CRN2_EHIP;
    rc = crn2_hc_color_selectors_device(blocks, endpoint_index, cluster_colors, n_blocks, 0, perceptual, selectors, k,
                                        selector_index, nullptr, details, used, sc.st);
    if (rc != CRN2_OK) return rc;
    rc = crn2_hc_selector_refine_device(details, k, 2, selectors, sc.st);
    if (rc != CRN2_OK) return rc;
    *n_selectors = k;
    return hipStreamSynchronize(sc.st) == hipSuccess ? CRN2_OK : CRN2_EHIP;
}

extern "C" int crn2_hc_alpha_selector_codebook_device(const uint64_t* block_selectors, uint32_t n_blocks,
                                                      uint32_t n_channels, uint32_t codebook_size,
                                                      const void* block_vals, const uint32_t* endpoint_index,
                                                      const void* alpha_values, const void* refined_values,
                                                      uint64_t* selectors, uint32_t* n_selectors,
                                                      uint32_t* selector_index, uint8_t* used, void* stream)
{
    if (!block_selectors || !n_blocks || !codebook_size || (n_channels != 1 && n_channels != 2) || !block_vals ||
        !endpoint_index || !alpha_values || !selectors || !n_selectors || !selector_index || !used)
        return CRN2_EINVAL;
    Scratch sc;
    sc.st = static_cast<hipStream_t>(stream);
    uint32_t k = 0;
    int rc = build_codebook(sc, block_selectors, n_blocks * n_channels, 3, codebook_size, selectors, &k);
    if (rc != CRN2_OK) return rc;
    uint32_t* details = sc.get<uint32_t>((size_t)k * 128);
    if (!details) return CRN2_ENOMEM;
    if (hipMemsetAsync(details, 0, (size_t)k * 128 * sizeof(uint32_t), sc.st) != hipSuccess ||
        hipMemsetAsync(used, 0, k, sc.st) != hipSuccess)
        return CRN2_EHIP;
    for (uint32_t c = 0; c < n_channels; c++) {
        rc = crn2_hc_alpha_selectors_device(static_cast<const uint8_t*>(block_vals) + (size_t)c * n_blocks * 16,
                                            endpoint_index + (size_t)c * n_blocks, alpha_values, refined_values,
                                            n_blocks, selectors, k, selector_index + (size_t)c * n_blocks, nullptr,
                                            details, used, sc.st);
        if (rc != CRN2_OK) return rc;
    }
    rc = crn2_hc_selector_refine_device(details, k, 3, selectors, sc.st);
    if (rc != CRN2_OK) return rc;
    *n_selectors = k;
    return hipStreamSynchronize(sc.st) == hipSuccess ? CRN2_OK : CRN2_EHIP;
}
