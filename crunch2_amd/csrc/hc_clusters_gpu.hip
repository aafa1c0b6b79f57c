// hc_clusters_gpu.hip -- dxt_hc's endpoint clustering on gfx950:
// determine_color_endpoints (crnlib/crn_dxt_hc.cpp:888-972, vec6F tile
// endpoints) and determine_alpha_endpoints (:1160-1239, vec2F, both alpha
// channels' tile endpoints together).
//
//  1. merge: the (vector, weight) pairs sorted ascending -- components
//     lexicographically, as vec::operator< orders them -- by a stable radix
//     sort per component from the last to the first; equal vectors merged,
//     weights summed and held at 2^32 - 1 (:930-946, :1211-1227);
//  2. the tree clusterizer's codebook of the distinct vectors
//     (crn2_tree_codebook_device, at most codebook_size entries);
//  3. every input vector assigned its nearest codebook entry, lowest index on
//     ties (determine_color_endpoint_clusters_task :836-886, whose partial-
//     distance pruning against the vector's own leaf never skips a nearest
//     entry, and determine_alpha_endpoint_clusters_task :1130-1158).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdint>

#include "crn2_internal.h"

extern "C" int crn2_tree_codebook_device(const float* vectors, const uint32_t* weights, uint32_t n, uint32_t dim,
                                         uint32_t max_splits, float* codebook, uint32_t* codebook_size,
                                         uint32_t* vector_leaf, void* stream);

namespace {

struct SatAdd {
    __host__ __device__ uint32_t operator()(uint32_t a, uint32_t b) const
    {
        return a > 0xFFFFFFFFu - b ? 0xFFFFFFFFu : a + b;
    }
};

__global__ void __launch_bounds__(256) iota_kernel(uint32_t* __restrict__ p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

// keys[i] = component c of vector order[i]
__global__ void __launch_bounds__(256) gather_comp_kernel(const float* __restrict__ vec, uint32_t dim, uint32_t c,
                                                         const uint32_t* __restrict__ order, uint32_t n,
                                                         float* __restrict__ keys)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = vec[(size_t)order[i] * dim + c];
}

// run[i] = 1 where sorted vector i differs (float !=) from vector i - 1
__global__ void __launch_bounds__(256) heads_kernel(const float* __restrict__ vec, uint32_t dim,
                                                   const uint32_t* __restrict__ order, uint32_t n,
                                                   const uint32_t* __restrict__ w, uint32_t* __restrict__ head,
                                                   uint32_t* __restrict__ wsorted)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t h = i == 0;
    if (i) {
        const float* a = vec + (size_t)order[i] * dim;
        const float* b = vec + (size_t)order[i - 1] * dim;
        for (uint32_t k = 0; k < dim; k++) h |= a[k] != b[k];
    }
    head[i] = h;
    wsorted[i] = w[order[i]];
}

// run ids (inclusive scan of heads, from 1) -> the first vector of each run
__global__ void __launch_bounds__(256) unique_vec_kernel(const float* __restrict__ vec, uint32_t dim,
                                                        const uint32_t* __restrict__ order,
                                                        const uint32_t* __restrict__ head,
                                                        const uint32_t* __restrict__ run, uint32_t n,
                                                        float* __restrict__ uvec)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !head[i]) return;
    for (uint32_t k = 0; k < dim; k++) uvec[(size_t)(run[i] - 1) * dim + k] = vec[(size_t)order[i] * dim + k];
}

// Nearest codebook entry of every vector, lowest index on ties; the codebook
// is staged in LDS in slices.
template <int D>
__global__ void __launch_bounds__(256) nearest_kernel(const float* __restrict__ vec, uint32_t n,
                                                     const float* __restrict__ cb, uint32_t k,
                                                     uint32_t* __restrict__ out)
{
    constexpr uint32_t kSlice = 2048;
    __shared__ float s[kSlice * D];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    float v[D];
    for (int c = 0; c < D; c++) v[c] = i < n ? vec[(size_t)i * D + c] : 0.0f;
    float best = 0.0f;
    uint32_t bi = 0;
    bool have = false;
    for (uint32_t s0 = 0; s0 < k; s0 += kSlice) {
        const uint32_t sn = min(kSlice, k - s0);
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < sn * D; j += blockDim.x) s[j] = cb[(size_t)s0 * D + j];
        __syncthreads();
        for (uint32_t j = 0; j < sn; j++) {
            float d = 0.0f;
#pragma unroll
            for (int c = 0; c < D; c++) {
                const float t = s[j * D + c] - v[c];
                d += t * t;
            }
            if (!have || d < best) {
                best = d;
                bi = s0 + j;
                have = true;
            }
        }
    }
    if (i < n) out[i] = bi;
}

struct Scratch {
    hipStream_t st;
    void* p[16] = {};
    int n = 0;
    template <class T>
    T* get(size_t count)
    {
        void* q = nullptr;
        if (n >= 16 || hipMallocAsync(&q, count ? count * sizeof(T) : 1, st) != hipSuccess) return nullptr;
        p[n++] = q;
        return static_cast<T*>(q);
    }
    ~Scratch()
    {
        for (int i = 0; i < n; i++) (void)hipFreeAsync(p[i], st);
        (void)hipStreamSynchronize(st);
    }
};

}  // namespace

extern "C" int crn2_hc_endpoint_clusters_device(const float* vectors, const uint32_t* weights, uint32_t n,
                                                uint32_t dim, uint32_t codebook_size, uint32_t* cluster,
                                                float* codebook, uint32_t* n_clusters, void* stream)
{
    if (!vectors || !weights || !n || !codebook_size || !cluster || !n_clusters || (dim != 2 && dim != 6))
        return CRN2_EINVAL;
    Scratch sc;
    sc.st = static_cast<hipStream_t>(stream);
    hipStream_t st = sc.st;
    uint32_t* order = sc.get<uint32_t>(n);
    uint32_t* order2 = sc.get<uint32_t>(n);
    float* keys = sc.get<float>(n);
    float* keys2 = sc.get<float>(n);
    uint32_t* head = sc.get<uint32_t>(n);
    uint32_t* run = sc.get<uint32_t>(n);
    uint32_t* ws = sc.get<uint32_t>(n);
    uint32_t* uid = sc.get<uint32_t>(n);
    uint32_t* uw = sc.get<uint32_t>(n);
    uint32_t* nrun = sc.get<uint32_t>(1);
    float* uvec = sc.get<float>((size_t)n * dim);
    float* cb = codebook ? codebook : sc.get<float>((size_t)codebook_size * dim);
    if (!order || !order2 || !keys || !keys2 || !head || !run || !ws || !uid || !uw || !nrun || !uvec || !cb)
        return CRN2_ENOMEM;
    size_t t1 = 0, t2 = 0, t3 = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, t1, keys, keys2, order, order2, (int)n, 0, 32, st) !=
            hipSuccess ||
        hipcub::DeviceScan::InclusiveSum(nullptr, t2, head, run, (int)n, st) != hipSuccess ||
        hipcub::DeviceReduce::ReduceByKey(nullptr, t3, run, uid, ws, uw, nrun, SatAdd(), (int)n, st) != hipSuccess)
        return CRN2_EHIP;
    size_t tb = t1 > t2 ? t1 : t2;
    if (t3 > tb) tb = t3;
    uint8_t* temp = sc.get<uint8_t>(tb);
    if (!temp) return CRN2_ENOMEM;
    const dim3 grid((n + 255) / 256), blk(256);
    hipLaunchKernelGGL(iota_kernel, grid, blk, 0, st, order, n);
    // stable LSD passes: last component first
    for (int c = (int)dim - 1; c >= 0; c--) {
        hipLaunchKernelGGL(gather_comp_kernel, grid, blk, 0, st, vectors, dim, (uint32_t)c, order, n, keys);
        size_t t = t1;
        if (hipcub::DeviceRadixSort::SortPairs(temp, t, keys, keys2, order, order2, (int)n, 0, 32, st) != hipSuccess)
            return CRN2_EHIP;
        uint32_t* x = order;
        order = order2;
        order2 = x;
    }
    hipLaunchKernelGGL(heads_kernel, grid, blk, 0, st, vectors, dim, order, n, weights, head, ws);
    size_t t = t2;
    if (hipcub::DeviceScan::InclusiveSum(temp, t, head, run, (int)n, st) != hipSuccess) return CRN2_EHIP;
    t = t3;
    if (hipcub::DeviceReduce::ReduceByKey(temp, t, run, uid, ws, uw, nrun, SatAdd(), (int)n, st) != hipSuccess)
        return CRN2_EHIP;
    hipLaunchKernelGGL(unique_vec_kernel, grid, blk, 0, st, vectors, dim, order, head, run, n, uvec);
    uint32_t nu = 0;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&nu, nrun, sizeof(nu), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return CRN2_EHIP;
    uint32_t k = 0;
    int rc = crn2_tree_codebook_device(uvec, uw, nu, dim, codebook_size, cb, &k, nullptr, st);
    if (rc != CRN2_OK) return rc;
    if (dim == 2) hipLaunchKernelGGL(nearest_kernel<2>, grid, blk, 0, st, vectors, n, cb, k, cluster);
    else hipLaunchKernelGGL(nearest_kernel<6>, grid, blk, 0, st, vectors, n, cb, k, cluster);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) return CRN2_EHIP;
    *n_clusters = k;
    return CRN2_OK;
}
