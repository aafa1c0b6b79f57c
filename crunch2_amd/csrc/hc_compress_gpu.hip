// hc_compress_gpu.hip -- dxt_hc::compress (crnlib/crn_dxt_hc.cpp:97-316) on
// gfx950: the whole hierarchical quantizer from texel blocks to palettes and
// per-block palette indices, every stage on the device.
//
//   1. adaptive tiles (crn2_hc_tiles_device);
//   2. the kept tiles listed in slot order (the reference's m_tiles order);
//   3. colour endpoint clustering of the tiles' vec6F endpoints, weights
//      (uint)(texels * level weight) (:888-972), the clusters' texel lists in
//      tile order and block lists in block order, then the colour endpoint
//      codebook with its refiner (crn2_hc_color_endpoints_device);
//   4. alpha endpoint clustering of every alpha channel's vec2F endpoints,
//      weights = texels (:1160-1239), then the alpha endpoint codebook per
//      channel (crn2_hc_alpha_endpoints_device);
//   5. colour and alpha selector codebooks (crn2_hc_*_selector_codebook_device);
//   6. the output (:210-312): endpoints of clusters with texels and used
//      selectors, each list without repeats in first-appearance order, and per
//      block the remapped indices with its reference code (1 = same endpoints
//      as the previous block -- at a row start the previous row's last --,
//      2 = as the block above, 0 = new).
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "crn2_internal.h"

extern "C" {
int crn2_hc_tiles_device(const void* blocks, uint32_t num_blocks, uint32_t nlevels, const uint32_t* level_desc,
                         uint32_t num_faces, int has_color, uint32_t num_alpha, const uint32_t* alpha_comp,
                         int perceptual, float color_derating, float alpha_derating, float ratio,
                         uint8_t* block_encodings, uint32_t* tile_indices, uint32_t* tile_src, float* tile_vec,
                         void* stream);
int crn2_hc_endpoint_clusters_device(const float* vectors, const uint32_t* weights, uint32_t n, uint32_t dim,
                                     uint32_t codebook_size, uint32_t* cluster, float* codebook,
                                     uint32_t* n_clusters, void* stream);
int crn2_hc_color_endpoints_device(const void* cluster_px, const uint32_t* cluster_ofs, uint32_t nclusters,
                                   uint32_t total_px, const uint32_t* block_cluster, uint32_t nblocks,
                                   const void* blocks, const float* block_weights, const uint8_t* block_encodings,
                                   int perceptual, uint32_t* endpoints, void* color_values, uint64_t* block_selectors,
                                   uint32_t* not_encoded, void* stream);
int crn2_hc_alpha_endpoints_device(const void* cluster_vals, const uint32_t* cluster_ofs, uint32_t nclusters,
                                   uint32_t total_vals, const uint32_t* block_cluster, uint32_t nblocks,
                                   const void* block_vals, const uint8_t* block_encodings, uint32_t* endpoints,
                                   void* alpha_values, void* refined_values, uint64_t* block_selectors, void* stream);
int crn2_hc_color_selector_codebook_device(const uint64_t* block_selectors, uint32_t n_blocks, uint32_t codebook_size,
                                           const void* blocks, const uint32_t* endpoint_index,
                                           const void* cluster_colors, int perceptual, uint32_t* selectors,
                                           uint32_t* n_selectors, uint32_t* selector_index, uint8_t* used,
                                           void* stream);
int crn2_hc_alpha_selector_codebook_device(const uint64_t* block_selectors, uint32_t n_blocks, uint32_t n_channels,
                                           uint32_t codebook_size, const void* block_vals,
                                           const uint32_t* endpoint_index, const void* alpha_values,
                                           const void* refined_values, uint64_t* selectors, uint32_t* n_selectors,
                                           uint32_t* selector_index, uint8_t* used, void* stream);
}

namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;

struct Levels {
    uint32_t n;
    uint32_t first[16], width[16];
    float weight[16];
};

__device__ inline uint32_t level_of(const Levels& L, uint32_t b)
{
    uint32_t l = 0;
    for (uint32_t i = 1; i < L.n; i++)
        if (b >= L.first[i]) l = i;
    return l;
}

constexpr uint32_t kOffsets[9] = {0, 16, 32, 48, 0, 32, 64, 96, 64};

__device__ inline uint32_t tile_texel(const uint32_t* blocks, uint32_t b, uint32_t w, uint32_t t, uint32_t i)
{
    const uint32_t idx = kOffsets[t] + i;
    if (idx < 64) return blocks[(size_t)(b + ((idx & 16) ? w : 0) + ((idx & 32) ? 1 : 0)) * 16 + (idx & 15)];
    const uint32_t j = idx - 64;
    return blocks[(size_t)(b + ((j & 32) ? w : 0) + ((j & 4) ? 1 : 0)) * 16 + ((j >> 1) & 12) + (j & 3)];
}

__global__ void __launch_bounds__(256) used_kernel(const uint32_t* __restrict__ src, uint32_t n,
                                                  uint32_t* __restrict__ flag)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = src[i] != kNone;
}

// The kept tiles in slot order: slot, texel count, colour / alpha vectors and weights.
__global__ void __launch_bounds__(256) tile_list_kernel(const uint32_t* __restrict__ src,
                                                       const float* __restrict__ tvec,
                                                       const uint32_t* __restrict__ pos, uint32_t n, Levels L,
                                                       uint32_t ntiles, int num_alpha, uint32_t* __restrict__ slot,
                                                       uint32_t* __restrict__ size, float* __restrict__ cvec,
                                                       uint32_t* __restrict__ cw, float* __restrict__ avec,
                                                       uint32_t* __restrict__ aw)
{
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n || src[s] == kNone) return;
    const uint32_t i = pos[s], b = src[s] >> 4, t = src[s] & 15u;
    const uint32_t sz = 16u << (t >> 2);
    slot[i] = s;
    size[i] = sz;
    for (int k = 0; k < 6; k++) cvec[(size_t)i * 6 + k] = tvec[(size_t)s * 10 + k];
    cw[i] = (uint32_t)((float)sz * L.weight[level_of(L, b)]);
    for (int a = 0; a < num_alpha; a++) {
        const size_t j = (size_t)a * ntiles + i;
        avec[j * 2] = tvec[(size_t)s * 10 + 6 + 2 * a];
        avec[j * 2 + 1] = tvec[(size_t)s * 10 + 7 + 2 * a];
        aw[j] = sz;
    }
}

__global__ void __launch_bounds__(256) iota_kernel(uint32_t* __restrict__ p, uint32_t n)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

__global__ void __launch_bounds__(256) cluster_count_kernel(const uint32_t* __restrict__ cl,
                                                           const uint32_t* __restrict__ size, uint32_t m,
                                                           uint32_t ntiles, uint32_t* __restrict__ count)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m) atomicAdd(&count[cl[i]], size[i % ntiles]);
}

__global__ void __launch_bounds__(256) sorted_size_kernel(const uint32_t* __restrict__ order, uint32_t m,
                                                         uint32_t ntiles, const uint32_t* __restrict__ size,
                                                         uint32_t* __restrict__ out)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < m) out[p] = size[order[p] % ntiles];
}

// Copies each entry's tile texels (colour words, or one channel's bytes) to
// its place in the clusters' concatenated lists.
__global__ void __launch_bounds__(256) gather_px_kernel(const uint32_t* __restrict__ blocks,
                                                       const uint32_t* __restrict__ order,
                                                       const uint32_t* __restrict__ off, uint32_t m, uint32_t ntiles,
                                                       const uint32_t* __restrict__ slot,
                                                       const uint32_t* __restrict__ src,
                                                       const uint32_t* __restrict__ swidth, int alpha,
                                                       const uint32_t* __restrict__ comp, uint32_t* __restrict__ cpx,
                                                       uint8_t* __restrict__ apx)
{
    const uint32_t p = blockIdx.x;
    if (p >= m) return;
    const uint32_t e = order[p], i = e % ntiles, s = slot[i];
    const uint32_t b = src[s] >> 4, t = src[s] & 15u, n = 16u << (t >> 2), w = swidth[s];
    for (uint32_t k = threadIdx.x; k < n; k += blockDim.x) {
        const uint32_t px = tile_texel(blocks, b, w, t, k);
        if (alpha) apx[off[p] + k] = (uint8_t)(px >> (8 * comp[e / ntiles]));
        else cpx[off[p] + k] = px;
    }
}

// Per block: level weight, its tile's colour cluster and per alpha channel
// its tile's alpha cluster, and that channel's 16 values.
__global__ void __launch_bounds__(256) block_info_kernel(const uint32_t* __restrict__ blocks,
                                                        const uint32_t* __restrict__ tile_idx,
                                                        const uint32_t* __restrict__ pos, uint32_t n, Levels L,
                                                        uint32_t ntiles, const uint32_t* __restrict__ ccl,
                                                        const uint32_t* __restrict__ acl, int has_color, int num_alpha,
                                                        const uint32_t* __restrict__ comp, float* __restrict__ bw,
                                                        uint32_t* __restrict__ bcc, uint32_t* __restrict__ bac,
                                                        uint8_t* __restrict__ bav)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    bw[b] = L.weight[level_of(L, b)];
    const uint32_t i = pos[tile_idx[b]];
    if (has_color) bcc[b] = ccl[i];
    for (int a = 0; a < num_alpha; a++) {
        bac[(size_t)a * n + b] = acl[(size_t)a * ntiles + i];
        for (int p = 0; p < 16; p++)
            bav[((size_t)a * n + b) * 16 + p] = (uint8_t)(blocks[(size_t)b * 16 + p] >> (8 * comp[a]));
    }
}

// ---------------------------------------------------------- dedupe

__global__ void __launch_bounds__(256) ep_keys_kernel(const uint32_t* __restrict__ ep, const uint32_t* __restrict__ ofs,
                                                     uint32_t k, uint32_t mask, uint64_t* __restrict__ key,
                                                     uint32_t* __restrict__ valid)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    key[i] = ep[i] & mask;
    valid[i] = ofs[i + 1] > ofs[i];
}

__global__ void __launch_bounds__(256) sel_keys_kernel(const void* __restrict__ sel, int wide,
                                                      const uint8_t* __restrict__ used, uint32_t k,
                                                      uint64_t* __restrict__ key, uint32_t* __restrict__ valid)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k) return;
    key[i] = wide ? static_cast<const uint64_t*>(sel)[i] : static_cast<const uint32_t*>(sel)[i];
    valid[i] = used[i] != 0;
}

__global__ void __launch_bounds__(256) compact_kernel(const uint32_t* __restrict__ valid,
                                                     const uint32_t* __restrict__ vpos, const uint64_t* __restrict__ key,
                                                     uint32_t k, uint32_t* __restrict__ list, uint64_t* __restrict__ lkey,
                                                     uint32_t* __restrict__ lidx)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= k || !valid[i]) return;
    list[vpos[i]] = i;
    lkey[vpos[i]] = key[i];
    lidx[vpos[i]] = vpos[i];
}

__global__ void __launch_bounds__(256) run_heads_kernel(const uint64_t* __restrict__ sk, uint32_t m,
                                                       uint32_t* __restrict__ head)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < m) head[p] = p == 0 || sk[p] != sk[p - 1];
}

__global__ void __launch_bounds__(256) run_first_kernel(const uint32_t* __restrict__ head,
                                                       const uint32_t* __restrict__ run,
                                                       const uint32_t* __restrict__ sj, uint32_t m,
                                                       uint32_t* __restrict__ first)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < m && head[p]) first[run[p] - 1] = sj[p];
}

__global__ void __launch_bounds__(256) rep_kernel(const uint32_t* __restrict__ run, const uint32_t* __restrict__ sj,
                                                 const uint32_t* __restrict__ first, uint32_t m,
                                                 uint32_t* __restrict__ rep, uint32_t* __restrict__ isrep)
{
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= m) return;
    const uint32_t j = sj[p], r = first[run[p] - 1];
    rep[j] = r;
    isrep[j] = r == j;
}

__global__ void __launch_bounds__(256) remap_kernel(const uint32_t* __restrict__ list, const uint64_t* __restrict__ lkey,
                                                   const uint32_t* __restrict__ rep, const uint32_t* __restrict__ isrep,
                                                   const uint32_t* __restrict__ opos, uint32_t m, int wide,
                                                   uint32_t* __restrict__ remap, void* __restrict__ out)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= m) return;
    remap[list[j]] = opos[rep[j]];
    if (isrep[j]) {
        if (wide) static_cast<uint64_t*>(out)[opos[j]] = lkey[j];
        else static_cast<uint32_t*>(out)[opos[j]] = (uint32_t)lkey[j];
    }
}

// Per block: remapped indices and the reference code (crn_dxt_hc.cpp:284-308).
__global__ void __launch_bounds__(256) indices_kernel(uint32_t n, Levels L, int has_color, int num_alpha,
                                                     const uint32_t* __restrict__ bcc, const uint32_t* __restrict__ bac,
                                                     const uint32_t* __restrict__ csel, const uint32_t* __restrict__ asel,
                                                     const uint32_t* __restrict__ cep_remap,
                                                     const uint32_t* __restrict__ aep_remap,
                                                     const uint32_t* __restrict__ csel_remap,
                                                     const uint32_t* __restrict__ asel_remap,
                                                     uint16_t* __restrict__ ep_out, uint16_t* __restrict__ sel_out)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const uint32_t l = level_of(L, b), rel = b - L.first[l], w = L.width[l];
    const uint32_t bx = rel % w, by = rel / w;
    bool top = by != 0;
    bool left = top || bx;
    uint16_t ep[3] = {0, 0, 0}, sl[3] = {0, 0, 0};
    auto epi = [&](uint32_t blk, int c) -> uint16_t {
        return (uint16_t)(c == 0 ? cep_remap[bcc[blk]] : aep_remap[bac[(size_t)(c - 1) * n + blk]]);
    };
    for (int c = has_color ? 0 : 1; c < 1 + num_alpha; c++) {
        const uint16_t e = epi(b, c);
        left = left && e == epi(b - 1, c);
        top = top && e == epi(b - w, c);
        ep[c] = e;
        sl[c] = (uint16_t)(c == 0 ? csel_remap[csel[b]] : asel_remap[asel[(size_t)(c - 1) * n + b]]);
    }
    for (int c = 0; c < 3; c++) {
        ep_out[(size_t)b * 4 + c] = ep[c];
        sel_out[(size_t)b * 3 + c] = sl[c];
    }
    ep_out[(size_t)b * 4 + 3] = left ? 1 : top ? 2 : 0;
}

struct Scratch {
    hipStream_t st;
    std::vector<void*> p;
    template <class T>
    T* get(size_t count)
    {
        void* q = nullptr;
        if (hipMallocAsync(&q, count ? count * sizeof(T) : 1, st) != hipSuccess) return nullptr;
        p.push_back(q);
        return static_cast<T*>(q);
    }
    ~Scratch()
    {
        for (void* q : p) (void)hipFreeAsync(q, st);
        (void)hipStreamSynchronize(st);
    }
};

inline dim3 grid(uint64_t n) { return dim3((uint32_t)((n + 255) / 256)); }

// Entries keyed by key[0..k) with valid[] set: the keys without repeats in
// first-appearance order into out (u32 or u64), remap[i] = position of i's key.
int dedupe(Scratch& sc, const uint64_t* key, const uint32_t* valid, uint32_t k, int wide, uint32_t* remap, void* out,
           uint32_t* n_out)
{
    hipStream_t st = sc.st;
    uint32_t* vpos = sc.get<uint32_t>(k + 1);
    uint32_t* list = sc.get<uint32_t>(k);
    uint64_t* lkey = sc.get<uint64_t>(k);
    uint32_t* lidx = sc.get<uint32_t>(k);
    uint64_t* sk = sc.get<uint64_t>(k);
    uint32_t* sj = sc.get<uint32_t>(k);
    uint32_t* head = sc.get<uint32_t>(k);
    uint32_t* run = sc.get<uint32_t>(k);
    uint32_t* first = sc.get<uint32_t>(k);
    uint32_t* rep = sc.get<uint32_t>(k);
    uint32_t* isrep = sc.get<uint32_t>(k + 1);
    uint32_t* opos = sc.get<uint32_t>(k + 1);
    if (!vpos || !list || !lkey || !lidx || !sk || !sj || !head || !run || !first || !rep || !isrep || !opos)
        return CRN2_ENOMEM;
    size_t t1 = 0, t2 = 0, t3 = 0;
    if (hipcub::DeviceScan::ExclusiveSum(nullptr, t1, valid, vpos, (int)k + 1, st) != hipSuccess ||
        hipcub::DeviceRadixSort::SortPairs(nullptr, t2, lkey, sk, lidx, sj, (int)k, 0, 64, st) != hipSuccess ||
        hipcub::DeviceScan::InclusiveSum(nullptr, t3, head, run, (int)k, st) != hipSuccess)
        return CRN2_EHIP;
    size_t tb = t1 > t2 ? t1 : t2;
    if (t3 > tb) tb = t3;
    uint8_t* temp = sc.get<uint8_t>(tb);
    if (!temp) return CRN2_ENOMEM;
    // valid has k + 1 entries readable: the scan's last element counts them all
    size_t t = t1;
    if (hipcub::DeviceScan::ExclusiveSum(temp, t, valid, vpos, (int)k + 1, st) != hipSuccess) return CRN2_EHIP;
    uint32_t m = 0;
    if (hipMemcpyAsync(&m, vpos + k, 4, hipMemcpyDeviceToHost, st) != hipSuccess || hipStreamSynchronize(st) != hipSuccess)
        return CRN2_EHIP;
    if (!m) {
        *n_out = 0;
        return CRN2_OK;
    }
    hipLaunchKernelGGL(compact_kernel, grid(k), dim3(256), 0, st, valid, vpos, key, k, list, lkey, lidx);
    t = t2;
    if (hipcub::DeviceRadixSort::SortPairs(temp, t, lkey, sk, lidx, sj, (int)m, 0, 64, st) != hipSuccess)
        return CRN2_EHIP;
    hipLaunchKernelGGL(run_heads_kernel, grid(m), dim3(256), 0, st, sk, m, head);
    t = t3;
    if (hipcub::DeviceScan::InclusiveSum(temp, t, head, run, (int)m, st) != hipSuccess) return CRN2_EHIP;
    hipLaunchKernelGGL(run_first_kernel, grid(m), dim3(256), 0, st, head, run, sj, m, first);
    hipLaunchKernelGGL(rep_kernel, grid(m), dim3(256), 0, st, run, sj, first, m, rep, isrep);
    if (hipMemsetAsync(isrep + m, 0, 4, st) != hipSuccess) return CRN2_EHIP;
    t = t1;
    if (hipcub::DeviceScan::ExclusiveSum(temp, t, isrep, opos, (int)m + 1, st) != hipSuccess) return CRN2_EHIP;
    hipLaunchKernelGGL(remap_kernel, grid(m), dim3(256), 0, st, list, lkey, rep, isrep, opos, m, wide, remap, out);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(n_out, opos + m, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return CRN2_EHIP;
    return CRN2_OK;
}

// Clusters' concatenated texel lists for m entries (entry e = channel * ntiles
// + tile, cluster cl[e]): offsets per cluster (k + 1) and the texels.
int cluster_lists(Scratch& sc, const uint32_t* blocks, const uint32_t* cl, uint32_t m, uint32_t ntiles, uint32_t k,
                  const uint32_t* size, const uint32_t* slot, const uint32_t* src, const uint32_t* swidth, int alpha,
                  const uint32_t* comp, uint32_t* ofs, void* px, uint32_t total)
{
    hipStream_t st = sc.st;
    uint32_t* count = sc.get<uint32_t>(k + 1);
    uint32_t* idx = sc.get<uint32_t>(m);
    uint32_t* skey = sc.get<uint32_t>(m);
    uint32_t* order = sc.get<uint32_t>(m);
    uint32_t* ssz = sc.get<uint32_t>(m + 1);
    uint32_t* off = sc.get<uint32_t>(m + 1);
    if (!count || !idx || !skey || !order || !ssz || !off) return CRN2_ENOMEM;
    size_t t1 = 0, t2 = 0;
    if (hipcub::DeviceRadixSort::SortPairs(nullptr, t1, cl, skey, idx, order, (int)m, 0, 32, st) != hipSuccess ||
        hipcub::DeviceScan::ExclusiveSum(nullptr, t2, count, ofs, (int)(k > m ? k : m) + 1, st) != hipSuccess)
        return CRN2_EHIP;
    uint8_t* temp = sc.get<uint8_t>(t1 > t2 ? t1 : t2);
    if (!temp) return CRN2_ENOMEM;
    if (hipMemsetAsync(count, 0, (size_t)(k + 1) * 4, st) != hipSuccess) return CRN2_EHIP;
    hipLaunchKernelGGL(cluster_count_kernel, grid(m), dim3(256), 0, st, cl, size, m, ntiles, count);
    size_t t = t2;
    if (hipcub::DeviceScan::ExclusiveSum(temp, t, count, ofs, (int)k + 1, st) != hipSuccess) return CRN2_EHIP;
    hipLaunchKernelGGL(iota_kernel, grid(m), dim3(256), 0, st, idx, m);
    t = t1;
    if (hipcub::DeviceRadixSort::SortPairs(temp, t, cl, skey, idx, order, (int)m, 0, 32, st) != hipSuccess)
        return CRN2_EHIP;
    hipLaunchKernelGGL(sorted_size_kernel, grid(m), dim3(256), 0, st, order, m, ntiles, size, ssz);
    if (hipMemsetAsync(ssz + m, 0, 4, st) != hipSuccess) return CRN2_EHIP;
    t = t2;
    if (hipcub::DeviceScan::ExclusiveSum(temp, t, ssz, off, (int)m + 1, st) != hipSuccess) return CRN2_EHIP;
    hipLaunchKernelGGL(gather_px_kernel, dim3(m), dim3(64), 0, st, blocks, order, off, m, ntiles, slot, src, swidth,
                       alpha, comp, static_cast<uint32_t*>(px), static_cast<uint8_t*>(px));
    (void)total;
    return hipGetLastError() == hipSuccess ? CRN2_OK : CRN2_EHIP;
}

}  // namespace

extern "C" int crn2_hc_compress_device(const void* blocks_v, uint32_t num_blocks, uint32_t nlevels,
                                       const uint32_t* level_desc, const float* level_weight, uint32_t num_faces,
                                       int has_color, uint32_t num_alpha, const uint32_t* alpha_comp, int perceptual,
                                       const uint32_t* codebook_sizes, const float* tile_params,
                                       uint16_t* endpoint_indices, uint16_t* selector_indices,
                                       uint32_t* color_endpoints, uint32_t* alpha_endpoints,
                                       uint32_t* color_selectors, uint64_t* alpha_selectors, uint32_t* counts,
                                       void* stream)
{
    if (!blocks_v || !num_blocks || !nlevels || nlevels > 16 || !level_desc || !level_weight || !num_faces ||
        num_alpha > 2 || (!has_color && !num_alpha) || (num_alpha && !alpha_comp) || !codebook_sizes ||
        !endpoint_indices || !selector_indices || !counts || (has_color && (!color_endpoints || !color_selectors)) ||
        (num_alpha && (!alpha_endpoints || !alpha_selectors)))
        return CRN2_EINVAL;
    for (int i = 0; i < 4; i++)
        if (!codebook_sizes[i]) return CRN2_EINVAL;
    const float cder = tile_params ? tile_params[0] : 2.0f, ader = tile_params ? tile_params[1] : 2.0f,
                ratio = tile_params ? tile_params[2] : 3.0f;
    const uint32_t* blocks = static_cast<const uint32_t*>(blocks_v);
    const uint32_t n = num_blocks;
    Scratch sc;
    sc.st = static_cast<hipStream_t>(stream);
    hipStream_t st = sc.st;
    Levels L;
    L.n = nlevels;
    for (uint32_t l = 0; l < 16; l++) {
        L.first[l] = l < nlevels ? level_desc[3 * l] : 0;
        L.width[l] = l < nlevels ? level_desc[3 * l + 2] : 1;
        L.weight[l] = l < nlevels ? level_weight[l] : 0.0f;
    }
    uint32_t comp_h[2] = {num_alpha > 0 ? alpha_comp[0] : 0, num_alpha > 1 ? alpha_comp[1] : 0};
    for (int i = 0; i < 4; i++) counts[i] = 0;

    // 1. tiles
    uint8_t* enc = sc.get<uint8_t>(n);
    uint32_t* tidx = sc.get<uint32_t>(n);
    uint32_t* src = sc.get<uint32_t>(n);
    float* tvec = sc.get<float>((size_t)n * 10);
    uint32_t* flag = sc.get<uint32_t>(n + 1);
    uint32_t* pos = sc.get<uint32_t>(n + 1);
    uint32_t* comp = sc.get<uint32_t>(2);
    if (!enc || !tidx || !src || !tvec || !flag || !pos || !comp) return CRN2_ENOMEM;
    int rc = crn2_hc_tiles_device(blocks, n, nlevels, level_desc, num_faces, has_color, num_alpha, comp_h, perceptual,
                                  cder, ader, ratio, enc, tidx, src, tvec, st);
    if (rc != CRN2_OK) return rc;
    if (hipMemcpyAsync(comp, comp_h, 8, hipMemcpyHostToDevice, st) != hipSuccess) return CRN2_EHIP;

    // 2. kept tiles in slot order
    hipLaunchKernelGGL(used_kernel, grid(n), dim3(256), 0, st, src, n, flag);
    if (hipMemsetAsync(flag + n, 0, 4, st) != hipSuccess) return CRN2_EHIP;
    {
        size_t t = 0;
        if (hipcub::DeviceScan::ExclusiveSum(nullptr, t, flag, pos, (int)n + 1, st) != hipSuccess) return CRN2_EHIP;
        uint8_t* temp = sc.get<uint8_t>(t);
        if (!temp || hipcub::DeviceScan::ExclusiveSum(temp, t, flag, pos, (int)n + 1, st) != hipSuccess)
            return CRN2_EHIP;
    }
    uint32_t ntiles = 0;
    if (hipMemcpyAsync(&ntiles, pos + n, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return CRN2_EHIP;
    if (!ntiles) return CRN2_EINVAL;
    uint32_t* slot = sc.get<uint32_t>(ntiles);
    uint32_t* size = sc.get<uint32_t>(ntiles);
    float* cvec = sc.get<float>((size_t)ntiles * 6);
    uint32_t* cw = sc.get<uint32_t>(ntiles);
    float* avec = sc.get<float>((size_t)ntiles * 2 * (num_alpha ? num_alpha : 1));
    uint32_t* aw = sc.get<uint32_t>((size_t)ntiles * (num_alpha ? num_alpha : 1));
    uint32_t* swidth = sc.get<uint32_t>(n);
    if (!slot || !size || !cvec || !cw || !avec || !aw || !swidth) return CRN2_ENOMEM;
    hipLaunchKernelGGL(tile_list_kernel, grid(n), dim3(256), 0, st, src, tvec, pos, n, L, ntiles, (int)num_alpha, slot,
                       size, cvec, cw, avec, aw);
    {
        // level width of every slot's chunk (for regathering tile texels)
        std::vector<uint32_t> w(n, 0);
        for (uint32_t l = 0; l < nlevels; l++) {
            const uint32_t f = level_desc[3 * l], nb = level_desc[3 * l + 1], wd = level_desc[3 * l + 2];
            if ((uint64_t)f + nb > n) return CRN2_EINVAL;
            for (uint32_t b = f; b < f + nb; b++) w[b] = wd;
        }
        if (hipMemcpyAsync(swidth, w.data(), (size_t)n * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return CRN2_EHIP;
    }

    float* bw = sc.get<float>(n);
    uint32_t* bcc = sc.get<uint32_t>(n);
    uint32_t* bac = sc.get<uint32_t>((size_t)n * (num_alpha ? num_alpha : 1));
    uint8_t* bav = sc.get<uint8_t>((size_t)n * 16 * (num_alpha ? num_alpha : 1));
    uint32_t* ccl = sc.get<uint32_t>(ntiles);
    uint32_t* acl = sc.get<uint32_t>((size_t)ntiles * (num_alpha ? num_alpha : 1));
    if (!bw || !bcc || !bac || !bav || !ccl || !acl) return CRN2_ENOMEM;

    // 3. colour endpoints
    uint32_t kc = 0, ka = 0;
    uint32_t *cofs = nullptr, *cep = nullptr, *aofs = nullptr, *aep = nullptr;
    uint32_t* ccolors = nullptr;
    uint8_t *avals = nullptr, *rvals = nullptr;
    uint64_t *cbsel = nullptr, *absel = nullptr;
    if (has_color) {
        const uint32_t cap = ntiles < codebook_sizes[0] ? ntiles : codebook_sizes[0];
        rc = crn2_hc_endpoint_clusters_device(cvec, cw, ntiles, 6, cap, ccl, nullptr, &kc, st);
        if (rc != CRN2_OK) return rc;
    }
    if (num_alpha) {
        const uint32_t cap = ntiles < codebook_sizes[2] ? ntiles : codebook_sizes[2];
        rc = crn2_hc_endpoint_clusters_device(avec, aw, ntiles * num_alpha, 2, cap, acl, nullptr, &ka, st);
        if (rc != CRN2_OK) return rc;
    }
    hipLaunchKernelGGL(block_info_kernel, grid(n), dim3(256), 0, st, blocks, tidx, pos, n, L, ntiles, ccl, acl,
                       has_color, (int)num_alpha, comp, bw, bcc, bac, bav);
    if (hipGetLastError() != hipSuccess) return CRN2_EHIP;
    if (has_color) {
        uint64_t total = 0;
        {
            std::vector<uint32_t> hs(ntiles);
            if (hipMemcpyAsync(hs.data(), size, (size_t)ntiles * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                return CRN2_EHIP;
            for (uint32_t v : hs) total += v;
        }
        cofs = sc.get<uint32_t>((size_t)(kc > ntiles ? kc : ntiles) + 1);
        uint32_t* cpx = sc.get<uint32_t>(total);
        cep = sc.get<uint32_t>(kc);
        ccolors = sc.get<uint32_t>((size_t)kc * 4);
        cbsel = sc.get<uint64_t>(n);
        uint32_t* ne = sc.get<uint32_t>(1);
        if (!cofs || !cpx || !cep || !ccolors || !cbsel || !ne) return CRN2_ENOMEM;
        rc = cluster_lists(sc, blocks, ccl, ntiles, ntiles, kc, size, slot, src, swidth, 0, comp, cofs, cpx,
                           (uint32_t)total);
        if (rc != CRN2_OK) return rc;
        if (hipMemsetAsync(cbsel, 0, (size_t)n * 8, st) != hipSuccess) return CRN2_EHIP;
        rc = crn2_hc_color_endpoints_device(cpx, cofs, kc, (uint32_t)total, bcc, n, blocks, bw, enc, perceptual, cep,
                                            ccolors, cbsel, ne, st);
        if (rc != CRN2_OK) return rc;
        uint32_t nne = 0;
        if (hipMemcpyAsync(&nne, ne, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return CRN2_EHIP;
        if (nne) return CRN2_EUNSUPPORTED;     // a cluster with more distinct colours than the search holds
    }

    // 4. alpha endpoints
    if (num_alpha) {
        uint64_t total = 0;
        {
            std::vector<uint32_t> hs(ntiles);
            if (hipMemcpyAsync(hs.data(), size, (size_t)ntiles * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                return CRN2_EHIP;
            for (uint32_t v : hs) total += v;
            total *= num_alpha;
        }
        const uint32_t m = ntiles * num_alpha;
        aofs = sc.get<uint32_t>((size_t)(ka > m ? ka : m) + 1);
        uint8_t* apx = sc.get<uint8_t>(total);
        aep = sc.get<uint32_t>(ka);
        avals = sc.get<uint8_t>((size_t)ka * 8);
        rvals = sc.get<uint8_t>((size_t)ka * 8);
        absel = sc.get<uint64_t>((size_t)n * num_alpha);
        if (!aofs || !apx || !aep || !avals || !rvals || !absel) return CRN2_ENOMEM;
        rc = cluster_lists(sc, blocks, acl, m, ntiles, ka, size, slot, src, swidth, 1, comp, aofs, apx,
                           (uint32_t)total);
        if (rc != CRN2_OK) return rc;
        if (hipMemsetAsync(absel, 0, (size_t)n * num_alpha * 8, st) != hipSuccess) return CRN2_EHIP;
        for (uint32_t a = 0; a < num_alpha; a++) {
            rc = crn2_hc_alpha_endpoints_device(apx, aofs, ka, (uint32_t)total, bac + (size_t)a * n, n,
                                                bav + (size_t)a * n * 16, enc, aep, avals, rvals,
                                                absel + (size_t)a * n, st);
            if (rc != CRN2_OK) return rc;
        }
    }

    // 5. selector codebooks
    uint32_t ncs = 0, nas = 0;
    uint32_t *csel = nullptr, *cs_idx = nullptr, *asel_idx = nullptr;
    uint8_t *cused = nullptr, *aused = nullptr;
    uint64_t* asel = nullptr;
    if (has_color) {
        csel = sc.get<uint32_t>(codebook_sizes[1]);
        cs_idx = sc.get<uint32_t>(n);
        cused = sc.get<uint8_t>(codebook_sizes[1]);
        if (!csel || !cs_idx || !cused) return CRN2_ENOMEM;
        rc = crn2_hc_color_selector_codebook_device(cbsel, n, codebook_sizes[1], blocks, bcc, ccolors, perceptual, csel,
                                                    &ncs, cs_idx, cused, st);
        if (rc != CRN2_OK) return rc;
    }
    if (num_alpha) {
        asel = sc.get<uint64_t>(codebook_sizes[3]);
        asel_idx = sc.get<uint32_t>((size_t)n * num_alpha);
        aused = sc.get<uint8_t>(codebook_sizes[3]);
        if (!asel || !asel_idx || !aused) return CRN2_ENOMEM;
        rc = crn2_hc_alpha_selector_codebook_device(absel, n, num_alpha, codebook_sizes[3], bav, bac, avals, rvals,
                                                    asel, &nas, asel_idx, aused, st);
        if (rc != CRN2_OK) return rc;
    }

    // 6. output lists and per-block indices
    const uint32_t kmax = std::max(std::max(kc, ka), std::max(ncs, nas)) + 1;
    uint64_t* key = sc.get<uint64_t>(kmax);
    uint32_t* valid = sc.get<uint32_t>(kmax);
    uint32_t* cep_remap = sc.get<uint32_t>(kc + 1);
    uint32_t* aep_remap = sc.get<uint32_t>(ka + 1);
    uint32_t* csel_remap = sc.get<uint32_t>(ncs + 1);
    uint32_t* asel_remap = sc.get<uint32_t>(nas + 1);
    if (!key || !valid || !cep_remap || !aep_remap || !csel_remap || !asel_remap) return CRN2_ENOMEM;
    if (has_color) {
        hipLaunchKernelGGL(ep_keys_kernel, grid(kc), dim3(256), 0, st, cep, cofs, kc, 0xFFFFFFFFu, key, valid);
        if (hipMemsetAsync(valid + kc, 0, 4, st) != hipSuccess) return CRN2_EHIP;
        rc = dedupe(sc, key, valid, kc, 0, cep_remap, color_endpoints, &counts[0]);
        if (rc != CRN2_OK) return rc;
        hipLaunchKernelGGL(sel_keys_kernel, grid(ncs), dim3(256), 0, st, csel, 0, cused, ncs, key, valid);
        if (hipMemsetAsync(valid + ncs, 0, 4, st) != hipSuccess) return CRN2_EHIP;
        rc = dedupe(sc, key, valid, ncs, 0, csel_remap, color_selectors, &counts[1]);
        if (rc != CRN2_OK) return rc;
    }
    if (num_alpha) {
        hipLaunchKernelGGL(ep_keys_kernel, grid(ka), dim3(256), 0, st, aep, aofs, ka, 0xFFFFu, key, valid);
        if (hipMemsetAsync(valid + ka, 0, 4, st) != hipSuccess) return CRN2_EHIP;
        rc = dedupe(sc, key, valid, ka, 0, aep_remap, alpha_endpoints, &counts[2]);
        if (rc != CRN2_OK) return rc;
        hipLaunchKernelGGL(sel_keys_kernel, grid(nas), dim3(256), 0, st, asel, 1, aused, nas, key, valid);
        if (hipMemsetAsync(valid + nas, 0, 4, st) != hipSuccess) return CRN2_EHIP;
        rc = dedupe(sc, key, valid, nas, 1, asel_remap, alpha_selectors, &counts[3]);
        if (rc != CRN2_OK) return rc;
    }
    hipLaunchKernelGGL(indices_kernel, grid(n), dim3(256), 0, st, n, L, has_color, (int)num_alpha, bcc, bac, cs_idx,
                       asel_idx, cep_remap, aep_remap, csel_remap, asel_remap, endpoint_indices, selector_indices);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) return CRN2_EHIP;
    return CRN2_OK;
}
