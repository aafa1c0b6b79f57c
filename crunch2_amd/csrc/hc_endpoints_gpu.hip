// hc_endpoints_gpu.hip -- dxt_hc's colour endpoint codebook on gfx950.
//
// Reference: dxt_hc::determine_color_endpoint_codebook_task
// (crnlib/crn_dxt_hc.cpp:663-754) and dxt_endpoint_refiner::refine /
// optimize_dxt1 (crnlib/crn_dxt_endpoint_refiner.cpp:36-144, 203-302).
// Three launches on the caller's stream:
//   1. crn2_dxt1_clusters_device: the uber endpoint search per cluster, one
//      wave per cluster (dxt1c_gpu.hip), with every texel's selector;
//   2. hc_ep_blocks_kernel, one thread per block: the cluster's four colours
//      (rounded middle colours when the search chose alternate rounding), the
//      block's weight and its 16 nearest-colour selectors;
//   3. hc_ep_refine_kernel, one thread per cluster: the refiner -- double
//      least-squares sums in texel order (so that they round as the
//      reference's do), then up to 8 rounds over the +-1 neighbours of either
//      endpoint scored through per-selector histograms with the selectors
//      fixed; its endpoints replace the search's when their error is lower.
// All float and double expressions keep the reference's types and operation
// order; the library is built with -ffp-contract=off and IEEE division.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "crn2_internal.h"

extern "C" int crn2_dxt1_clusters_device(const void* rgba, const uint32_t* offsets, uint32_t nclusters, int flags,
                                         void* out, uint8_t* selectors, void* stream);

namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;

struct ClusterRes {          // crn2_dxt1_clusters_device output, 24 bytes per cluster
    uint64_t error;
    uint32_t lo_hi;
    uint32_t nunique;
    uint32_t flags;          // bit 0 reordered, bit 2 alternate rounding, bit 8 not encoded
    uint32_t pad;
};

__constant__ uint8_t k_to_linear[4] = {0, 3, 1, 2};
__constant__ uint8_t k_from_linear[4] = {0, 2, 3, 1};

__device__ inline void unpack565(uint32_t p, int* o)
{
    const uint32_t b = p & 31u, g = (p >> 5) & 63u, r = (p >> 11) & 31u;
    o[0] = (int)((r << 3) | (r >> 2));
    o[1] = (int)((g << 2) | (g >> 4));
    o[2] = (int)((b << 3) | (b >> 2));
}

// dxt1_block::get_block_colors4 (crn_dxt.cpp:247-260)
__device__ inline void colors4(uint32_t c0, uint32_t c1, int (*o)[3])
{
    unpack565(c0, o[0]);
    unpack565(c1, o[1]);
    for (int k = 0; k < 3; k++) {
        o[2][k] = (o[0][k] * 2 + o[1][k]) / 3;
        o[3][k] = (o[1][k] * 2 + o[0][k]) / 3;
    }
}

__device__ inline uint32_t color_distance(bool perceptual, const int* a, const int* b)
{
    const int dr = a[0] - b[0], dg = a[1] - b[1], db = a[2] - b[2];
    return perceptual ? (uint32_t)(8 * dr * dr + 25 * dg * dg + db * db) : (uint32_t)(dr * dr + dg * dg + db * db);
}

// The cluster's colours in linear order (crn_dxt_hc.cpp:692-702).
__device__ inline void cluster_colors(const ClusterRes& r, int (*cv)[3], bool rounded)
{
    int bv[4][3];
    colors4(r.lo_hi & 0xFFFFu, r.lo_hi >> 16, bv);
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 3; k++) cv[i][k] = bv[k_from_linear[i]][k];
    if (rounded && (r.flags & 4u))
        for (int k = 0; k < 3; k++) {
            cv[1][k] = ((cv[0][k] << 1) + cv[3][k] + 1) / 3;
            cv[2][k] = ((cv[3][k] << 1) + cv[0][k] + 1) / 3;
        }
}

__global__ void __launch_bounds__(256) hc_ep_blocks_kernel(const ClusterRes* __restrict__ res,
                                                           const uint32_t* __restrict__ ofs,
                                                           const uint32_t* __restrict__ block_cluster, uint32_t nblocks,
                                                           const uint8_t* __restrict__ blocks,
                                                           const float* __restrict__ weights,
                                                           const uint8_t* __restrict__ encodings, bool perceptual,
                                                           uint64_t* __restrict__ block_selectors)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblocks) return;
    const uint32_t c = block_cluster[b];
    if (c == kNone || ofs[c + 1] == ofs[c]) return;
    const ClusterRes r = res[c];
    if (r.flags & 256u) return;
    int cv[4][3];
    cluster_colors(r, cv, true);
    const uint32_t endpoint_weight = color_distance(perceptual, cv[0], cv[3]) / 2000;
    const float enc_w = 1.15f + (1.0f - 1.15f) * ((float)encodings[b] / 7.0f);
    const uint32_t w0 = (uint32_t)((float)endpoint_weight * weights[b]);
    const uint32_t wc = w0 < 1 ? 1u : (w0 > 2048 ? 2048u : w0);
    const uint32_t weight = (uint32_t)((float)wc * enc_w);
    const bool reordered = (r.flags & 1u) != 0;
    uint32_t selector = 0;
    const uint8_t* px = blocks + (size_t)b * 64;
    for (int p = 0; p < 16; p++) {
        const int x[3] = {px[4 * p], px[4 * p + 1], px[4 * p + 2]};
        uint32_t best = 0xFFFFFFFFu, sb = 0;
        for (int t = 0; t < 4; t++) {
            const uint32_t s = reordered ? 3u - k_to_linear[t] : k_to_linear[t];
            const uint32_t e = color_distance(perceptual, x, cv[s]);
            if (e < best) {
                sb = s;
                best = e;
            }
        }
        selector = selector << 2 | sb;
    }
    block_selectors[b] = (uint64_t)selector << 32 | weight;
}

__device__ inline int clamp_int(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// optimize_dxt1 (crn_dxt_endpoint_refiner.cpp:203-302).
__device__ void optimize_dxt1(const uint64_t* hist, const uint64_t (*D2)[3], const uint64_t (*DD)[3], bool perceptual,
                              const float* l, const float* h, uint32_t& out_lo, uint32_t& out_hi, uint64_t& out_err)
{
    uint32_t L0 = (uint32_t)(clamp_int((int)(l[0] * 32.0f), 0, 31) << 11 | clamp_int((int)(l[1] * 64.0f), 0, 63) << 5 |
                             clamp_int((int)(l[2] * 32.0f), 0, 31));
    uint32_t H0 = (uint32_t)(clamp_int((int)(h[0] * 32.0f), 0, 31) << 11 | clamp_int((int)(h[1] * 64.0f), 0, 63) << 5 |
                             clamp_int((int)(h[2] * 32.0f), 0, 31));
    const bool preserveL = hist[0] + hist[2] > hist[1] + hist[3];
    uint32_t sol[52];
    out_err = ~0ull;
    bool improved = true;
    for (uint32_t iterations = 8; improved && iterations; iterations--) {
        improved = false;
        uint32_t cnt = 0;
        for (int pass = 0; pass < 2; pass++) {
            const uint32_t E = pass ? H0 : L0, O = pass ? L0 : H0;
            const uint32_t b0 = E & 31u, g0 = (E >> 5) & 63u, r0 = (E >> 11) & 31u;
            for (uint32_t b = b0 ? b0 - 1 : b0; b <= b0 + 1 && b <= 31; b++)
                for (uint32_t g = g0 ? g0 - 1 : g0; g <= g0 + 1 && g <= 63; g++)
                    for (uint32_t r = r0 ? r0 - 1 : r0; r <= r0 + 1 && r <= 31; r++) {
                        const uint32_t X = r << 11 | g << 5 | b;
                        if (X != E) sol[cnt++] = X > O ? X | O << 16 : O | X << 16;
                    }
        }
        for (uint32_t i = 1; i < cnt; i++) {
            const uint32_t v = sol[i];
            uint32_t j = i;
            while (j && sol[j - 1] > v) {
                sol[j] = sol[j - 1];
                j--;
            }
            sol[j] = v;
        }
        for (uint32_t i = 0; i < cnt; i++) {
            if (i && sol[i] == sol[i - 1]) continue;
            uint32_t L = sol[i] & 0xFFFFu, H = sol[i] >> 16;
            if (L == H) {
                uint32_t addL, subH;
                if (!preserveL)
                    addL = (~L & 0x1Fu) ? 0x1u : (~L & 0xF800u) ? 0x800u : (~L & 0x7E0u) ? 0x20u : 0u;
                else
                    addL = !L ? 0x1u : 0u;
                L = (L + addL) & 0xFFFFu;
                if (preserveL)
                    subH = (H & 0x1Fu) ? 0x1u : (H & 0xF800u) ? 0x800u : (H & 0x7E0u) ? 0x20u : 0u;
                else
                    subH = H == 0xFFFFu ? 0x1u : 0u;
                H = (H - subH) & 0xFFFFu;
            }
            int bc[4][3];
            colors4(L, H, bc);
            uint64_t error = 0;
            for (int s = 0; s < 4; s++) {
                uint64_t d[3];
                for (int c = 0; c < 3; c++) {
                    const uint64_t v = (uint64_t)bc[s][c];
                    d[c] = hist[s] * v * v - D2[s][c] * v + DD[s][c];
                }
                error += perceptual ? d[0] * 8 + d[1] * 25 + d[2] : d[0] + d[1] + d[2];
            }
            if (error < out_err) {
                out_lo = L0 = L;
                out_hi = H0 = H;
                out_err = error;
                if (!error) return;
                improved = true;
            }
        }
    }
}

__global__ void __launch_bounds__(64) hc_ep_refine_kernel(const uint32_t* __restrict__ px_all,
                                                          const uint32_t* __restrict__ ofs, uint32_t nclusters,
                                                          const ClusterRes* __restrict__ res,
                                                          const uint8_t* __restrict__ sel_all, bool perceptual,
                                                          uint32_t* __restrict__ endpoints,
                                                          uint32_t* __restrict__ color_values,
                                                          uint32_t* __restrict__ not_encoded)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nclusters) return;
    const uint32_t o = ofs[c], n = ofs[c + 1] - o;
    endpoints[c] = 0;
    for (int i = 0; i < 4; i++) color_values[4 * c + i] = 0;
    if (!n) return;
    const ClusterRes r = res[c];
    if (r.flags & 256u) {
        atomicAdd(not_encoded, 1u);
        return;
    }
    int cv[4][3];
    cluster_colors(r, cv, false);
    for (int i = 0; i < 4; i++)
        color_values[4 * c + i] = (uint32_t)cv[i][0] | (uint32_t)cv[i][1] << 8 | (uint32_t)cv[i][2] << 16 | 0xFF000000u;
    // dxt_endpoint_refiner::refine (crn_dxt_endpoint_refiner.cpp:36-144)
    const uint32_t* px = px_all + o;
    const uint8_t* sel = sel_all + o;
    double alpha2_sum = 0.0f, beta2_sum = 0.0f, alphabeta_sum = 0.0f;
    double ax[3] = {0, 0, 0}, bx[3] = {0, 0, 0}, first[3] = {0, 0, 0};
    uint64_t hist[4] = {0, 0, 0, 0}, D2[4][3], DD[4][3];
    for (int s = 0; s < 4; s++)
        for (int k = 0; k < 3; k++) D2[s][k] = DD[s][k] = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t p = px[i];
        const uint32_t s = sel[i] & 3u;
        const int pc[3] = {(int)(p & 0xFFu), (int)((p >> 8) & 0xFFu), (int)((p >> 16) & 0xFFu)};
        const double k = (double)((float)k_to_linear[s] * 1.0f / 3.0f);
        const double alpha = 1.0f - k;
        const double beta = k;
        double x[3];
        for (int q = 0; q < 3; q++) x[q] = (double)((float)pc[q] * 1.0f / 255.0f);
        if (!i)
            for (int q = 0; q < 3; q++) first[q] = x[q];
        alpha2_sum += alpha * alpha;
        beta2_sum += beta * beta;
        alphabeta_sum += alpha * beta;
        for (int q = 0; q < 3; q++) {
            ax[q] += alpha * x[q];
            bx[q] += beta * x[q];
        }
        hist[s]++;
        for (int q = 0; q < 3; q++) {
            D2[s][q] += (uint64_t)(pc[q] * 2);
            DD[s][q] += (uint64_t)(pc[q] * pc[q]);
        }
    }
    double a[3], b[3];
    if (beta2_sum == 0.0f) {
        for (int q = 0; q < 3; q++) {
            a[q] = ax[q] / alpha2_sum;
            b[q] = 0.0;
        }
    } else if (alpha2_sum == 0.0f) {
        for (int q = 0; q < 3; q++) {
            a[q] = 0.0;
            b[q] = bx[q] / beta2_sum;
        }
    } else {
        const double factor = alpha2_sum * beta2_sum - alphabeta_sum * alphabeta_sum;
        if (factor != 0.0f) {
            for (int q = 0; q < 3; q++) {
                a[q] = (ax[q] * beta2_sum - bx[q] * alphabeta_sum) / factor;
                b[q] = (bx[q] * alpha2_sum - ax[q] * alphabeta_sum) / factor;
            }
        } else {
            for (int q = 0; q < 3; q++) a[q] = b[q] = first[q];
        }
    }
    float l[3], h[3];
    for (int q = 0; q < 3; q++) {
        l[q] = (float)a[q];
        h[q] = (float)b[q];
        l[q] = l[q] < 0.0f ? 0.0f : (l[q] > 1.0f ? 1.0f : l[q]);
        h[q] = h[q] < 0.0f ? 0.0f : (h[q] > 1.0f ? 1.0f : h[q]);
    }
    uint32_t lo = 0, hi = 0;
    uint64_t err;
    optimize_dxt1(hist, D2, DD, perceptual, l, h, lo, hi, err);
    endpoints[c] = err < r.error ? (lo | hi << 16) : r.lo_hi;
}

inline uint32_t grid1(uint64_t n, uint32_t tb) { return (uint32_t)((n + tb - 1) / tb); }

}  // namespace

extern "C" int crn2_hc_color_endpoints_device(const void* cluster_px, const uint32_t* cluster_ofs, uint32_t nclusters,
                                              uint32_t total_px, const uint32_t* block_cluster, uint32_t nblocks,
                                              const void* blocks, const float* block_weights,
                                              const uint8_t* block_encodings, int perceptual, uint32_t* endpoints,
                                              void* color_values, uint64_t* block_selectors, uint32_t* not_encoded,
                                              void* stream)
{
    if (!nclusters) return CRN2_OK;
    if (!cluster_px || !cluster_ofs || !endpoints || !color_values || !not_encoded || (perceptual & ~1) ||
        (nblocks && (!block_cluster || !blocks || !block_weights || !block_encodings || !block_selectors)))
        return CRN2_EINVAL;
    hipStream_t s = static_cast<hipStream_t>(stream);
    ClusterRes* res = nullptr;
    uint8_t* sel = nullptr;
    if (hipMallocAsync(reinterpret_cast<void**>(&res), (size_t)nclusters * sizeof(ClusterRes), s) != hipSuccess)
        return CRN2_ENOMEM;
    if (hipMallocAsync(reinterpret_cast<void**>(&sel), (size_t)total_px + 1, s) != hipSuccess) {
        (void)hipFreeAsync(res, s);
        return CRN2_ENOMEM;
    }
    int rc = CRN2_OK;
    if (hipMemsetAsync(not_encoded, 0, 4, s) != hipSuccess) rc = CRN2_EHIP;
    if (rc == CRN2_OK) rc = crn2_dxt1_clusters_device(cluster_px, cluster_ofs, nclusters, perceptual, res, sel, stream);
    if (rc == CRN2_OK && nblocks) {
        hipLaunchKernelGGL(hc_ep_blocks_kernel, dim3(grid1(nblocks, 256)), dim3(256), 0, s, res, cluster_ofs,
                           block_cluster, nblocks, static_cast<const uint8_t*>(blocks), block_weights, block_encodings,
                           perceptual != 0, block_selectors);
        if (hipGetLastError() != hipSuccess) rc = CRN2_EHIP;
    }
    if (rc == CRN2_OK) {
        hipLaunchKernelGGL(hc_ep_refine_kernel, dim3(grid1(nclusters, 64)), dim3(64), 0, s,
                           static_cast<const uint32_t*>(cluster_px), cluster_ofs, nclusters, res, sel, perceptual != 0,
                           endpoints, static_cast<uint32_t*>(color_values), not_encoded);
        if (hipGetLastError() != hipSuccess) rc = CRN2_EHIP;
    }
    (void)hipFreeAsync(sel, s);
    (void)hipFreeAsync(res, s);
    return rc;
}
