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
//   3. hc_ep_refine_kernel, one wave per cluster: the refiner -- double
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
    // one wave per cluster, as hc_alpha_refine_kernel: the serial texel-order
    // sums run on every lane, the texels arrive 64 at a time
    const uint32_t c = blockIdx.x, lane = threadIdx.x;
    if (c >= nclusters) return;
    const uint32_t o = ofs[c], n = ofs[c + 1] - o;
    if (lane == 0) {
        endpoints[c] = 0;
        for (int i = 0; i < 4; i++) color_values[4 * c + i] = 0;
    }
    if (!n) return;
    const ClusterRes r = res[c];
    if (r.flags & 256u) {
        if (lane == 0) atomicAdd(not_encoded, 1u);
        return;
    }
    int cv[4][3];
    cluster_colors(r, cv, false);
    if (lane == 0)
        for (int i = 0; i < 4; i++)
            color_values[4 * c + i] =
                (uint32_t)cv[i][0] | (uint32_t)cv[i][1] << 8 | (uint32_t)cv[i][2] << 16 | 0xFF000000u;
    // dxt_endpoint_refiner::refine (crn_dxt_endpoint_refiner.cpp:36-144)
    const uint32_t* px = px_all + o;
    const uint8_t* sel = sel_all + o;
    double alpha2_sum = 0.0f, beta2_sum = 0.0f, alphabeta_sum = 0.0f;
    double ax[3] = {0, 0, 0}, bx[3] = {0, 0, 0}, first[3] = {0, 0, 0};
    uint64_t hist[4] = {0, 0, 0, 0}, D2[4][3], DD[4][3];
    for (int s = 0; s < 4; s++)
        for (int k = 0; k < 3; k++) D2[s][k] = DD[s][k] = 0;
    for (uint32_t i0 = 0; i0 < n; i0 += 64) {
      const uint32_t mine = i0 + lane < n ? (px[i0 + lane] & 0x00FFFFFFu) | ((uint32_t)(sel[i0 + lane] & 3u) << 24) : 0u;
      const uint32_t m = min(64u, n - i0);
      for (uint32_t t = 0; t < m; t++) {
        const uint32_t p = (uint32_t)__builtin_amdgcn_readlane((int)mine, (int)t);
        const uint32_t i = i0 + t;
        const uint32_t s = p >> 24;
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
    if (lane == 0) endpoints[c] = err < r.error ? (lo | hi << 16) : r.lo_hi;
}

// ------------------------------------------------------------ alpha (H.8)
// dxt_hc::determine_alpha_endpoint_codebook_task (crn_dxt_hc.cpp:1014-1130,
// without ETC colour blocks) for one alpha channel:
//   hc_alpha_search_kernel, one wave per cluster: the uber dxt5_endpoint_
//     optimizer (crn_dxt5a.cpp:40-262, 8-value blocks only) over the cluster's
//     values -- distinct values in first-appearance order with their counts
//     (LDS tables over the 256 values), every distinct pair in that order,
//     then the +-16 box around the current best (centre re-read at every
//     step); each candidate is scored by all lanes at once (first minimum per
//     value, wave sum) and kept on a strict improvement, a zero error ends the
//     search; canonical order with the eight-value invert table;
//   hc_alpha_blocks_kernel, one thread per block: values in linear order,
//     encoding weights from the endpoint spread, 3-bit nearest-value selectors;
//   hc_alpha_refine_kernel, one wave per cluster: the DXT5 refiner
//     (least squares on the alpha values in texel order, then the canonical
//     23 x 23 box scored through per-selector histograms).
struct AlphaRes {
    uint64_t error;
    uint32_t first, second, reordered, pad;
};

__constant__ uint8_t k5_to_linear[8] = {0, 7, 1, 2, 3, 4, 5, 6};
__constant__ uint8_t k5_from_linear[8] = {0, 2, 3, 4, 5, 6, 7, 1};
__constant__ uint8_t k5_eight_invert[8] = {1, 0, 7, 6, 5, 4, 3, 2};

__device__ inline void values8(uint32_t l, uint32_t h, uint32_t* v)
{
    v[0] = l;
    v[1] = h;
    for (uint32_t j = 2; j < 8; j++) v[j] = (l * (8 - j) + h * (j - 1)) / 7;
}

// dxt5_block::get_block_values (crn_dxt.cpp:369-455)
__device__ inline void alpha_values_of(uint32_t l, uint32_t h, uint32_t* v)
{
    if (l > h) {
        values8(l, h, v);
    } else {
        v[0] = l;
        v[1] = h;
        for (uint32_t j = 2; j < 6; j++) v[j] = (l * (6 - j) + h * (j - 1)) / 5;
        v[6] = 0;
        v[7] = 255;
    }
}

struct AlphaLds {
    uint32_t count[256];
    uint32_t first[256];
    uint32_t uval[256];
    uint32_t uw[256];
    uint32_t rank[256];
    uint8_t trial[256];
    uint8_t best[256];
};

__device__ inline uint64_t wave_sum64(uint64_t v)
{
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, off), hi = (uint32_t)__shfl_xor((int)(v >> 32), off);
        v += (uint64_t)lo | ((uint64_t)hi << 32);
    }
    return v;
}

__global__ void __launch_bounds__(64) hc_alpha_search_kernel(const uint8_t* __restrict__ vals_all,
                                                             const uint32_t* __restrict__ ofs, uint32_t nclusters,
                                                             AlphaRes* __restrict__ res, uint8_t* __restrict__ sel_all)
{
    __shared__ AlphaLds L;
    const uint32_t c = blockIdx.x, lane = threadIdx.x;
    const uint32_t o = ofs[c], n = ofs[c + 1] - o;
    if (!n) return;
    const uint8_t* vals = vals_all + o;
    for (uint32_t v = lane; v < 256; v += 64) {
        L.count[v] = 0;
        L.first[v] = 0xFFFFFFFFu;
    }
    __syncthreads();
    for (uint32_t i = lane; i < n; i += 64) {
        atomicAdd(&L.count[vals[i]], 1u);
        atomicMin(&L.first[vals[i]], i);
    }
    __syncthreads();
    // rank of each present value by first appearance
    for (uint32_t v = lane; v < 256; v += 64) {
        if (!L.count[v]) continue;
        const uint32_t f = L.first[v];
        uint32_t r = 0;
        for (uint32_t u = 0; u < 256; u++) r += (L.count[u] && L.first[u] < f) ? 1u : 0u;
        L.rank[v] = r;
        L.uval[r] = v;
        L.uw[r] = L.count[v];
    }
    __syncthreads();
    uint32_t nu = 0;
    for (uint32_t v = 0; v < 256; v++) nu += L.count[v] ? 1u : 0u;
    AlphaRes out;
    out.pad = 0;
    if (nu == 1) {
        out.error = 0;
        out.first = out.second = L.uval[0];
        out.reordered = 0;
        if (lane == 0) res[c] = out;
        for (uint32_t i = lane; i < n; i += 64) sel_all[o + i] = 0;
        return;
    }
    uint64_t best_err = ~0ull;
    uint32_t best_first = 0, best_second = 0;
    bool done = false;
    // evaluate_solution (crn_dxt5a.cpp:198-262) with the 8-value palette only
    auto evaluate = [&](uint32_t lo, uint32_t hi) {
        uint32_t pal[8];
        values8(lo, hi, pal);
        uint64_t part = 0;
        for (uint32_t i = lane; i < nu; i += 64) {
            const int val = (int)L.uval[i], w = (int)L.uw[i];
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
            L.trial[i] = (uint8_t)bsel;
            part += best;
        }
        const uint64_t err = wave_sum64(part);
        if (err < best_err) {
            best_err = err;
            best_first = lo;
            best_second = hi;
            for (uint32_t i = lane; i < nu; i += 64) L.best[i] = L.trial[i];
            if (!err) done = true;
        }
    };
    for (uint32_t i = 0; i + 1 < nu && !done; i++)
        for (uint32_t j = i + 1; j < nu && !done; j++) evaluate(L.uval[i], L.uval[j]);
    if (best_err && !done) {
        for (int dl = -16; dl <= 16 && !done; dl++) {
            const int l = (int)best_first + dl;
            if (l < 0) continue;
            if (l > 255) break;
            for (int dh = -16; dh <= 16 && !done; dh++) {
                const int h = (int)best_second + dh;
                if (h < 0) continue;
                if (h > 255) break;
                evaluate((uint32_t)l, (uint32_t)h);
            }
        }
    }
    // canonical order (crn_dxt5a.cpp:151-184), 8-value blocks
    uint32_t f = best_first, sc = best_second, reordered = 0;
    bool zero = f == sc, invert = false;
    if (!zero && f <= sc) {
        const uint32_t t = f;
        f = sc;
        sc = t;
        reordered = 1;
        invert = true;
    }
    __syncthreads();
    for (uint32_t i = lane; i < n; i += 64) {
        const uint32_t u = L.rank[vals[i]];
        uint32_t s = zero ? 0u : L.best[u];
        if (invert) s = k5_eight_invert[s];
        sel_all[o + i] = (uint8_t)s;
    }
    out.error = best_err;
    out.first = f;
    out.second = sc;
    out.reordered = reordered;
    if (lane == 0) res[c] = out;
}

__global__ void __launch_bounds__(256) hc_alpha_blocks_kernel(const AlphaRes* __restrict__ res,
                                                              const uint32_t* __restrict__ ofs,
                                                              const uint32_t* __restrict__ block_cluster,
                                                              uint32_t nblocks, const uint8_t* __restrict__ block_vals,
                                                              const uint8_t* __restrict__ encodings,
                                                              uint64_t* __restrict__ block_selectors)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblocks) return;
    const uint32_t c = block_cluster[b];
    if (c == kNone || ofs[c + 1] == ofs[c]) return;
    const AlphaRes r = res[c];
    uint32_t bv[8], av[8];
    alpha_values_of(r.first, r.second, bv);
    for (int i = 0; i < 8; i++) av[i] = bv[k5_from_linear[i]];
    const int delta = (int)r.first - (int)r.second;
    const uint32_t dd = (uint32_t)(delta * delta >> 3);
    const uint32_t ew = dd < 1 ? 1u : (dd > 2048 ? 2048u : dd);
    const uint32_t weight = (uint32_t)((float)ew * (1.15f + (1.0f - 1.15f) * ((float)encodings[b] / 7.0f)));
    uint64_t selector = 0;
    const uint8_t* v = block_vals + (size_t)b * 16;
    for (int p = 0; p < 16; p++) {
        uint32_t best = 0xFFFFFFFFu, sb = 0;
        for (int t = 0; t < 8; t++) {
            const uint32_t s = r.reordered ? 7u - k5_to_linear[t] : k5_to_linear[t];
            const int d = (int)v[p] - (int)av[s];
            const uint32_t e = (uint32_t)(d >= 0 ? d : -d);
            if (e < best) {
                sb = s;
                best = e;
            }
        }
        selector = selector << 3 | sb;
    }
    block_selectors[b] = selector << 16 | weight;
}

__global__ void __launch_bounds__(64) hc_alpha_refine_kernel(const uint8_t* __restrict__ vals_all,
                                                             const uint32_t* __restrict__ ofs, uint32_t nclusters,
                                                             const AlphaRes* __restrict__ res,
                                                             const uint8_t* __restrict__ sel_all,
                                                             uint32_t* __restrict__ endpoints,
                                                             uint32_t* __restrict__ alpha_values,
                                                             uint32_t* __restrict__ refined_values)
{
    // one wave per cluster: every lane runs the same (serial, texel-order)
    // computation; the texels arrive 64 at a time by one coalesced load and
    // are broadcast lane by lane, so no step waits on a lone global load
    const uint32_t c = blockIdx.x, lane = threadIdx.x;
    if (c >= nclusters) return;
    const uint32_t o = ofs[c], n = ofs[c + 1] - o;
    if (lane == 0) {
        endpoints[c] = 0;
        alpha_values[2 * c] = alpha_values[2 * c + 1] = 0;
        refined_values[2 * c] = refined_values[2 * c + 1] = 0;
    }
    if (!n) return;
    const AlphaRes r = res[c];
    uint32_t bv[8], av[8];
    alpha_values_of(r.first, r.second, bv);
    for (int i = 0; i < 8; i++) av[i] = bv[k5_from_linear[i]];
    if (lane == 0) {
        alpha_values[2 * c] = av[0] | av[1] << 8 | av[2] << 16 | av[3] << 24;
        alpha_values[2 * c + 1] = av[4] | av[5] << 8 | av[6] << 16 | av[7] << 24;
    }
    // dxt_endpoint_refiner::refine, DXT5 selectors (crn_dxt_endpoint_refiner.cpp:36-144)
    const uint8_t* vals = vals_all + o;
    const uint8_t* sel = sel_all + o;
    // Every product the sums take depends only on the selector s and the
    // value a: alpha = 1 - k(s), beta = k(s), x = a / 255.  The rounded
    // products are tabulated once (LDS), so each texel adds five table
    // entries to the five running sums in texel order -- the same sums, bit
    // for bit, as multiplying and adding per texel.
    __shared__ double t_ax[8][256], t_bx[8][256], t_ab[3][8];
    for (uint32_t e = lane; e < 8 * 256; e += 64) {
        const uint32_t s = e >> 8, a = e & 255u;
        const double k = (double)((float)k5_to_linear[s] * 1.0f / 7.0f);
        const double alpha = 1.0f - k, beta = k, x = (double)((float)a / 255.0f);
        t_ax[s][a] = alpha * x;
        t_bx[s][a] = beta * x;
        if (a == 0) {
            t_ab[0][s] = alpha * alpha;
            t_ab[1][s] = beta * beta;
            t_ab[2][s] = alpha * beta;
        }
    }
    __syncthreads();
    double alpha2_sum = 0.0f, beta2_sum = 0.0f, alphabeta_sum = 0.0f, ax = 0, bx = 0;
    const double first = (double)((float)vals[0] / 255.0f);
    uint64_t hist[8], D2[8], DD[8];
    for (int s = 0; s < 8; s++) hist[s] = D2[s] = DD[s] = 0;
    for (uint32_t i0 = 0; i0 < n; i0 += 64) {
      const uint32_t mine = i0 + lane < n ? ((uint32_t)sel[i0 + lane] & 7u) | ((uint32_t)vals[i0 + lane] << 8) : 0u;
      const uint32_t m = min(64u, n - i0);
      for (uint32_t t = 0; t < m; t++) {
        const uint32_t sv = (uint32_t)__builtin_amdgcn_readlane((int)mine, (int)t);
        const uint32_t s = sv & 7u, a = sv >> 8;
        alpha2_sum += t_ab[0][s];
        beta2_sum += t_ab[1][s];
        alphabeta_sum += t_ab[2][s];
        ax += t_ax[s][a];
        bx += t_bx[s][a];
        hist[s]++;
        D2[s] += (uint64_t)(a * 2);
        DD[s] += (uint64_t)(a * a);
      }
    }
    double a, b;
    if (beta2_sum == 0.0f) {
        a = ax / alpha2_sum;
        b = 0.0;
    } else if (alpha2_sum == 0.0f) {
        a = 0.0;
        b = bx / beta2_sum;
    } else {
        const double factor = alpha2_sum * beta2_sum - alphabeta_sum * alphabeta_sum;
        if (factor != 0.0f) {
            a = (ax * beta2_sum - bx * alphabeta_sum) / factor;
            b = (bx * alpha2_sum - ax * alphabeta_sum) / factor;
        } else {
            a = b = first;
        }
    }
    float lf = (float)a, hf = (float)b;
    lf = lf < 0.0f ? 0.0f : (lf > 1.0f ? 1.0f : lf);
    hf = hf < 0.0f ? 0.0f : (hf > 1.0f ? 1.0f : hf);
    // optimize_dxt5 (crn_dxt_endpoint_refiner.cpp:146-201)
    const uint32_t L0 = (uint32_t)clamp_int((int)(lf * 256.0f), 0, 255);
    const uint32_t H0 = (uint32_t)clamp_int((int)(hf * 256.0f), 0, 255);
    auto canon = [](uint32_t L, uint32_t H) -> uint32_t {
        return L == H ? (H ? ((H - 1) << 8 | L) : 1u) : (L > H ? (H << 8 | L) : (L << 8 | H));
    };
    uint64_t rerr = ~0ull;
    uint32_t rlo = 0, rhi = 0;
    auto score = [&](uint32_t sol) -> bool {   // true: error reached zero
        const uint32_t L = sol & 0xFFu, H = sol >> 8;
        uint32_t v[8];
        values8(L, H, v);
        uint64_t error = 0;
        for (int s = 0; s < 8; s++) error += hist[s] * v[s] * v[s] - D2[s] * v[s] + DD[s];
        if (error < rerr) {
            rlo = L;
            rhi = H;
            rerr = error;
            if (!error) return true;
        }
        return false;
    };
    bool stop = score(canon(L0, H0));
    const uint32_t minL = L0 <= 11 ? 0 : L0 - 11, maxL = L0 >= 244 ? 255 : L0 + 11;
    const uint32_t minH = H0 <= 11 ? 0 : H0 - 11, maxH = H0 >= 244 ? 255 : H0 + 11;
    for (uint32_t L = minL; L <= maxL && !stop; L++)
        for (uint32_t H = minH; H <= maxH && !stop; H++)
            if ((maxH < L || L <= H || H < minL) && (L != L0 || H != H0) && (L != H0 || H != L0)) stop = score(canon(L, H));
    const bool refined = rerr < r.error;
    uint32_t f = r.first, sc = r.second;
    if (refined) {
        f = rlo;
        sc = rhi;
        alpha_values_of(f, sc, bv);
        for (int i = 0; i < 8; i++) av[i] = bv[k5_from_linear[i]];
    }
    if (lane == 0) {
        refined_values[2 * c] = av[0] | av[1] << 8 | av[2] << 16 | av[3] << 24;
        refined_values[2 * c + 1] = av[4] | av[5] << 8 | av[6] << 16 | av[7] << 24;
        endpoints[c] = f | sc << 8 | (refined ? 1u : 0u) << 16;
    }
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
        hipLaunchKernelGGL(hc_ep_refine_kernel, dim3(nclusters), dim3(64), 0, s,
                           static_cast<const uint32_t*>(cluster_px), cluster_ofs, nclusters, res, sel, perceptual != 0,
                           endpoints, static_cast<uint32_t*>(color_values), not_encoded);
        if (hipGetLastError() != hipSuccess) rc = CRN2_EHIP;
    }
    (void)hipFreeAsync(sel, s);
    (void)hipFreeAsync(res, s);
    return rc;
}

extern "C" int crn2_hc_alpha_endpoints_device(const void* cluster_vals, const uint32_t* cluster_ofs, uint32_t nclusters,
                                              uint32_t total_vals, const uint32_t* block_cluster, uint32_t nblocks,
                                              const void* block_vals, const uint8_t* block_encodings,
                                              uint32_t* endpoints, void* alpha_values, void* refined_values,
                                              uint64_t* block_selectors, void* stream)
{
    if (!nclusters) return CRN2_OK;
    if (!cluster_vals || !cluster_ofs || !endpoints || !alpha_values || !refined_values ||
        (nblocks && (!block_cluster || !block_vals || !block_encodings || !block_selectors)))
        return CRN2_EINVAL;
    hipStream_t s = static_cast<hipStream_t>(stream);
    AlphaRes* res = nullptr;
    uint8_t* sel = nullptr;
    if (hipMallocAsync(reinterpret_cast<void**>(&res), (size_t)nclusters * sizeof(AlphaRes), s) != hipSuccess)
        return CRN2_ENOMEM;
    if (hipMallocAsync(reinterpret_cast<void**>(&sel), (size_t)total_vals + 1, s) != hipSuccess) {
        (void)hipFreeAsync(res, s);
        return CRN2_ENOMEM;
    }
    int rc = CRN2_OK;
    hipLaunchKernelGGL(hc_alpha_search_kernel, dim3(nclusters), dim3(64), 0, s, static_cast<const uint8_t*>(cluster_vals),
                       cluster_ofs, nclusters, res, sel);
    if (hipGetLastError() != hipSuccess) rc = CRN2_EHIP;
    if (rc == CRN2_OK && nblocks) {
        hipLaunchKernelGGL(hc_alpha_blocks_kernel, dim3(grid1(nblocks, 256)), dim3(256), 0, s, res, cluster_ofs,
                           block_cluster, nblocks, static_cast<const uint8_t*>(block_vals), block_encodings,
                           block_selectors);
        if (hipGetLastError() != hipSuccess) rc = CRN2_EHIP;
    }
    if (rc == CRN2_OK) {
        hipLaunchKernelGGL(hc_alpha_refine_kernel, dim3(nclusters), dim3(64), 0, s,
                           static_cast<const uint8_t*>(cluster_vals), cluster_ofs, nclusters, res, sel, endpoints,
                           static_cast<uint32_t*>(alpha_values), static_cast<uint32_t*>(refined_values));
        if (hipGetLastError() != hipSuccess) rc = CRN2_EHIP;
    }
    (void)hipFreeAsync(sel, s);
    (void)hipFreeAsync(res, s);
    return rc;
}
