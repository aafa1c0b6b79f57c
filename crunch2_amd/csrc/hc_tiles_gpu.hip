// hc_tiles_gpu.hip -- dxt_hc's adaptive tile choice (dxt_hc::determine_tiles_task,
// crnlib/crn_dxt_hc.cpp:386-561) on gfx950.
//
// Each 2x2-block chunk of a level offers nine tiles (the four blocks, the two
// block columns, the two halves, the whole 8x8) and eight ways of covering
// the chunk with them.  The reference scores every tile (colour: the error of
// dxt_fast::compress_color_block without refinement against the 4-colour
// palette, crn_dxt_fast.cpp:25-764; alpha: the error of
// dxt5_endpoint_optimizer at normal quality, crn_dxt5a.cpp:40-262), turns the
// eight coverings' summed errors into PSNR minus a per-level derating and
// keeps the best covering; each kept tile gets palettized endpoints
// (palettize_color / palettize_alpha :318-376, with split_vectors,
// crn_tree_clusterizer.h:573-753).
//
// Three kernels, all exact restatements with the reference's operation order
// (the library is built with -ffp-contract=off):
//   tile_error_kernel   one thread per (chunk, tile): colour errors;
//   tile_alpha_kernel   one wave per (chunk, tile, alpha channel): alpha errors;
//   tile_choose_kernel  one thread per chunk: PSNR, derating, best covering,
//                       block encodings, tile indices, kept tiles' sources;
//   tile_vec_kernel     one thread per kept tile: palettized endpoints.
// The chunk order and the tile slots (tile_offset walks 4 slots per chunk,
// backwards on every other row pair of a face) are laid out on the host.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include "crn2_internal.h"
#include "dxt_fast_dev.h"

namespace {

using namespace dxtf;

__constant__ uint8_t c_om5[256][2], c_om6[256][2];
__constant__ float c_u2f[256];

constexpr uint32_t kOffsets[9] = {0, 16, 32, 48, 0, 32, 64, 96, 64};
__constant__ uint8_t c_tiles[8][4] = {{8}, {6, 7}, {4, 5}, {6, 1, 3}, {7, 0, 2}, {4, 2, 3}, {5, 0, 1}, {0, 2, 1, 3}};
__constant__ uint8_t c_tile_map[8][2][2] = {{{0, 0}, {0, 0}}, {{0, 0}, {1, 1}}, {{0, 1}, {0, 1}}, {{0, 0}, {1, 2}},
                                            {{1, 2}, {0, 0}}, {{0, 1}, {0, 2}}, {{1, 0}, {2, 0}}, {{0, 1}, {2, 3}}};

struct Chunk {
    uint32_t b, width, level, toff;
};

struct TileArgs {
    const uint32_t* blocks;      // RGBA8 words, 16 per block
    const Chunk* chunks;
    uint32_t nchunks;
    int has_color, num_alpha, perceptual;
    uint32_t alpha_comp[2];
    float color_der[16][8], alpha_der[8], ratio;
    uint32_t* terr;              // nchunks x 3 x 9
    uint8_t* block_encodings;
    uint32_t* tile_indices;
    uint32_t* tile_src;          // per slot: b << 4 | t, 0xFFFFFFFF empty
    float* tile_vec;             // per slot: 10 floats
    uint32_t* slot_width;        // per slot: the level width of its chunk
    uint32_t num_blocks;
};

// Texel i of tile t of the chunk whose top-left block is b (level width w).
__device__ inline uint32_t tile_texel(const uint32_t* blocks, uint32_t b, uint32_t w, uint32_t t, uint32_t i)
{
    const uint32_t idx = kOffsets[t] + i;
    if (idx < 64) return blocks[(size_t)(b + ((idx & 16) ? w : 0) + ((idx & 32) ? 1 : 0)) * 16 + (idx & 15)];
    const uint32_t j = idx - 64;
    return blocks[(size_t)(b + ((j & 32) ? w : 0) + ((j & 4) ? 1 : 0)) * 16 + ((j >> 1) & 12) + (j & 3)];
}

// dxt_fast::compress_color_block (refine = false), then the error against the
// 4-colour palette of the result (crn_dxt_hc.cpp:457-470).
template <int N>
__device__ uint32_t tile_color_error(const uint32_t (&px)[N])
{
    uint8_t sel[N];
    uint32_t lo, hi;
    compress_color<N>(px, lo, hi, sel, c_om5, c_om6);
    int col[4][3];
    eval_colors(col, lo, hi);
    uint32_t err = 0;
#pragma unroll
    for (int i = 0; i < N; i++)
        for (int c = 0; c < 3; c++) {
            const uint32_t d = (uint32_t)ch(px[i], c) - (uint32_t)col[sel[i]][c];
            err += d * d;
        }
    return err;
}

template <int N>
__device__ void tile_errors(const TileArgs& a, const Chunk& c, uint32_t t, uint32_t* out)
{
    uint32_t px[N];
#pragma unroll
    for (int i = 0; i < N; i++) px[i] = tile_texel(a.blocks, c.b, c.width, t, (uint32_t)i);
    out[t] = tile_color_error<N>(px);
}

// Colour errors, one thread per (chunk, tile).
__global__ void __launch_bounds__(256) tile_error_kernel(TileArgs a)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.nchunks * 9) return;
    const uint32_t ci = i / 9, t = i % 9;
    const Chunk c = a.chunks[ci];
    uint32_t* out = a.terr + (size_t)ci * 27;
    if (t < 4) tile_errors<16>(a, c, t, out);
    else if (t < 8) tile_errors<32>(a, c, t, out);
    else tile_errors<64>(a, c, t, out);
}

// Alpha errors, one wave per (chunk, tile, alpha channel): the error of
// dxt5_endpoint_optimizer at normal quality, 8-value blocks only
// (crn_dxt5a.cpp:53-103, 198-262): the best error over every pair of distinct
// values, per value the nearest palette entry.  The reference walks the pairs
// in first-appearance order with early exits; its result is the smallest
// exact error over the unordered pairs of distinct values (the 8-value
// palette of (l, h) is the palette of (h, l), and the early exits only skip
// pairs that cannot lower the minimum), so here the distinct values come from
// a 256-bin histogram in LDS, the lanes split the pairs, and a wave minimum
// finishes.
constexpr uint32_t kAlphaWaves = 4;
__global__ void __launch_bounds__(64 * kAlphaWaves) tile_alpha_kernel(TileArgs a)
{
    __shared__ uint32_t hist[kAlphaWaves][256];
    __shared__ uint8_t uv[kAlphaWaves][64];
    __shared__ uint8_t uw[kAlphaWaves][64];
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const uint32_t job = blockIdx.x * kAlphaWaves + wave;
    const uint32_t njobs = a.nchunks * 9 * (uint32_t)a.num_alpha;
    if (job >= njobs) return;   // whole waves leave together
    const uint32_t k = job % (uint32_t)a.num_alpha, ct = job / (uint32_t)a.num_alpha;
    const uint32_t ci = ct / 9, t = ct % 9;
    const Chunk c = a.chunks[ci];
    const uint32_t n = t < 4 ? 16 : (t < 8 ? 32 : 64);
    uint32_t* h = hist[wave];
    for (uint32_t v = lane; v < 256; v += 64) h[v] = 0;
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    if (lane < n) atomicAdd(&h[(tile_texel(a.blocks, c.b, c.width, t, lane) >> (8 * a.alpha_comp[k])) & 255u], 1u);
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    // compact the occupied bins (any order: the result does not depend on it)
    uint32_t nu = 0;
    for (uint32_t v0 = 0; v0 < 256; v0 += 64) {
        const uint32_t cnt = h[v0 + lane];
        const uint64_t occ = __ballot(cnt != 0);
        if (cnt) {
            const uint32_t r = nu + (uint32_t)__popcll(occ & ((1ull << lane) - 1ull));
            uv[wave][r] = (uint8_t)(v0 + lane);
            uw[wave][r] = (uint8_t)cnt;
        }
        nu += (uint32_t)__popcll(occ);
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    uint64_t best = ~0ull;
    if (nu > 1) {
        const uint32_t npairs = nu * (nu - 1) / 2;
        for (uint32_t p = lane; p < npairs; p += 64) {
            // pair p in row-major order of the strict upper triangle
            uint32_t i = 0, left = p;
            while (left >= nu - 1 - i) {
                left -= nu - 1 - i;
                i++;
            }
            const uint32_t j = i + 1 + left;
            const uint32_t l = uv[wave][i], hv = uv[wave][j];
            uint32_t pal[8];
            pal[0] = l;
            pal[1] = hv;
            pal[2] = (l * 6 + hv) / 7;
            pal[3] = (l * 5 + hv * 2) / 7;
            pal[4] = (l * 4 + hv * 3) / 7;
            pal[5] = (l * 3 + hv * 4) / 7;
            pal[6] = (l * 2 + hv * 5) / 7;
            pal[7] = (l + hv * 6) / 7;
            uint64_t err = 0;
            for (uint32_t u = 0; u < nu && err < best; u++) {
                const int val = uv[wave][u], w = uw[wave][u];
                uint32_t b = 0xFFFFFFFFu;
#pragma unroll
                for (int s = 0; s < 8; s++) {
                    const int e = val - (int)pal[s];
                    b = min(b, (uint32_t)(e * e * w));
                }
                err += b;
            }
            best = min(best, err);
        }
    } else {
        best = 0;
    }
    for (int o = 32; o > 0; o >>= 1) best = min(best, (uint64_t)__shfl_xor(best, o, 64));
    if (lane == 0) a.terr[(size_t)ci * 27 + 9 * (1 + k) + t] = (uint32_t)best;
}

__global__ void __launch_bounds__(256) tile_choose_kernel(TileArgs a)
{
    const uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (ci >= a.nchunks) return;
    const Chunk c = a.chunks[ci];
    const uint32_t* te = a.terr + (size_t)ci * 27;
    float best_q = 0.0f;
    uint32_t best_e = 0;
    for (uint32_t e = 0; e < 8; e++) {
        uint32_t tot[3] = {0, 0, 0};
        for (uint32_t k = a.has_color ? 0 : 1; k < 1u + (uint32_t)a.num_alpha; k++)
            for (uint32_t t = 0, s = e + 1; s; s >>= 1, t++) tot[k] += te[9 * k + c_tiles[e][t]];
        float q = 0;
        if (a.has_color) {
            const double psnr = tot[0] ? log10(255.0f / sqrt(tot[0] / 192.0)) * 20.0f : 999999.0f;
            const double x = psnr - a.color_der[c.level][e];
            q = (float)(x > 0.0f ? x : 0.0f);
            if (a.num_alpha) q *= a.ratio;
        }
        for (int k = 0; k < a.num_alpha; k++) {
            const double psnr = tot[1 + k] ? log10(255.0f / sqrt(tot[1 + k] / 64.0)) * 20.0f : 999999.0f;
            const double x = psnr - a.alpha_der[e];
            q += (float)(x > 0.0f ? x : 0.0f);
        }
        if (q > best_q) {
            best_q = q;
            best_e = e;
        }
    }
    for (uint32_t ti = 0, s = best_e + 1; s; s >>= 1, ti++) {
        a.tile_src[c.toff | ti] = c.b << 4 | c_tiles[best_e][ti];
        a.slot_width[c.toff | ti] = c.width;
    }
    for (uint32_t by = 0; by < 2; by++)
        for (uint32_t bx = 0; bx < 2; bx++) {
            const uint32_t blk = c.b + (by ? c.width : 0) + bx;
            a.block_encodings[blk] = (uint8_t)best_e;
            a.tile_indices[blk] = c.toff | c_tile_map[best_e][by][bx];
        }
}

// ---------------------------------------------------------- palettizing

__device__ inline float sqd(const float* x, const float* y, int d)
{
    float s = 0.0f;
    for (int i = 0; i < d; i++) {
        const float t = x[i] - y[i];
        s += t * t;
    }
    return s;
}

__device__ inline float dotv(const float* x, const float* y, int d)
{
    float s = x[0] * y[0];
    for (int i = 1; i < d; i++) s += x[i] * y[i];
    return s;
}

// split_vectors (crn_tree_clusterizer.h:573-753), D = 1 or 3, sums in vector order.
template <int D>
__device__ void split_vectors(const float* v, const uint32_t* w, uint32_t size, float* result)
{
    float wv[64 * D];
    double wd[64];
    float c[D];
    for (int k = 0; k < D; k++) c[k] = 0.0f;
    double tt = 0.0;
    uint64_t tw = 0;
    for (uint32_t i = 0; i < size; i++) {
        const float fw = (float)w[i];
        for (int k = 0; k < D; k++) {
            wv[i * D + k] = v[i * D + k] * fw;
            c[k] += wv[i * D + k];
        }
        tw += w[i];
        wd[i] = dotv(v + i * D, v + i * D, D) * fw;
        tt += wd[i];
    }
    const float var = (float)(tt - (dotv(c, c, D) / (float)tw));
    const float inv = 1.0f / (float)tw;
    for (int k = 0; k < D; k++) c[k] *= inv;
    for (int k = 0; k < D; k++) result[k] = result[D + k] = c[k];
    if (var <= 0.0f || size == 1) return;
    const float *fur = v, *opp = v;
    double fd = -1.0f, od = -1.0f;
    for (uint32_t i = 0; i < size; i++) {
        const double dd = sqd(v + i * D, c, D);
        if (dd > fd) {
            fd = dd;
            fur = v + i * D;
        }
    }
    for (uint32_t i = 0; i < size; i++) {
        const double dd = sqd(v + i * D, fur, D);
        if (dd > od) {
            od = dd;
            opp = v + i * D;
        }
    }
    float lc[D], rc[D];
    for (int k = 0; k < D; k++) {
        lc[k] = (fur[k] + c[k]) * .5f;
        rc[k] = (opp[k] + c[k]) * .5f;
    }
    if (size > 2) {
        float cov[D][D];
        for (int x = 0; x < D; x++)
            for (int y = 0; y < D; y++) cov[x][y] = 0.0f;
        for (uint32_t i = 0; i < size; i++) {
            float x[D], y[D];
            for (int k = 0; k < D; k++) {
                x[k] = v[i * D + k] - c[k];
                y[k] = x[k] * (float)w[i];
            }
            for (int p = 0; p < D; p++)
                for (int q = p; q < D; q++) cov[p][q] = cov[p][q] + x[p] * y[q];
        }
        const float div = (float)tw;
        for (int p = 0; p < D; p++)
            for (int q = p; q < D; q++) {
                cov[p][q] /= div;
                cov[q][p] = cov[p][q];
            }
        float axis[D];
        for (int k = 0; k < D; k++) axis[k] = 1.0f;
        for (int it = 0; it < 10; it++) {
            float x[D];
            double ms = 0;
            for (int i = 0; i < D; i++) {
                double sum = 0;
                for (int j = 0; j < D; j++) sum += axis[j] * cov[i][j];
                x[i] = (float)sum;
                ms = i ? (ms > sum ? ms : sum) : sum;
            }
            if (ms != 0.0f) {
                const float sc = (float)(1.0f / ms);
                for (int i = 0; i < D; i++) x[i] *= sc;
            }
            for (int i = 0; i < D; i++) axis[i] = x[i];
        }
        double n2 = axis[0] * axis[0];
        for (int i = 1; i < D; i++) n2 += axis[i] * axis[i];
        if (n2 != 0) {
            const float sc = (float)(1.0f / sqrt(n2));
            for (int i = 0; i < D; i++) axis[i] *= sc;
        }
        float nl[D], nr[D];
        for (int k = 0; k < D; k++) nl[k] = nr[k] = 0.0f;
        double lw = 0.0f, rw = 0.0f;
        for (uint32_t i = 0; i < size; i++) {
            float x[D];
            for (int k = 0; k < D; k++) x[k] = v[i * D + k] - c[k];
            const float fw = (float)w[i];
            if (dotv(x, axis, D) < 0.0f) {
                for (int k = 0; k < D; k++) nl[k] += wv[i * D + k];
                lw += fw;
            } else {
                for (int k = 0; k < D; k++) nr[k] += wv[i * D + k];
                rw += fw;
            }
        }
        if (lw > 0.0f && rw > 0.0f) {
            const float sl = (float)(1.0f / lw), sr = (float)(1.0f / rw);
            for (int k = 0; k < D; k++) {
                lc[k] = nl[k] * sl;
                rc[k] = nr[k] * sr;
            }
        }
    }
    float prev = 1e+10f;
    for (int loop = 0; loop < 1024; loop++) {
        float nl[D], nr[D];
        for (int k = 0; k < D; k++) nl[k] = nr[k] = 0.0f;
        double lt = 0.0f, rt = 0.0f;
        uint64_t lw = 0, rw = 0;
        for (uint32_t i = 0; i < size; i++) {
            const float* x = v + i * D;
            if (sqd(lc, x, D) < sqd(rc, x, D)) {
                for (int k = 0; k < D; k++) nl[k] += wv[i * D + k];
                lt += wd[i];
                lw += w[i];
            } else {
                for (int k = 0; k < D; k++) nr[k] += wv[i * D + k];
                rt += wd[i];
                rw += w[i];
            }
        }
        if (!lw || !rw) return;
        const float lv = (float)(lt - (dotv(nl, nl, D) / (float)lw));
        const float rv = (float)(rt - (dotv(nr, nr, D) / (float)rw));
        const float il = 1.0f / (float)lw, ir = 1.0f / (float)rw;
        for (int k = 0; k < D; k++) {
            lc[k] = nl[k] * il;
            rc[k] = nr[k] * ir;
        }
        const float total = lv + rv;
        if (total < .00001f) break;
        if (((prev - total) / total) < .00001f) break;
        prev = total;
    }
    for (int k = 0; k < D; k++) {
        result[k] = lc[k];
        result[D + k] = rc[k];
    }
}

__device__ void sort_u32(uint32_t* x, uint32_t n)
{
    for (uint32_t i = 1; i < n; i++) {
        const uint32_t v = x[i];
        uint32_t j = i;
        while (j && x[j - 1] > v) {
            x[j] = x[j - 1];
            j--;
        }
        x[j] = v;
    }
}

__global__ void __launch_bounds__(64) tile_vec_kernel(TileArgs a)
{
    const uint32_t slot = blockIdx.x * blockDim.x + threadIdx.x;
    if (slot >= a.num_blocks) return;
    const uint32_t src = a.tile_src[slot];
    if (src == 0xFFFFFFFFu) return;
    const uint32_t b = src >> 4, t = src & 15u;
    const uint32_t width = a.slot_width[slot];
    const uint32_t n = 16u << (t >> 2);
    float* out = a.tile_vec + (size_t)slot * 10;
    uint32_t col[64], wts[64], size;
    if (a.has_color) {
        for (uint32_t i = 0; i < n; i++) col[i] = tile_texel(a.blocks, b, width, t, i) & 0xFFFFFFu;
        // crnlib orders colours as r << 16 | g << 8 | b
        for (uint32_t i = 0; i < n; i++) col[i] = (col[i] & 255u) << 16 | (col[i] & 0xFF00u) | (col[i] >> 16);
        sort_u32(col, n);
        float v[64 * 3];
        size = 0;
        for (uint32_t i = 0; i < n; i++) {
            if (!i || col[i] != col[i - 1]) {
                v[size * 3] = a.perceptual ? c_u2f[col[i] >> 16] * 0.5f : c_u2f[col[i] >> 16];
                v[size * 3 + 1] = c_u2f[col[i] >> 8 & 0xFF];
                v[size * 3 + 2] = a.perceptual ? c_u2f[col[i] & 0xFF] * 0.25f : c_u2f[col[i] & 0xFF];
                wts[size++] = 1;
            } else {
                wts[size - 1]++;
            }
        }
        float r[6];
        split_vectors<3>(v, wts, size, r);
        const float l0 = sqrtf(dotv(r, r, 3)), l1 = sqrtf(dotv(r + 3, r + 3, 3));
        if (l0 > l1) {
            for (int k = 0; k < 3; k++) {
                out[k] = r[3 + k];
                out[3 + k] = r[k];
            }
        } else {
            for (int k = 0; k < 6; k++) out[k] = r[k];
        }
    }
    for (int k = 0; k < a.num_alpha; k++) {
        for (uint32_t i = 0; i < n; i++) col[i] = (tile_texel(a.blocks, b, width, t, i) >> (8 * a.alpha_comp[k])) & 255u;
        sort_u32(col, n);
        float v[64];
        size = 0;
        for (uint32_t i = 0; i < n; i++) {
            if (!i || col[i] != col[i - 1]) {
                v[size] = c_u2f[col[i]];
                wts[size++] = 1;
            } else {
                wts[size - 1]++;
            }
        }
        float r[2];
        split_vectors<1>(v, wts, size, r);
        if (r[1] < r[0]) {
            out[6 + 2 * k] = r[1];
            out[7 + 2 * k] = r[0];
        } else {
            out[6 + 2 * k] = r[0];
            out[7 + 2 * k] = r[1];
        }
    }
}

bool g_tables_ready = false;

int init_tables()
{
    if (g_tables_ready) return CRN2_OK;
    static uint8_t t5[256][2], t6[256][2];
    static float u2f[256];
    // ryg single-colour tables for 4-colour blocks (crn_ryg_dxt.cpp:58-99)
    for (int which = 0; which < 2; which++) {
        const int size = which ? 64 : 32;
        uint8_t(*t)[2] = which ? t6 : t5;
        for (int i = 0; i < 256; i++) {
            int best = 256;
            for (int mn = 0; mn < size; mn++)
                for (int mx = 0; mx < size; mx++) {
                    const int me = size == 32 ? ((mn << 3) | (mn >> 2)) : ((mn << 2) | (mn >> 4));
                    const int xe = size == 32 ? ((mx << 3) | (mx >> 2)) : ((mx << 2) | (mx >> 4));
                    int err = abs(((xe * 2 + me) / 3) - i);
                    err += (abs(xe - me) * 8) >> 8;
                    if (err < best) {
                        t[i][0] = (uint8_t)mx;
                        t[i][1] = (uint8_t)mn;
                        best = err;
                    }
                }
        }
    }
    for (int i = 0; i < 256; i++) u2f[i] = i * 1.0f / 255.0f;
    if (hipMemcpyToSymbol(HIP_SYMBOL(c_om5), t5, 512) != hipSuccess ||
        hipMemcpyToSymbol(HIP_SYMBOL(c_om6), t6, 512) != hipSuccess ||
        hipMemcpyToSymbol(HIP_SYMBOL(c_u2f), u2f, sizeof(u2f)) != hipSuccess)
        return CRN2_EHIP;
    g_tables_ready = true;
    return CRN2_OK;
}

}  // namespace

extern "C" int crn2_hc_tiles_device(const void* blocks, uint32_t num_blocks, uint32_t nlevels,
                                    const uint32_t* level_desc, uint32_t num_faces,
                                    int has_color, uint32_t num_alpha, const uint32_t* alpha_comp, int perceptual,
                                    float color_derating, float alpha_derating, float ratio,
                                    uint8_t* block_encodings, uint32_t* tile_indices, uint32_t* tile_src,
                                    float* tile_vec, void* stream)
{
    if (!blocks || !num_blocks || !nlevels || nlevels > 16 || !level_desc || !num_faces || num_alpha > 2 ||
        (!has_color && !num_alpha) || (num_alpha && !alpha_comp) || !block_encodings || !tile_indices ||
        !tile_src || !tile_vec)
        return CRN2_EINVAL;
    const int rc0 = init_tables();
    if (rc0 != CRN2_OK) return rc0;
    hipStream_t st = static_cast<hipStream_t>(stream);
    TileArgs a;
    a.blocks = static_cast<const uint32_t*>(blocks);
    a.has_color = has_color ? 1 : 0;
    a.num_alpha = (int)num_alpha;
    a.perceptual = perceptual ? 1 : 0;
    a.alpha_comp[0] = num_alpha > 0 ? alpha_comp[0] : 0;
    a.alpha_comp[1] = num_alpha > 1 ? alpha_comp[1] : 0;
    if (a.alpha_comp[0] > 3 || a.alpha_comp[1] > 3) return CRN2_EINVAL;
    a.ratio = ratio;
    // derating tables (crn_dxt_hc.cpp:122-142)
    static const uint32_t kTileDerating[8] = {0, 1, 1, 2, 2, 2, 2, 3};
    for (uint32_t l = 0; l < nlevels; l++) {
        float d = color_derating;
        if (l && d > .25f) {
            const float q = d / powf(3.0f, (float)l);
            d = .25f > q ? .25f : q;
        }
        for (int e = 0; e < 8; e++) a.color_der[l][e] = 0.0f + (d - 0.0f) * (kTileDerating[e] / 3.0f);
    }
    for (int e = 0; e < 8; e++) a.alpha_der[e] = 0.0f + (alpha_derating - 0.0f) * (kTileDerating[e] / 3.0f);
    // chunks in the reference's order with their tile slots (crn_dxt_hc.cpp:404-423)
    std::vector<Chunk> chunks;
    chunks.reserve(num_blocks / 4);
    for (uint32_t l = 0; l < nlevels; l++) {
        const uint32_t first = level_desc[3 * l], nb = level_desc[3 * l + 1], width = level_desc[3 * l + 2];
        if (!width || nb % width) return CRN2_EINVAL;
        const uint32_t height = nb / width, face_h = height / num_faces;
        if ((width & 1) || (face_h & 1) || face_h * num_faces != height) return CRN2_EUNSUPPORTED;
        if ((uint64_t)first + nb > num_blocks) return CRN2_EINVAL;
        uint32_t hface = 0, b = first;
        for (uint32_t h = 0; h < height; h += 2, hface += 2, b += width) {
            uint32_t toff = b, tdelta = 4;
            if (hface == face_h) hface = 0;
            else if (hface & 2) {
                tdelta = (uint32_t)-4;
                toff += (width << 1) + tdelta;
            }
            for (uint32_t bn = b + width; b < bn; b += 2, toff += tdelta) chunks.push_back(Chunk{b, width, l, toff});
        }
    }
    if (chunks.empty()) return CRN2_OK;
    Chunk* dch = nullptr;
    uint32_t* terr = nullptr;
    uint32_t* swidth = nullptr;
    const uint32_t nch = (uint32_t)chunks.size();
    int rc = CRN2_OK;
    if (hipMallocAsync((void**)&dch, nch * sizeof(Chunk), st) != hipSuccess ||
        hipMallocAsync((void**)&terr, (size_t)nch * 27 * sizeof(uint32_t), st) != hipSuccess ||
        hipMallocAsync((void**)&swidth, (size_t)num_blocks * sizeof(uint32_t), st) != hipSuccess)
        rc = CRN2_ENOMEM;
    a.slot_width = swidth;
    a.chunks = dch;
    a.nchunks = nch;
    a.terr = terr;
    a.block_encodings = block_encodings;
    a.tile_indices = tile_indices;
    a.tile_src = tile_src;
    a.tile_vec = tile_vec;
    a.num_blocks = num_blocks;
    if (rc == CRN2_OK && (hipMemcpyAsync(dch, chunks.data(), nch * sizeof(Chunk), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemsetAsync(tile_src, 0xFF, (size_t)num_blocks * sizeof(uint32_t), st) != hipSuccess ||
        hipMemsetAsync(tile_vec, 0, (size_t)num_blocks * 10 * sizeof(float), st) != hipSuccess))
        rc = CRN2_EHIP;
    if (rc == CRN2_OK) {
        if (has_color) hipLaunchKernelGGL(tile_error_kernel, dim3((nch * 9 + 255) / 256), dim3(256), 0, st, a);
        if (num_alpha) {
            const uint32_t jobs = nch * 9 * num_alpha;
            hipLaunchKernelGGL(tile_alpha_kernel, dim3((jobs + kAlphaWaves - 1) / kAlphaWaves), dim3(64 * kAlphaWaves),
                               0, st, a);
        }
        hipLaunchKernelGGL(tile_choose_kernel, dim3((nch + 255) / 256), dim3(256), 0, st, a);
        hipLaunchKernelGGL(tile_vec_kernel, dim3((num_blocks + 63) / 64), dim3(64), 0, st, a);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) rc = CRN2_EHIP;
    }
    for (void* p : {(void*)dch, (void*)terr, (void*)swidth})
        if (p) (void)hipFreeAsync(p, st);
    (void)hipStreamSynchronize(st);
    return rc;
}
