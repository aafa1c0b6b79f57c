// qdxt_gpu.hip -- crnlib's clustered DXT1 compressor (qdxt1,
// crnlib/crn_qdxt1.cpp) and the vector quantizers it runs
// (clusterizer, crnlib/crn_clusterizer.h; threaded_clusterizer,
// crnlib/crn_threaded_clusterizer.h) on gfx950.
//
// Pipeline (one qdxt1 object = one init, any number of packs):
//   init   training vectors: per 2x2-block chunk one thread per tile layout
//          (dxt_fast compression + error), one thread per chunk for the
//          encoding choice and the chosen tiles' representative colours;
//          the 6-D endpoint tree grown to 65,535 leaves; the distinct
//          dxt_fast selector words counted in a device hash set.
//   pack   endpoint clusters cut from the tree on the host (depth first);
//          their texels gathered and searched one wave per cluster by the
//          cluster endpoint search (dxt1c_gpu.hip) or, at quality 255,
//          block by block (dxt1_gpu.hip); selector vectors per block; the
//          threaded selector clusters (three PCA divisions, four trees);
//          the selector optimisation one wave per selector cluster, lanes =
//          (texel, selector value).
//
// The tree.  The reference grows it best first: a max-heap on node
// variance pops the worst node, which is split (PCA start, up to eight
// rounds of 2-means) and its children pushed.  A split depends only on the
// node's own vectors, so the device splits nodes in rounds, one wave per
// node, and the host replays the heap (the reference's insert / down-heap
// on the same float variances) as far as the splits computed so far allow;
// when it pops a node not yet split it stops, and the next round splits
// that node together with every heap node that could still be popped
// within the remaining leaf budget (those with fewer than that many heap
// nodes of higher variance).  Node vectors live in a pool of index segments:
// a split writes its children (left part then right part, each in the
// parent's order) to a fresh segment, so every node's list stays in
// ascending vector order, which is the order of the reference's lists.
//
// Sums.  The reference accumulates float sums over a node's vectors in list
// order (centroids, the covariance, the 2-means centres).  The split kernel
// keeps that order: a batch of 64 vectors is staged in LDS by all lanes and
// lane k adds the batch's terms of component k (or covariance entry k) one
// after another.  Integer sums (weights, sums of squares: exact in double)
// are wave reductions.  Per-vector work (projections, distances, the side
// of each vector) is one vector per lane.  The file is compiled with
// -ffp-contract=off so that no multiply-add is fused, as in the reference's
// x86-64 build.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "crn2_internal.h"
#include "dxt_fast_dev.h"

namespace {

using namespace dxtf;

__constant__ uint8_t q_om5[256][2], q_om6[256][2];

// g_chunk_tile_layouts (crnlib/crn_dxt_hc_common.cpp:47-63): x, y, width, height
__constant__ uint8_t c_layout[9][4] = {{0, 0, 8, 8}, {0, 0, 8, 4}, {0, 4, 8, 4}, {0, 0, 4, 8}, {4, 0, 4, 8},
                                       {0, 0, 4, 4}, {4, 0, 4, 4}, {0, 4, 4, 4}, {4, 4, 4, 4}};
// g_chunk_encodings (:29-45): tile count, then the layout of each tile
__constant__ uint8_t c_enc[8][5] = {{1, 0}, {2, 1, 2}, {2, 3, 4}, {3, 1, 7, 8}, {3, 2, 5, 6},
                                    {3, 3, 6, 8}, {3, 4, 5, 7}, {4, 5, 6, 7, 8}};
__constant__ float c_linear[4] = {0.0f, 3.0f, 1.0f, 2.0f};   // g_dxt1_to_linear

__device__ inline uint32_t lane() { return threadIdx.x & 63u; }

// ------------------------------------------------------------ init: training
struct TChunk {
    uint32_t first, bw, bh, cx, cy;
    float der;    // the level's colour PSNR derating
};

// Chunk texel (x, y), clamped to the level's block area (crn_qdxt1.cpp:127-147).
__device__ inline uint32_t chunk_texel(const uint32_t* __restrict__ blocks, const TChunk& c, uint32_t x, uint32_t y)
{
    const uint32_t lw = c.bw * 4, lh = c.bh * 4;
    uint32_t px = c.cx * 8 + x, py = c.cy * 8 + y;
    px = px < lw - 1 ? px : lw - 1;
    py = py < lh - 1 ? py : lh - 1;
    return blocks[(size_t)(c.first + (py >> 2) * c.bw + (px >> 2)) * 16 + (py & 3) * 4 + (px & 3)];
}

// dxt1_block::get_block_colors: 4 colours when low > high, else 3 + black
__device__ inline void block_colors(uint32_t lo, uint32_t hi, int (&col)[4][3])
{
    unpack565(lo, col[0]);
    unpack565(hi, col[1]);
    for (int k = 0; k < 3; k++) {
        if (lo > hi) {
            col[2][k] = (col[0][k] * 2 + col[1][k]) / 3;
            col[3][k] = (col[1][k] * 2 + col[0][k]) / 3;
        } else {
            col[2][k] = (col[0][k] + col[1][k]) >> 1;
            col[3][k] = 0;
        }
    }
}

// One tile layout: dxt_fast compression and the RGB squared error against
// the resulting block colours (crn_qdxt1.cpp:172-197), texels in raster
// order within the tile.
template <int N>
__device__ uint32_t layout_error(const uint32_t* __restrict__ blocks, const TChunk& c, uint32_t l)
{
    const uint32_t xo = c_layout[l][0], yo = c_layout[l][1], w = c_layout[l][2];
    uint32_t px[N];
#pragma unroll
    for (int i = 0; i < N; i++) px[i] = chunk_texel(blocks, c, xo + (uint32_t)i % w, yo + (uint32_t)i / w);
    uint8_t sel[N];
    uint32_t lo, hi;
    compress_color<N>(px, lo, hi, sel, q_om5, q_om6);
    int col[4][3];
    block_colors(lo, hi, col);
    uint32_t err = 0;
#pragma unroll
    for (int i = 0; i < N; i++)
        for (int k = 0; k < 3; k++) {
            const int d = ch(px[i], k) - col[sel[i]][k];
            err += (uint32_t)(d * d);
        }
    return err;
}

__global__ void __launch_bounds__(256) qdxt_layout_kernel(const uint32_t* __restrict__ blocks,
                                                          const TChunk* __restrict__ chunks, uint32_t nchunks,
                                                          uint32_t* __restrict__ lerr)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nchunks * 9) return;
    const TChunk c = chunks[i / 9];
    const uint32_t l = i % 9;
    lerr[i] = l == 0 ? layout_error<64>(blocks, c, l) : (l < 5 ? layout_error<32>(blocks, c, l)
                                                               : layout_error<16>(blocks, c, l));
}

// dxt_fast::find_representative_colors (crn_dxt_fast.cpp:855-997) over n
// texels given by texel(i); colours packed r | g << 8 | b << 16.
template <class F>
__device__ void rep_colors(uint32_t n, F texel, uint32_t& lo_out, uint32_t& hi_out)
{
    uint64_t s[3] = {0, 0, 0};
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t p = texel(i);
        for (int k = 0; k < 3; k++) s[k] += (uint32_t)ch(p, k);
    }
    int ave[3];
    for (int k = 0; k < 3; k++) ave[k] = (int)((s[k] + (n / 2)) / n);
    int fd = -1;
    uint32_t fp = texel(0);
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t p = texel(i);
        const int r = ch(p, 0) - ave[0], g = ch(p, 1) - ave[1], b = ch(p, 2) - ave[2];
        const int d = r * r + g * g + b * b;
        if (d > fd) {
            fd = d;
            fp = p;
        }
    }
    int l[3] = {ch(fp, 0), ch(fp, 1), ch(fp, 2)};
    int od = -1;
    uint32_t op = texel(0);
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t p = texel(i);
        const int r = ch(p, 0) - l[0], g = ch(p, 1) - l[1], b = ch(p, 2) - l[2];
        const int d = r * r + g * g + b * b;
        if (d > od) {
            od = d;
            op = p;
        }
    }
    int h[3] = {ch(op, 0), ch(op, 1), ch(op, 2)};
    for (int k = 0; k < 3; k++) {
        l[k] = (l[k] + ave[k]) >> 1;
        h[k] = (h[k] + ave[k]) >> 1;
    }
    for (int iter = 0; iter < 4; iter++) {
        if (l[0] == h[0] && l[1] == h[1] && l[2] == h[2]) break;
        uint64_t nc[2][3] = {{0, 0, 0}, {0, 0, 0}};
        uint32_t wt[2] = {0, 0};
        int vr = h[0] - l[0], vg = h[1] - l[1], vb = h[2] - l[2];
        const int mid = (vr * l[0] + vg * l[1] + vb * l[2]) + (vr * h[0] + vg * h[1] + vb * h[2]);
        vr *= 2;
        vg *= 2;
        vb *= 2;
        for (uint32_t i = 0; i < n; i++) {
            const uint32_t p = texel(i);
            const int dot = ch(p, 0) * vr + ch(p, 1) * vg + ch(p, 2) * vb;
            const uint32_t m = dot > mid ? 1u : 0u;
            for (int k = 0; k < 3; k++) nc[m][k] += (uint32_t)ch(p, k);
            wt[m]++;
        }
        if (!wt[0] || !wt[1]) break;
        int n8[2][3];
        for (int j = 0; j < 2; j++)
            for (int k = 0; k < 3; k++) n8[j][k] = (int)(uint8_t)((nc[j][k] + (wt[j] / 2)) / wt[j]);
        if (n8[0][0] == l[0] && n8[0][1] == l[1] && n8[0][2] == l[2] && n8[1][0] == h[0] && n8[1][1] == h[1] &&
            n8[1][2] == h[2])
            break;
        for (int k = 0; k < 3; k++) {
            l[k] = n8[0][k];
            h[k] = n8[1][k];
        }
    }
    const uint32_t e0 = (uint32_t)(l[0] * l[0] + l[1] * l[1] + l[2] * l[2]);
    const uint32_t e1 = (uint32_t)(h[0] * h[0] + h[1] * h[1] + h[2] * h[2]);
    const uint32_t lp = (uint32_t)l[0] | (uint32_t)l[1] << 8 | (uint32_t)l[2] << 16;
    const uint32_t hp = (uint32_t)h[0] | (uint32_t)h[1] << 8 | (uint32_t)h[2] << 16;
    lo_out = e0 > e1 ? hp : lp;
    hi_out = e0 > e1 ? lp : hp;
}

__device__ inline void put_training(float* __restrict__ vecs, uint32_t* __restrict__ weights, uint32_t b,
                                    uint32_t lo, uint32_t hi)
{
    int dist = 0;
    for (int k = 0; k < 3; k++) {
        vecs[6 * (size_t)b + k] = (float)ch(lo, k);
        vecs[6 * (size_t)b + 3 + k] = (float)ch(hi, k);
        const int d = ch(lo, k) - ch(hi, k);
        dist += d * d;
    }
    const uint32_t w = (uint32_t)dist / 5000;
    weights[b] = w < 1 ? 1 : (w > 8 ? 8 : w);
}

// Encoding choice of one chunk (crn_qdxt1.cpp:199-247) and the training
// vectors of its blocks from the chosen tiles (:249-340).
__global__ void __launch_bounds__(256) qdxt_choose_kernel(const uint32_t* __restrict__ blocks,
                                                          const TChunk* __restrict__ chunks, uint32_t nchunks,
                                                          const uint32_t* __restrict__ lerr,
                                                          float* __restrict__ vecs, uint32_t* __restrict__ weights,
                                                          uint8_t* __restrict__ enc)
{
    const uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (ci >= nchunks) return;
    const TChunk c = chunks[ci];
    double best = -1.0f;
    uint32_t be = 0;
    for (uint32_t e = 0; e < 8; e++) {
        double total = 0;
        for (uint32_t t = 0; t < c_enc[e][0]; t++) total += (double)lerr[ci * 9 + c_enc[e][1 + t]];
        const double ms = total * (1.0f / (64.0f * 3.0f));
        const double rms = sqrt(ms);
        double psnr = 999999.0f;
        if (ms) {
            psnr = log10(255.0f / rms) * 20.0f;
            psnr = psnr < 0.0 ? 0.0 : (psnr > 500.0 ? 500.0 : psnr);
        }
        const float cder = 0.0f + (c.der - 0.0f) * ((float)(c_enc[e][0] - 1) / 3.0f);
        psnr = psnr - cder;
        if (psnr > best) {
            best = psnr;
            be = e;
        }
    }
    for (uint32_t t = 0; t < c_enc[be][0]; t++) {
        const uint32_t l = c_enc[be][1 + t];
        const uint32_t xo = c_layout[l][0], yo = c_layout[l][1], w = c_layout[l][2], h = c_layout[l][3];
        uint32_t lo, hi;
        rep_colors(w * h, [&](uint32_t i) { return chunk_texel(blocks, c, xo + i % w, yo + i / w); }, lo, hi);
        for (uint32_t y = 0; y < (h >> 2); y++) {
            const uint32_t by = c.cy * 2 + y + (yo >> 2);
            if (by >= c.bh) continue;
            for (uint32_t x = 0; x < (w >> 2); x++) {
                const uint32_t bx = c.cx * 2 + x + (xo >> 2);
                if (bx >= c.bw) break;
                const uint32_t b = c.first + bx + by * c.bw;
                put_training(vecs, weights, b, lo, hi);
                if (enc) enc[b] = (uint8_t)be;
            }
        }
    }
}

// Non-hierarchical training vectors: one block's 16 texels (crn_qdxt1.cpp:368-399).
__global__ void __launch_bounds__(256) qdxt_block_train_kernel(const uint32_t* __restrict__ blocks, uint32_t n,
                                                               float* __restrict__ vecs,
                                                               uint32_t* __restrict__ weights,
                                                               uint8_t* __restrict__ enc)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const uint32_t* p = blocks + (size_t)b * 16;
    uint32_t lo, hi;
    rep_colors(16, [&](uint32_t i) { return p[i]; }, lo, hi);
    put_training(vecs, weights, b, lo, hi);
    if (enc) enc[b] = 255;
}

__device__ inline uint32_t hash32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7feb352dU;
    x ^= x >> 15;
    x *= 0x846ca68bU;
    x ^= x >> 16;
    return x;
}

// The distinct selector words of dxt_fast::compress_color_block per block
// (crn_qdxt1.cpp:409-426), counted in an open-addressing set (key + 1, 0 empty).
__global__ void __launch_bounds__(256) qdxt_selword_kernel(const uint32_t* __restrict__ blocks, uint32_t n,
                                                           unsigned long long* __restrict__ table, uint32_t mask,
                                                           uint32_t* __restrict__ count)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    uint32_t px[16];
#pragma unroll
    for (int i = 0; i < 16; i++) px[i] = blocks[(size_t)b * 16 + i];
    uint8_t sel[16];
    uint32_t lo, hi;
    compress_color<16>(px, lo, hi, sel, q_om5, q_om6);
    uint32_t word = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) word |= (uint32_t)sel[i] << (2 * i);
    const unsigned long long key = (unsigned long long)word + 1ull;
    uint32_t h = hash32(word) & mask;
    for (;;) {
        const unsigned long long prev = atomicCAS(&table[h], 0ull, key);
        if (prev == 0ull) {
            atomicAdd(count, 1u);
            break;
        }
        if (prev == key) break;
        h = (h + 1) & mask;
    }
}

// --------------------------------------------------------------- the trees
struct VqSplit {
    uint32_t node;        // device slot of the node's centroid and weight
    uint32_t seg, n;      // its vectors: pool[seg, seg + n)
    uint32_t child_seg;   // the children's vectors go to pool[child_seg, child_seg + n)
    uint32_t left, right; // device slots for the children
};

struct VqResult {
    uint32_t split;       // 0: unsplittable
    uint32_t nl, nr;
    float lvar, rvar;
};

struct VqArgs {
    const float* v;                 // vectors, D floats each
    const uint32_t* w;              // weights
    uint32_t* pool;                 // index segments
    float* cen;                     // per slot: D floats
    unsigned long long* tw;         // per slot
    unsigned long long* passes;     // profiling: vector passes of the splits (NULL: off)
    float exact_max;                // > 0: components are integers in [0, exact_max] (see side_pass)
};

// Float sums of non-negative integers are exact in any order while the total
// stays within 2^24: a node whose weight times the largest component is at
// most that sums its vectors by a wave reduction (same bits as the list
// order); others keep the serial order.
__device__ inline bool exact_sums(const VqArgs& A, unsigned long long total_weight)
{
    return A.exact_max > 0.0f && (double)total_weight * (double)A.exact_max <= 16777216.0;
}

__device__ inline float wave_sum_f(float v)
{
    for (int o = 32; o; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

// Waves per node in the split and division kernels.  Every wave runs the
// whole node (the same loads, LDS stores of the same values and in-order
// sums, so each holds every result); only the covariance entries are split
// between them, which for D = 16 (136 entries) takes one round of lanes
// instead of three.
template <int D>
constexpr int kVqWaves = D >= 16 ? 3 : 1;

template <int D>
struct VqLds {
    static constexpr int kE = D * (D + 1) / 2;
    static constexpr int kR = D >= 16 ? 4 : 8;   // vectors per lane per batch
    static constexpr int kRow = 64 * kR + 4;     // component rows (padded: rows start in different banks)
    alignas(16) float a[D][kRow];   // staged batch, component-major: vector - centroid, or products
    alignas(16) float b[D][kRow];   // staged batch: (vector - centroid) * weight
    uint32_t side[64];
    float cov[D][D];
    float axis[D], prev[D], x[D];
    double red[64];
};

__device__ inline double wave_sum_d(double v)
{
    for (int o = 32; o; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ inline unsigned long long wave_sum_u64(unsigned long long v)
{
    for (int o = 32; o; o >>= 1) v += __shfl_xor(v, o);
    return v;
}
__device__ inline double wave_max_d(double v)
{
    for (int o = 32; o; o >>= 1) {
        const double t = __shfl_xor(v, o);
        v = t > v ? t : v;
    }
    return v;
}

template <int D>
__device__ inline float vdot(const float (&a)[D], const float (&b)[D])
{
    float r = a[0] * b[0];
#pragma unroll
    for (int i = 1; i < D; i++) r += a[i] * b[i];
    return r;
}

template <int D>
__device__ inline float vsqd(const float (&a)[D], const float (&b)[D])
{
    float s = 0;
#pragma unroll
    for (int i = 0; i < D; i++) {
        const float t = a[i] - b[i];
        s += t * t;
    }
    return s;
}

template <int D>
__device__ inline void load_vec(const VqArgs& A, uint32_t vi, float (&v)[D])
{
#pragma unroll
    for (int c = 0; c < D; c++) v[c] = A.v[(size_t)vi * D + c];
}

// The list walks below keep U chunks of 64 vectors in flight per round:
// their index and vector loads are issued together, so a pass over a long
// list waits for memory once per U chunks instead of once per chunk.

// Sum of the listed vectors' weights (integers: any order).
__device__ inline unsigned long long list_weight(const VqArgs& A, const uint32_t* __restrict__ list, uint32_t n)
{
    constexpr uint32_t U = 8;
    const uint32_t me = lane();
    unsigned long long s = 0;
    for (uint32_t j0 = 0; j0 < n; j0 += 64 * U) {
        uint32_t vi[U];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t j = j0 + 64 * u + me;
            vi[u] = j < n ? list[j] : 0xFFFFFFFFu;
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) s += vi[u] != 0xFFFFFFFFu ? A.w[vi[u]] : 0u;
    }
    return wave_sum_u64(s);
}

// Number of listed vectors for which is_left(v) holds.
template <int D, class F>
__device__ inline uint32_t list_count(const VqArgs& A, const uint32_t* __restrict__ list, uint32_t n, F&& is_left)
{
    constexpr uint32_t U = D >= 16 ? 4 : 8;
    const uint32_t me = lane();
    uint32_t cnt = 0;
    for (uint32_t j0 = 0; j0 < n; j0 += 64 * U) {
        uint32_t vi[U];
        float v[U][D];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t j = j0 + 64 * u + me;
            vi[u] = j < n ? list[j] : list[0];
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) load_vec<D>(A, vi[u], v[u]);
#pragma unroll
        for (uint32_t u = 0; u < U; u++) cnt += (j0 + 64 * u + me < n && is_left(v[u])) ? 1u : 0u;
    }
    return (uint32_t)wave_sum_u64(cnt);
}

// Stable partition of list[0, n) into out: the vectors for which is_left(v)
// holds from out[0] on, the others from out[nl] on (nl = their count), each
// side in list order.
template <int D, class F>
__device__ inline void list_partition(const VqArgs& A, const uint32_t* __restrict__ list, uint32_t n, uint32_t nl,
                                      uint32_t* __restrict__ out, F&& is_left)
{
    constexpr uint32_t U = D >= 16 ? 4 : 8;
    const uint32_t me = lane();
    const unsigned long long below = me ? (~0ull >> (64 - me)) : 0ull;
    uint32_t lpos = 0, rpos = nl;
    for (uint32_t j0 = 0; j0 < n; j0 += 64 * U) {
        uint32_t vi[U];
        float v[U][D];
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const uint32_t j = j0 + 64 * u + me;
            vi[u] = j < n ? list[j] : list[0];
        }
#pragma unroll
        for (uint32_t u = 0; u < U; u++) load_vec<D>(A, vi[u], v[u]);
#pragma unroll
        for (uint32_t u = 0; u < U; u++) {
            const bool ok = j0 + 64 * u + me < n;
            const bool left = ok && is_left(v[u]);
            const unsigned long long lm = __ballot(left), rm = __ballot(ok && !left);
            if (ok) {
                if (left) out[lpos + (uint32_t)__popcll(lm & below)] = vi[u];
                else out[rpos + (uint32_t)__popcll(rm & below)] = vi[u];
            }
            lpos += (uint32_t)__popcll(lm);
            rpos += (uint32_t)__popcll(rm);
        }
    }
}

// acc + x[0] + x[1] + ... + x[cnt - 1], added one after another; x is 16-byte
// aligned LDS.  Reads 16 values ahead (four 128-bit loads) while adding.
__device__ inline float serial_sum(float acc, const float* x, uint32_t cnt)
{
    const float4* x4 = reinterpret_cast<const float4*>(x);
    uint32_t k = 0;
    if (cnt >= 16) {
        float4 q0 = x4[0], q1 = x4[1], q2 = x4[2], q3 = x4[3];
        for (; k + 16 <= cnt; k += 16) {
            float4 n0 = q0, n1 = q1, n2 = q2, n3 = q3;
            if (k + 32 <= cnt) {
                const uint32_t g = (k + 16) >> 2;
                n0 = x4[g];
                n1 = x4[g + 1];
                n2 = x4[g + 2];
                n3 = x4[g + 3];
            }
            acc += q0.x; acc += q0.y; acc += q0.z; acc += q0.w;
            acc += q1.x; acc += q1.y; acc += q1.z; acc += q1.w;
            acc += q2.x; acc += q2.y; acc += q2.z; acc += q2.w;
            acc += q3.x; acc += q3.y; acc += q3.z; acc += q3.w;
            q0 = n0;
            q1 = n1;
            q2 = n2;
            q3 = n3;
        }
    }
    for (; k < cnt; k++) acc += x[k];
    return acc;
}

// Per-lane partial statistics of one pass (reduced by the caller).
struct PassStats {
    unsigned long long w[2];   // weights per side
    double tt[2];              // sums of w |v|^2 per side (integers: exact in any order)
    uint32_t cnt[2];           // vectors per side
};

// One pass over a node's list in order, 64 vectors per batch: every lane
// takes one vector, decides its side (side_of: 0 or 1) and stages v * w in
// the array of its side (0 in the other); lane k < 2D then adds the batch's
// terms of (side k / D, component k % D) one after another, so the float sums
// keep the list order.  The next batch's vectors are loaded while the
// current one is summed.  Returns the sum of lane k's (side, component).
template <int D, class SideF>
__device__ float side_pass(VqLds<D>& L, const VqArgs& A, const uint32_t* __restrict__ list, uint32_t n,
                           SideF side_of, PassStats& st, bool exact)
{
    const uint32_t me = lane();
    float acc = 0.0f;
    st.w[0] = st.w[1] = 0;
    st.tt[0] = st.tt[1] = 0.0;
    st.cnt[0] = st.cnt[1] = 0;
    if (exact) {
        // integer terms, total within 2^24: per-lane sums, then a wave reduction
        float pl[D], pr[D];
#pragma unroll
        for (int c = 0; c < D; c++) pl[c] = pr[c] = 0.0f;
        constexpr uint32_t U = D >= 16 ? 4 : 8;
        for (uint32_t j0 = 0; j0 < n; j0 += 64 * U) {
            uint32_t vi[U], wu[U];
            float vu[U][D];
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                const uint32_t j = j0 + 64 * u + me;
                vi[u] = j < n ? list[j] : list[0];
            }
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                load_vec<D>(A, vi[u], vu[u]);
                wu[u] = A.w[vi[u]];
            }
#pragma unroll
            for (uint32_t u = 0; u < U; u++) {
                if (j0 + 64 * u + me >= n) continue;
                const float (&v)[D] = vu[u];
                const uint32_t w = wu[u];
                const uint32_t sd = side_of(v);
                const float wf = (float)w;
                const double t = vdot<D>(v, v) * wf;
#pragma unroll
                for (int c = 0; c < D; c++) {
                    const float p = v[c] * wf;
                    pl[c] += sd == 0 ? p : 0.0f;
                    pr[c] += sd == 0 ? 0.0f : p;
                }
                if (sd == 0) {
                    st.w[0] += w;
                    st.tt[0] += t;
                    st.cnt[0]++;
                } else {
                    st.w[1] += w;
                    st.tt[1] += t;
                    st.cnt[1]++;
                }
            }
        }
#pragma unroll
        for (int c = 0; c < D; c++) {
            const float a = wave_sum_f(pl[c]), b = wave_sum_f(pr[c]);
            if (me == (uint32_t)c) acc = a;
            if (me == (uint32_t)(D + c)) acc = b;
        }
        return acc;
    }
    constexpr int R = VqLds<D>::kR;
    constexpr uint32_t kB = 64u * R;   // vectors per batch
    float v[R][D];
    uint32_t w[R], vn[R];   // vn: the indices of the batch after the one loaded
#pragma unroll
    for (int r = 0; r < R; r++) {
        w[r] = 0;
        const uint32_t j = (uint32_t)r * 64u + me;
        if (j < n) {
            const uint32_t vi = list[j];
            load_vec<D>(A, vi, v[r]);
            w[r] = A.w[vi];
        }
        const uint32_t jn = kB + j;
        vn[r] = jn < n ? list[jn] : 0u;
    }
    const float* src = me < D ? &L.a[me][0] : &L.b[me < 2 * D ? me - D : 0][0];
    for (uint32_t j0 = 0; j0 < n; j0 += kB) {
#pragma unroll
        for (int r = 0; r < R; r++) {
            const uint32_t k = (uint32_t)r * 64u + me;
            if (j0 + k < n) {
                const uint32_t sd = side_of(v[r]);
                const float wf = (float)w[r];
                const double t = vdot<D>(v[r], v[r]) * wf;
#pragma unroll
                for (int c = 0; c < D; c++) {
                    const float p = v[r][c] * wf;
                    L.a[c][k] = sd == 0 ? p : 0.0f;
                    L.b[c][k] = sd == 0 ? 0.0f : p;
                }
                if (sd == 0) {
                    st.w[0] += w[r];
                    st.tt[0] += t;
                    st.cnt[0]++;
                } else {
                    st.w[1] += w[r];
                    st.tt[1] += t;
                    st.cnt[1]++;
                }
            }
        }
        // the next batch's vectors (their indices are in vn already) and the
        // indices of the batch after it load while this one is summed
#pragma unroll
        for (int r = 0; r < R; r++) {
            const uint32_t jn = j0 + kB + (uint32_t)r * 64u + me;
            if (jn < n) {
                load_vec<D>(A, vn[r], v[r]);
                w[r] = A.w[vn[r]];
            }
            const uint32_t jnn = jn + kB;
            vn[r] = jnn < n ? list[jnn] : 0u;
        }
        __syncthreads();
        const uint32_t cnt = n - j0 < kB ? n - j0 : kB;
        // adding 0.0f to a sum of non-negative terms leaves it unchanged
        if (me < 2 * D) acc = serial_sum(acc, src, cnt);
        __syncthreads();
    }
    return acc;
}

// The D sums of side s as a vector on every lane.
template <int D>
__device__ inline void gather_sums(float acc, uint32_t s, float (&out)[D])
{
#pragma unroll
    for (int c = 0; c < D; c++) out[c] = __shfl(acc, (int)(s * D + c));
}

// acc + x[0] * y[0] + x[1] * y[1] + ..., each product rounded, then added one
// after another; 128-bit loads, 8 values ahead.
__device__ inline float serial_dot(float acc, const float* x, const float* y, uint32_t cnt)
{
    const float4* x4 = reinterpret_cast<const float4*>(x);
    const float4* y4 = reinterpret_cast<const float4*>(y);
    uint32_t k = 0;
    if (cnt >= 8) {
        float4 a0 = x4[0], a1 = x4[1], b0 = y4[0], b1 = y4[1];
        for (; k + 8 <= cnt; k += 8) {
            float4 na0 = a0, na1 = a1, nb0 = b0, nb1 = b1;
            if (k + 16 <= cnt) {
                const uint32_t g = (k + 8) >> 2;
                na0 = x4[g];
                na1 = x4[g + 1];
                nb0 = y4[g];
                nb1 = y4[g + 1];
            }
            const float p0 = a0.x * b0.x, p1 = a0.y * b0.y, p2 = a0.z * b0.z, p3 = a0.w * b0.w;
            const float p4 = a1.x * b1.x, p5 = a1.y * b1.y, p6 = a1.z * b1.z, p7 = a1.w * b1.w;
            acc = acc + p0; acc = acc + p1; acc = acc + p2; acc = acc + p3;
            acc = acc + p4; acc = acc + p5; acc = acc + p6; acc = acc + p7;
            a0 = na0;
            a1 = na1;
            b0 = nb0;
            b1 = nb1;
        }
    }
    for (; k < cnt; k++) acc = acc + x[k] * y[k];
    return acc;
}

// Weighted covariance about cen, entries in list order (crn_clusterizer.h:
// 494-509, crn_threaded_clusterizer.h:266-279); then scaled as mode asks,
// mirrored into L.cov.  mode 0: float one_over (clusterizer), 1: double
// one_over, result rounded to float (threaded clusterizer).  Lane e holds
// entries e, e + 64 and e + 128 of the upper triangle; batches of 64
// vectors as in side_pass, the next one loaded during the sums.
template <int D>
__device__ void covariance(VqLds<D>& L, const VqArgs& A, const uint32_t* __restrict__ list, uint32_t n,
                           const float (&cen)[D], float oo_f, double oo_d, int mode)
{
    constexpr int kE = VqLds<D>::kE;
    constexpr int W = kVqWaves<D>;
    const uint32_t me = lane(), wv = threadIdx.x >> 6;
    int ex[3], ey[3];
    float acc[3] = {0.0f, 0.0f, 0.0f};
    // entry e = lane + 64 (wave + W r): with W waves the rounds per wave drop
    // to ceil(kE / (64 W))
#pragma unroll
    for (int r = 0; r < 3; r++) {
        int e = (int)me + 64 * ((int)wv + W * r), x = 0;
        if (e < kE) {
            int row = D;
            while (e >= row) {
                e -= row;
                row--;
                x++;
            }
            ex[r] = x;
            ey[r] = x + e;
        } else {
            ex[r] = ey[r] = -1;
        }
    }
    constexpr int R = VqLds<D>::kR;
    constexpr uint32_t kB = 64u * R;
    float v[R][D];
    uint32_t w[R], vn[R];   // vn: the indices of the batch after the one loaded
#pragma unroll
    for (int r = 0; r < R; r++) {
        w[r] = 0;
        const uint32_t j = (uint32_t)r * 64u + me;
        if (j < n) {
            const uint32_t vi = list[j];
            load_vec<D>(A, vi, v[r]);
            w[r] = A.w[vi];
        }
        const uint32_t jn = kB + j;
        vn[r] = jn < n ? list[jn] : 0u;
    }
    for (uint32_t j0 = 0; j0 < n; j0 += kB) {
#pragma unroll
        for (int r = 0; r < R; r++) {
            const uint32_t k = (uint32_t)r * 64u + me;
            if (j0 + k < n) {
                const float wf = (float)w[r];
#pragma unroll
                for (int c = 0; c < D; c++) {
                    const float d = v[r][c] - cen[c];
                    L.a[c][k] = d;
                    L.b[c][k] = d * wf;
                }
            }
        }
#pragma unroll
        for (int r = 0; r < R; r++) {
            const uint32_t jn = j0 + kB + (uint32_t)r * 64u + me;
            if (jn < n) {
                load_vec<D>(A, vn[r], v[r]);
                w[r] = A.w[vn[r]];
            }
            const uint32_t jnn = jn + kB;
            vn[r] = jnn < n ? list[jnn] : 0u;
        }
        __syncthreads();
        const uint32_t cnt = n - j0 < kB ? n - j0 : kB;
#pragma unroll
        for (int r = 0; r < 3; r++)
            if (ex[r] >= 0) acc[r] = serial_dot(acc[r], &L.a[ex[r]][0], &L.b[ey[r]][0], cnt);
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 3; r++)
        if (ex[r] >= 0) {
            const float c = mode ? (float)(acc[r] * oo_d) : acc[r] * oo_f;
            L.cov[ex[r]][ey[r]] = c;
            L.cov[ey[r]][ex[r]] = c;
        }
    __syncthreads();
}

// Power iteration from lerp(.75, 1.25) and normalisation (crn_clusterizer.h:
// 523-566, crn_threaded_clusterizer.h:288-339; vec::norm is the sum of
// squares, vec::normalize crn_vec.h:674-690).  Result on every lane.
template <int D>
__device__ void pca_axis(VqLds<D>& L, float (&axis)[D])
{
    // every wave of the node runs the iteration; only wave 0 stores (the
    // last step reads L.axis[k] before storing it, which another wave's store
    // would overtake)
    const uint32_t me = lane();
    const bool owner = me < D && threadIdx.x < 64;
    if (owner) {
        const float t = (float)me * (1.0f / (float)(D - 1));
        L.axis[me] = .75f + (1.25f - .75f) * t;
        L.prev[me] = L.axis[me];
    }
    __syncthreads();
    for (int iter = 0; iter < 10; iter++) {
        double sum = 0;
        if (me < D)
            for (int j = 0; j < D; j++) sum += L.axis[j] * L.cov[me][j];
        const double max_sum = wave_max_d(me < D ? fabs(sum) : 0.0);
        float x = (float)sum;
        if (max_sum != 0.0f) x *= (float)(1.0f / max_sum);
        __syncthreads();
        if (owner) L.x[me] = L.prev[me] - x;   // delta
        __syncthreads();
        float nrm = L.x[0] * L.x[0];
        for (int i = 1; i < D; i++) nrm += L.x[i] * L.x[i];
        __syncthreads();
        if (owner) {
            L.prev[me] = L.axis[me];
            L.axis[me] = x;
        }
        __syncthreads();
        if (nrm < .0025f) break;
    }
#pragma unroll
    for (int i = 0; i < D; i++) axis[i] = L.axis[i];
    double nn = axis[0] * axis[0];
    for (int i = 1; i < D; i++) nn += axis[i] * axis[i];
    if (nn != 0) {
        const float s = (float)(1.0f / sqrt(nn));
#pragma unroll
        for (int i = 0; i < D; i++) axis[i] *= s;
    }
    __syncthreads();
}

// First vector of the list (in order) at the largest distance from ref
// (compute_split_estimate, crn_clusterizer.h:445-481).
template <int D>
__device__ void furthest(const VqArgs& A, const uint32_t* __restrict__ list, uint32_t n, const float (&ref)[D],
                         float (&out)[D])
{
    const uint32_t me = lane();
    float bd = -1.0f;
    uint32_t bj = 0xFFFFFFFFu;
    for (uint32_t j = me; j < n; j += 64) {
        float v[D];
        load_vec<D>(A, list[j], v);
        const float d = vsqd<D>(v, ref);
        if (d > bd) {
            bd = d;
            bj = j;
        }
    }
    for (int o = 32; o; o >>= 1) {
        const float od = __shfl_xor(bd, o);
        const uint32_t oj = (uint32_t)__shfl_xor((int)bj, o);
        if (od > bd || (od == bd && oj < bj)) {
            bd = od;
            bj = oj;
        }
    }
    if (bj == 0xFFFFFFFFu) {
#pragma unroll
        for (int c = 0; c < D; c++) out[c] = 0.0f;
    } else {
        load_vec<D>(A, list[bj], out);
    }
}

// split_node (crn_clusterizer.h:722-871) of one node per wave.
template <int D>
__global__ void __launch_bounds__(192) vq_split_kernel(VqArgs A, const VqSplit* __restrict__ splits,
                                                      VqResult* __restrict__ results)
{
    __shared__ VqLds<D> L;
    const VqSplit sp = splits[blockIdx.x];
    const uint32_t* list = A.pool + sp.seg;
    const uint32_t n = sp.n, me = lane();
    float cen[D];
#pragma unroll
    for (int c = 0; c < D; c++) cen[c] = A.cen[(size_t)sp.node * D + c];
    const unsigned long long ptw = A.tw[sp.node];
    const bool exact = exact_sums(A, ptw);
    float lc[D], rc[D];
    if (n == 2) {
        load_vec<D>(A, list[0], lc);
        load_vec<D>(A, list[1], rc);
    } else {
        // compute_split_pca (:483-601)
        covariance<D>(L, A, list, n, cen, 1.0f / (float)ptw, 0.0, 0);
        float axis[D];
        pca_axis<D>(L, axis);
        PassStats ps;
        const float acc = side_pass<D>(L, A, list, n, [&](const float (&v)[D]) {
            float d[D];
#pragma unroll
            for (int c = 0; c < D; c++) d[c] = v[c] - cen[c];
            const double t = vdot<D>(d, axis);
            return t < 0.0f ? 0u : 1u;
        }, ps, exact);
        const double lwp = (double)wave_sum_u64(ps.w[0]), rwp = (double)wave_sum_u64(ps.w[1]);
        if (lwp > 0.0f && rwp > 0.0f) {
            float sl[D], sr[D];
            gather_sums<D>(acc, 0, sl);
            gather_sums<D>(acc, 1, sr);
            const float fl = (float)(1.0f / lwp), fr = (float)(1.0f / rwp);
#pragma unroll
            for (int c = 0; c < D; c++) {
                lc[c] = sl[c] * fl;
                rc[c] = sr[c] * fr;
            }
        } else {
            float fur[D], opp[D];
            furthest<D>(A, list, n, cen, fur);
            furthest<D>(A, list, n, fur, opp);
#pragma unroll
            for (int c = 0; c < D; c++) {
                lc[c] = (fur[c] + cen[c]) * .5f;
                rc[c] = (opp[c] + cen[c]) * .5f;
            }
        }
    }
    uint32_t npass = n == 2 ? 0u : 2u;
    // 2-means (:746-832)
    float plc[D], prc[D];
    unsigned long long lw = 0, rw = 0;
    float prev_total = 1e+10f, lvar = 0.0f, rvar = 0.0f;
    uint32_t nl = 0;
    bool uns = false;
    for (int loop = 0; loop < 8; loop++) {
        npass++;
#pragma unroll
        for (int c = 0; c < D; c++) {
            plc[c] = lc[c];
            prc[c] = rc[c];
        }
        PassStats ps;
        const float acc = side_pass<D>(L, A, list, n, [&](const float (&v)[D]) {
            const double ld = vsqd<D>(plc, v), rd = vsqd<D>(prc, v);
            return ld < rd ? 0u : 1u;
        }, ps, exact);
        lw = wave_sum_u64(ps.w[0]);
        rw = wave_sum_u64(ps.w[1]);
        const double lt = wave_sum_d(ps.tt[0]), rt = wave_sum_d(ps.tt[1]);
        nl = (uint32_t)wave_sum_u64(ps.cnt[0]);
        if (!lw || !rw) {
            uns = true;
            break;
        }
        float sl[D], sr[D];
        gather_sums<D>(acc, 0, sl);
        gather_sums<D>(acc, 1, sr);
        lvar = (float)(lt - (vdot<D>(sl, sl) / (float)lw));
        rvar = (float)(rt - (vdot<D>(sr, sr) / (float)rw));
        const float fl = 1.0f / (float)lw, fr = 1.0f / (float)rw;
#pragma unroll
        for (int c = 0; c < D; c++) {
            lc[c] = sl[c] * fl;
            rc[c] = sr[c] * fr;
        }
        const float total = lvar + rvar;
        if (total < .00001f) break;
        if (((prev_total - total) / total) < .00125f) break;
        prev_total = total;
    }
    VqResult r;
    if (uns) {
        r.split = 0;
        r.nl = r.nr = 0;
        r.lvar = r.rvar = 0.0f;
        if (me == 0) {
            results[blockIdx.x] = r;
            if (A.passes && threadIdx.x == 0) atomicAdd(A.passes, (unsigned long long)n * npass);
        }
        return;
    }
    // the final round's sides (from the centres before its update), stable
    list_partition<D>(A, list, n, nl, A.pool + sp.child_seg, [&](const float (&v)[D]) {
        return (double)vsqd<D>(plc, v) < (double)vsqd<D>(prc, v);
    });
    if (me < D) {
        A.cen[(size_t)sp.left * D + me] = lc[me];
        A.cen[(size_t)sp.right * D + me] = rc[me];
    }
    if (me == 0) {
        if (A.passes && threadIdx.x == 0) atomicAdd(A.passes, (unsigned long long)n * (npass + 1));
        A.tw[sp.left] = lw;
        A.tw[sp.right] = rw;
        r.split = 1;
        r.nl = nl;
        r.nr = n - nl;
        r.lvar = lvar;
        r.rvar = rvar;
        results[blockIdx.x] = r;
    }
}

// The root of a tree (crn_clusterizer.h:77-96): centroid, weight, variance
// over the list in order.  One wave.
template <int D>
__global__ void __launch_bounds__(64) vq_root_kernel(VqArgs A, uint32_t seg, uint32_t n, uint32_t slot,
                                                     float* __restrict__ var_out)
{
    __shared__ VqLds<D> L;
    const uint32_t* list = A.pool + seg;
    const uint32_t me = lane();
    const unsigned long long tw0 = list_weight(A, list, n);
    PassStats ps;
    const float acc = side_pass<D>(L, A, list, n, [&](const float (&)[D]) { return 0u; }, ps, exact_sums(A, tw0));
    const unsigned long long tw = wave_sum_u64(ps.w[0]);
    const double tt = wave_sum_d(ps.tt[0]);
    float sum[D];
    gather_sums<D>(acc, 0, sum);
    const float var = (float)(tt - (vdot<D>(sum, sum) / (float)tw));
    const float f = 1.0f / (float)tw;
    if (me < D) A.cen[(size_t)slot * D + me] = sum[me] * f;
    if (me == 0) {
        A.tw[slot] = tw;
        *var_out = var;
    }
}

// threaded_clusterizer::compute_split (crn_threaded_clusterizer.h:237-379):
// PCA of the listed vectors and the division by the sign of the projection,
// left part then right part to pool[out_seg, out_seg + n).  One wave per job.
struct DivJob {
    uint32_t seg, n, out_seg;
};

template <int D>
__global__ void __launch_bounds__(192) vq_division_kernel(VqArgs A, const DivJob* __restrict__ jobs,
                                                         uint32_t* __restrict__ nleft)
{
    __shared__ VqLds<D> L;
    const DivJob jb = jobs[blockIdx.x];
    const uint32_t* list = A.pool + jb.seg;
    const uint32_t n = jb.n, me = lane();
    const unsigned long long tw0 = list_weight(A, list, n);
    PassStats ps;
    const float acc = side_pass<D>(L, A, list, n, [&](const float (&)[D]) { return 0u; }, ps, exact_sums(A, tw0));
    const double tw = (double)wave_sum_u64(ps.w[0]);
    float cen[D], axis[D];
    gather_sums<D>(acc, 0, cen);
    if (tw == 0.0f) {
#pragma unroll
        for (int c = 0; c < D; c++) axis[c] = 0.0f;
    } else {
        const double oo = 1.0f / tw;
#pragma unroll
        for (int c = 0; c < D; c++) cen[c] = (float)(cen[c] * oo);
        covariance<D>(L, A, list, n, cen, 0.0f, oo, 1);
        pca_axis<D>(L, axis);
    }
    auto is_left = [&](const float (&v)[D]) {
        float d[D];
#pragma unroll
        for (int c = 0; c < D; c++) d[c] = v[c] - cen[c];
        return vdot<D>(d, axis) < 0.0f;
    };
    const uint32_t nl = list_count<D>(A, list, n, is_left);
    list_partition<D>(A, list, n, nl, A.pool + jb.out_seg, is_left);
    if (me == 0) nleft[blockIdx.x] = nl;
}

// Clusters as contiguous lists: cluster k's segment copied to idx[ofs[k], ...).
struct SegCopy {
    uint32_t seg, n, dst;
};

__global__ void __launch_bounds__(256) seg_copy_kernel(const uint32_t* __restrict__ pool,
                                                       const SegCopy* __restrict__ segs, uint32_t* __restrict__ idx)
{
    const SegCopy s = segs[blockIdx.x];
    for (uint32_t j = threadIdx.x; j < s.n; j += blockDim.x) idx[s.dst + j] = pool[s.seg + j];
}

// ------------------------------------------------------------ pack kernels
// Texels of the listed blocks, 16 per block, in list order.
__global__ void __launch_bounds__(256) gather_texels_kernel(const uint32_t* __restrict__ blocks,
                                                            const uint32_t* __restrict__ idx, uint32_t n,
                                                            uint32_t* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * 16) return;
    out[i] = blocks[(size_t)idx[i >> 4] * 16 + (i & 15)];
}

struct Dxt1Rec {   // crn2_dxt1_blocks_device / crn2_dxt1_clusters_device records
    unsigned long long error;
    uint32_t lo_hi, selectors_or_unique, flags, pad;
};

// Per block of cluster c: the cluster's endpoints and the block's 16
// selectors from the per-texel selectors (crn_qdxt1.cpp:633-650).
__global__ void __launch_bounds__(256) cluster_blocks_kernel(const uint32_t* __restrict__ idx,
                                                             const uint32_t* __restrict__ block_cluster,
                                                             uint32_t n, const Dxt1Rec* __restrict__ rec,
                                                             const uint8_t* __restrict__ sel,
                                                             unsigned long long* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t word = 0;
    for (int t = 0; t < 16; t++) word |= (uint32_t)(sel[(size_t)i * 16 + t] & 3u) << (2 * t);
    out[idx[i]] = (unsigned long long)rec[block_cluster[i]].lo_hi | ((unsigned long long)word << 32);
}

__global__ void __launch_bounds__(256) single_blocks_kernel(const Dxt1Rec* __restrict__ rec, uint32_t n,
                                                            unsigned long long* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    out[i] = (unsigned long long)rec[i].lo_hi | ((unsigned long long)rec[i].selectors_or_unique << 32);
}

__device__ inline uint32_t cdist(int perceptual, int r0, int g0, int b0, int r1, int g1, int b1)
{
    const int dr = r0 - r1, dg = g0 - g1, db = b0 - b1;
    return perceptual ? (uint32_t)(dr * dr * 8 + dg * dg * 25 + db * db) : (uint32_t)(dr * dr + dg * dg + db * db);
}

// Selector vectors and weights (crn_qdxt1.cpp:843-874).
__global__ void __launch_bounds__(256) selector_vec_kernel(const unsigned long long* __restrict__ blk, uint32_t n,
                                                           int perceptual, float* __restrict__ v,
                                                           uint32_t* __restrict__ w)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const unsigned long long x = blk[b];
    for (int t = 0; t < 16; t++) v[(size_t)b * 16 + t] = c_linear[(x >> (32 + 2 * t)) & 3];
    int c0[3], c1[3];
    unpack565((uint32_t)(x & 0xFFFF), c0);
    unpack565((uint32_t)((x >> 16) & 0xFFFF), c1);
    const uint32_t d = cdist(perceptual, c0[0], c0[1], c0[2], c1[0], c1[1], c1[2]) / 2000;
    w[b] = d < 1 ? 1 : (d > 2048 ? 2048 : d);
}

// optimize_selectors_task (crn_qdxt1.cpp:714-840), one workgroup per
// selector cluster: 16 block slots x 16 texels, each thread summing the
// errors of all four selector values of its texel over its slot's blocks,
// then the slots' sums added per (type, texel, value).  The sums are exact
// integers, so their order does not matter.
constexpr uint32_t kSelSlots = 16;

__global__ void __launch_bounds__(256) selector_opt_kernel(const uint32_t* __restrict__ blocks,
                                                           const uint32_t* __restrict__ idx,
                                                           const uint32_t* __restrict__ ofs,
                                                           unsigned long long* __restrict__ blk, int perceptual,
                                                           uint32_t thresh)
{
    __shared__ unsigned long long part[kSelSlots][2][16][4];
    __shared__ uint32_t pcnt[kSelSlots][2];
    __shared__ uint32_t words[2];
    __shared__ uint32_t counts[2];
    const uint32_t c = blockIdx.x, tid = threadIdx.x;
    const uint32_t b0 = ofs[c], m = ofs[c + 1] - b0;
    if (m <= 1) return;
    const uint32_t slot = tid >> 4, t = tid & 15;
    unsigned long long err[2][4] = {{0, 0, 0, 0}, {0, 0, 0, 0}};
    uint32_t cnt[2] = {0, 0};
#pragma unroll 2
    for (uint32_t j = slot; j < m; j += kSelSlots) {
        const uint32_t b = idx[b0 + j];
        const unsigned long long x = blk[b];
        const uint32_t p = blocks[(size_t)b * 16 + t];
        const uint32_t lo = (uint32_t)(x & 0xFFFF), hi = (uint32_t)((x >> 16) & 0xFFFF);
        // a 3-colour block with a transparent texel keeps its selectors; the
        // slot's 16 lanes are 16 consecutive lanes of one wave
        const unsigned long long ta = __ballot(thresh > 0 && (p >> 24) < thresh);
        const bool transparent = ((ta >> (lane() & ~15u)) & 0xFFFFull) != 0;
        if (lo <= hi && transparent) continue;
        const int typ = lo <= hi ? 1 : 0;
        int col[4][3];
        block_colors(lo, hi, col);
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const unsigned long long e =
                cdist(perceptual, ch(p, 0), ch(p, 1), ch(p, 2), col[v][0], col[v][1], col[v][2]);
            err[0][v] += typ ? 0ull : e;
            err[1][v] += typ ? e : 0ull;
        }
        cnt[typ]++;
    }
#pragma unroll
    for (int v = 0; v < 4; v++) {
        part[slot][0][t][v] = err[0][v];
        part[slot][1][t][v] = err[1][v];
    }
    if (t == 0) {
        pcnt[slot][0] = cnt[0];
        pcnt[slot][1] = cnt[1];
    }
    __syncthreads();
    if (tid < 32) {
        // the first value of least error among the allowed ones (ascending)
        const uint32_t typ = tid >> 4, tt = tid & 15;
        uint32_t best_v = 0;
        unsigned long long best_e = 0xFFFFFFFFFFull;
        for (uint32_t v = 0; v < (typ ? 3u : 4u); v++) {
            unsigned long long e = 0;
            for (uint32_t k = 0; k < kSelSlots; k++) e += part[k][typ][tt][v];
            if (e < best_e) {
                best_e = e;
                best_v = v;
            }
        }
        uint32_t n = 0;
        for (uint32_t k = 0; k < kSelSlots; k++) n += pcnt[k][typ];
        // the 16 texels of a type are 16 consecutive lanes of wave 0
        uint32_t word = best_v << (2 * tt);
        for (uint32_t o = 1; o < 16; o <<= 1) word |= (uint32_t)__shfl_xor((int)word, (int)o);
        if (tt == 0) {
            words[typ] = word;
            counts[typ] = n;
        }
    }
    __syncthreads();
    for (uint32_t j = tid; j < m; j += blockDim.x) {
        const uint32_t b = idx[b0 + j];
        const unsigned long long x = blk[b];
        const uint32_t lo = (uint32_t)(x & 0xFFFF), hi = (uint32_t)((x >> 16) & 0xFFFF);
        const uint32_t typ = lo <= hi ? 1u : 0u;
        if (counts[typ] <= 1) continue;
        bool mine = true;
        if (typ == 1 && thresh > 0)
            for (int k = 0; k < 16; k++) mine &= (blocks[(size_t)b * 16 + k] >> 24) >= thresh;
        if (mine) blk[b] = (x & 0xFFFFFFFFull) | ((unsigned long long)words[typ] << 32);
    }
}

// ------------------------------------------------------------------- qdxt5
// dxt5_block::get_block_values (crn_dxt.cpp:395-455)
__device__ inline void alpha_values(uint32_t l, uint32_t h, uint32_t (&v)[8])
{
    v[0] = l;
    v[1] = h;
    if (l > h) {
        for (uint32_t j = 2; j < 8; j++) v[j] = (l * (8 - j) + h * (j - 1)) / 7;
    } else {
        for (uint32_t j = 2; j < 6; j++) v[j] = (l * (6 - j) + h * (j - 1)) / 5;
        v[6] = 0;
        v[7] = 255;
    }
}

// dxt_fast::compress_alpha_block (crn_dxt_fast.cpp:788-826) of N values
template <int N>
__device__ inline void fast_alpha(const uint32_t (&a)[N], uint32_t& low8, uint32_t& high8, uint8_t (&sel)[N])
{
    int mn = (int)a[0], mx = (int)a[0];
#pragma unroll
    for (int i = 1; i < N; i++) {
        mn = (int)a[i] < mn ? (int)a[i] : mn;
        mx = (int)a[i] > mx ? (int)a[i] : mx;
    }
    low8 = (uint32_t)mx;
    high8 = (uint32_t)mn;
    const int dist = mx - mn, bias = mn * 7 - (dist >> 1), dist4 = dist * 4, dist2 = dist * 2;
#pragma unroll
    for (int i = 0; i < N; i++) {
        int v = (int)a[i] * 7 - bias, ind, t;
        t = (dist4 - v) >> 31;
        ind = t & 4;
        v -= dist4 & t;
        t = (dist2 - v) >> 31;
        ind += t & 2;
        v -= dist2 & t;
        t = (dist - v) >> 31;
        ind += t & 1;
        ind = -ind & 7;
        ind ^= (2 > ind);
        sel[i] = (uint8_t)ind;
    }
}

template <int N>
__device__ uint32_t alpha_layout_error(const uint32_t* __restrict__ blocks, const TChunk& c, uint32_t l, uint32_t comp)
{
    const uint32_t xo = c_layout[l][0], yo = c_layout[l][1], w = c_layout[l][2];
    uint32_t a[N];
#pragma unroll
    for (int i = 0; i < N; i++) a[i] = (uint32_t)ch(chunk_texel(blocks, c, xo + (uint32_t)i % w, yo + (uint32_t)i / w),
                                                   (int)comp);
    uint8_t sel[N];
    uint32_t lo, hi, v[8];
    fast_alpha<N>(a, lo, hi, sel);
    alpha_values(lo, hi, v);
    uint32_t err = 0;
#pragma unroll
    for (int i = 0; i < N; i++) {
        const int d = (int)a[i] - (int)v[sel[i]];
        err += (uint32_t)(d * d);
    }
    return err;
}

__global__ void __launch_bounds__(256) qdxt5_layout_kernel(const uint32_t* __restrict__ blocks,
                                                           const TChunk* __restrict__ chunks, uint32_t nchunks,
                                                           uint32_t comp, uint32_t* __restrict__ lerr)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nchunks * 9) return;
    const TChunk c = chunks[i / 9];
    const uint32_t l = i % 9;
    lerr[i] = l == 0 ? alpha_layout_error<64>(blocks, c, l, comp)
                     : (l < 5 ? alpha_layout_error<32>(blocks, c, l, comp) : alpha_layout_error<16>(blocks, c, l, comp));
}

__device__ inline void put_alpha_training(float* __restrict__ vecs, uint32_t* __restrict__ weights, uint32_t b,
                                          uint32_t lo, uint32_t hi)
{
    const int l0 = ch(lo, 0), h0 = ch(hi, 0);
    vecs[2 * (size_t)b] = (float)l0;
    vecs[2 * (size_t)b + 1] = (float)h0;
    const uint32_t w = (uint32_t)((l0 - h0) * (l0 - h0)) / 8;
    weights[b] = w < 1 ? 1 : (w > 8 ? 8 : w);
}

// qdxt5::init's encoding choice (crn_qdxt5.cpp:196-246) and training vectors
// (:248-318): tiles of grey texels (the channel in r, g and b).
__global__ void __launch_bounds__(256) qdxt5_choose_kernel(const uint32_t* __restrict__ blocks,
                                                           const TChunk* __restrict__ chunks, uint32_t nchunks,
                                                           uint32_t comp, const uint32_t* __restrict__ lerr,
                                                           float* __restrict__ vecs, uint32_t* __restrict__ weights,
                                                           uint8_t* __restrict__ enc)
{
    const uint32_t ci = blockIdx.x * blockDim.x + threadIdx.x;
    if (ci >= nchunks) return;
    const TChunk c = chunks[ci];
    double best = -1.0f;
    uint32_t be = 0;
    for (uint32_t e = 0; e < 8; e++) {
        double total = 0;
        for (uint32_t t = 0; t < c_enc[e][0]; t++) total += (double)lerr[ci * 9 + c_enc[e][1 + t]];
        const double ms = total * (1.0f / 64.0f);
        const double rms = sqrt(ms);
        double psnr = 999999.0f;
        if (ms) {
            psnr = log10(255.0f / rms) * 20.0f;
            psnr = psnr < 0.0 ? 0.0 : (psnr > 500.0 ? 500.0 : psnr);
        }
        const float ader = 0.0f + (c.der - 0.0f) * ((float)(c_enc[e][0] - 1) / 3.0f);
        psnr = psnr - ader;
        if (psnr > best) {
            best = psnr;
            be = e;
        }
    }
    for (uint32_t t = 0; t < c_enc[be][0]; t++) {
        const uint32_t l = c_enc[be][1 + t];
        const uint32_t xo = c_layout[l][0], yo = c_layout[l][1], w = c_layout[l][2], h = c_layout[l][3];
        uint32_t lo, hi;
        rep_colors(w * h,
                   [&](uint32_t i) {
                       const uint32_t a = (uint32_t)ch(chunk_texel(blocks, c, xo + i % w, yo + i / w), (int)comp);
                       return a | a << 8 | a << 16 | 0xFF000000u;
                   },
                   lo, hi);
        for (uint32_t y = 0; y < (h >> 2); y++) {
            const uint32_t by = c.cy * 2 + y + (yo >> 2);
            if (by >= c.bh) continue;
            for (uint32_t x = 0; x < (w >> 2); x++) {
                const uint32_t bx = c.cx * 2 + x + (xo >> 2);
                if (bx >= c.bw) break;
                const uint32_t b = c.first + bx + by * c.bw;
                put_alpha_training(vecs, weights, b, lo, hi);
                if (enc) enc[b] = (uint8_t)be;
            }
        }
    }
}

__global__ void __launch_bounds__(256) qdxt5_block_train_kernel(const uint32_t* __restrict__ blocks, uint32_t n,
                                                                uint32_t comp, float* __restrict__ vecs,
                                                                uint32_t* __restrict__ weights,
                                                                uint8_t* __restrict__ enc)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const uint32_t* p = blocks + (size_t)b * 16;
    uint32_t lo, hi;
    rep_colors(16,
               [&](uint32_t i) {
                   const uint32_t a = (uint32_t)ch(p[i], (int)comp);
                   return a | a << 8 | a << 16 | 0xFF000000u;
               },
               lo, hi);
    put_alpha_training(vecs, weights, b, lo, hi);
    if (enc) enc[b] = 255;
}

__global__ void __launch_bounds__(256) qdxt5_selword_kernel(const uint32_t* __restrict__ blocks, uint32_t n,
                                                            uint32_t comp, unsigned long long* __restrict__ table,
                                                            uint32_t mask, uint32_t* __restrict__ count)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    uint32_t a[16];
#pragma unroll
    for (int i = 0; i < 16; i++) a[i] = (uint32_t)ch(blocks[(size_t)b * 16 + i], (int)comp);
    uint8_t sel[16];
    uint32_t lo, hi;
    fast_alpha<16>(a, lo, hi, sel);
    unsigned long long word = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) word |= (unsigned long long)sel[i] << (3 * i);
    const unsigned long long key = word + 1ull;
    uint32_t h = hash32((uint32_t)word ^ hash32((uint32_t)(word >> 32))) & mask;
    for (;;) {
        const unsigned long long prev = atomicCAS(&table[h], 0ull, key);
        if (prev == 0ull) {
            atomicAdd(count, 1u);
            break;
        }
        if (prev == key) break;
        h = (h + 1) & mask;
    }
}

// Channel values of the listed blocks, 16 per block, in list order.
__global__ void __launch_bounds__(256) gather_values_kernel(const uint32_t* __restrict__ blocks,
                                                            const uint32_t* __restrict__ idx, uint32_t n, uint32_t comp,
                                                            uint8_t* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * 16) return;
    out[i] = (uint8_t)ch(blocks[(size_t)idx[i >> 4] * 16 + (i & 15)], (int)comp);
}

__constant__ uint8_t c_six_invert[8] = {1, 0, 5, 4, 3, 2, 6, 7};
__constant__ uint8_t c_eight_invert[8] = {1, 0, 7, 6, 5, 4, 3, 2};

struct E2Lds {
    uint32_t count[256], first[256], uval[256], uw[256], rank[256];
    uint8_t trial[256], best[256];
};

// dxt5_endpoint_optimizer::compute (crnlib/crn_dxt5a.cpp:40-262) over one
// cluster's values, one wave: distinct values in first-appearance order with
// their counts, every distinct pair, then (better and uber, error left) the
// +-8 / +-16 box that follows the best; each candidate scored by all lanes
// (first minimum per value, exact wave sum) and kept on a strict improvement;
// both palettes when use_both; the canonical order with the invert tables.
__global__ void __launch_bounds__(64) qdxt5_search_kernel(const uint8_t* __restrict__ vals_all,
                                                          const uint32_t* __restrict__ ofs, int quality, int use_both,
                                                          uint32_t* __restrict__ res, uint8_t* __restrict__ sel_all)
{
    __shared__ E2Lds L;
    const uint32_t c = blockIdx.x, me = lane();
    const uint32_t o = ofs[c], n = ofs[c + 1] - o;
    if (!n) return;
    const uint8_t* vals = vals_all + o;
    for (uint32_t v = me; v < 256; v += 64) {
        L.count[v] = 0;
        L.first[v] = 0xFFFFFFFFu;
    }
    __syncthreads();
    for (uint32_t i = me; i < n; i += 64) {
        atomicAdd(&L.count[vals[i]], 1u);
        atomicMin(&L.first[vals[i]], i);
    }
    __syncthreads();
    for (uint32_t v = me; v < 256; v += 64) {
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
    if (nu == 1) {
        if (me == 0) res[c] = L.uval[0] | L.uval[0] << 8;
        for (uint32_t i = me; i < n; i += 64) sel_all[o + i] = 0;
        return;
    }
    unsigned long long best_err = ~0ull;
    uint32_t bf = 0, bs = 0, bt = 0;
    bool done = false;
    auto evaluate = [&](uint32_t lo, uint32_t hi) {
        for (uint32_t type = 0; type < (use_both ? 2u : 1u) && !done; type++) {
            uint32_t pal[8];
            pal[0] = lo;
            pal[1] = hi;
            if (!type) {
                for (uint32_t j = 2; j < 8; j++) pal[j] = (lo * (8 - j) + hi * (j - 1)) / 7;
            } else {
                for (uint32_t j = 2; j < 6; j++) pal[j] = (lo * (6 - j) + hi * (j - 1)) / 5;
                pal[6] = 0;
                pal[7] = 255;
            }
            unsigned long long part = 0;
            for (uint32_t i = me; i < nu; i += 64) {
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
            const unsigned long long err = wave_sum_u64(part);
            if (err < best_err) {
                best_err = err;
                bf = lo;
                bs = hi;
                bt = type;
                for (uint32_t i = me; i < nu; i += 64) L.best[i] = L.trial[i];
                if (!err) done = true;
            }
        }
    };
    for (uint32_t i = 0; i + 1 < nu && !done; i++)
        for (uint32_t j = i + 1; j < nu && !done; j++) evaluate(L.uval[i], L.uval[j]);
    if (quality >= 3 && best_err && !done) {
        const int probe = quality == 4 ? 16 : 8;
        for (int dl = -probe; dl <= probe && !done; dl++) {
            const int l = (int)bf + dl;
            if (l < 0) continue;
            if (l > 255) break;
            for (int dh = -probe; dh <= probe && !done; dh++) {
                const int h = (int)bs + dh;
                if (h < 0) continue;
                if (h > 255) break;
                evaluate((uint32_t)l, (uint32_t)h);
            }
        }
    }
    uint32_t f = bf, sc = bs;
    int inv = 0;   // 0 none, 1 six, 2 eight
    if (f != sc) {
        if (bt) {
            if (f > sc) {
                f = bs;
                sc = bf;
                inv = 1;
            }
        } else if (f <= sc) {
            f = bs;
            sc = bf;
            inv = 2;
        }
    }
    __syncthreads();
    for (uint32_t i = me; i < n; i += 64) {
        uint32_t s = bf == bs ? 0u : L.best[L.rank[vals[i]]];
        if (inv == 1) s = c_six_invert[s];
        else if (inv == 2) s = c_eight_invert[s];
        sel_all[o + i] = (uint8_t)s;
    }
    if (me == 0) res[c] = f | sc << 8;
}

// dxt_fast::compress_alpha_block (crn_dxt_fast.cpp:788-826) over one
// cluster's values, qdxt5's superfast endpoint pass (crn_qdxt5.cpp:532-541):
// low = the largest value, high = the smallest, each selector from the
// value's position between them.  One wave per cluster.
__global__ void __launch_bounds__(64) alpha_fast_kernel(const uint8_t* __restrict__ vals,
                                                        const uint32_t* __restrict__ vofs, uint32_t* __restrict__ res,
                                                        uint8_t* __restrict__ sel_all)
{
    const uint32_t c = blockIdx.x, me = lane();
    const uint32_t o = vofs[c], n = vofs[c + 1] - o;
    int mn = 255, mx = 0;
    for (uint32_t j = me; j < n; j += 64) {
        const int v = vals[o + j];
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
    }
    for (int k = 32; k; k >>= 1) {
        const int omn = __shfl_xor(mn, k), omx = __shfl_xor(mx, k);
        mn = omn < mn ? omn : mn;
        mx = omx > mx ? omx : mx;
    }
    const int dist = mx - mn, bias = mn * 7 - (dist >> 1), dist4 = dist * 4, dist2 = dist * 2;
    for (uint32_t j = me; j < n; j += 64) {
        int a = (int)vals[o + j] * 7 - bias;
        int t = (dist4 - a) >> 31;
        int ind = t & 4;
        a -= dist4 & t;
        t = (dist2 - a) >> 31;
        ind += t & 2;
        a -= dist2 & t;
        t = (dist - a) >> 31;
        ind += t & 1;
        ind = -ind & 7;
        ind ^= (2 > ind) ? 1 : 0;
        sel_all[o + j] = (uint8_t)ind;
    }
    if (me == 0) res[c] = (uint32_t)mx | ((uint32_t)mn << 8);
}

__global__ void __launch_bounds__(256) alpha_blocks_kernel(const uint32_t* __restrict__ idx,
                                                           const uint32_t* __restrict__ block_cluster, uint32_t n,
                                                           const uint32_t* __restrict__ res,
                                                           const uint8_t* __restrict__ sel,
                                                           unsigned long long* __restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    unsigned long long word = 0;
    for (int t = 0; t < 16; t++) word |= (unsigned long long)(sel[(size_t)i * 16 + t] & 7u) << (3 * t);
    out[idx[i]] = (unsigned long long)(res[block_cluster[i]] & 0xFFFFu) | (word << 16);
}

__constant__ float c_d5_linear[8] = {0, 7, 1, 2, 3, 4, 5, 6};
__constant__ float c_d5_linear6[8] = {0, 5, 1, 2, 3, 4, 0, 0};

// Selector vectors of DXT5 alpha blocks (crn_qdxt5.cpp:708-766): kind[b] =
// 0 / 1 for 8- / 6-value blocks, 2 for a 6-value block using the absolute
// values 0 and 255 (left out).
__global__ void __launch_bounds__(256) alpha_selector_vec_kernel(const unsigned long long* __restrict__ blk, uint32_t n,
                                                                 float* __restrict__ v, uint32_t* __restrict__ w,
                                                                 uint8_t* __restrict__ kind)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    const unsigned long long x = blk[b];
    const uint32_t lo = (uint32_t)(x & 255), hi = (uint32_t)((x >> 8) & 255);
    const bool a6 = lo <= hi;
    bool absolute = false;
    for (int t = 0; t < 16; t++) {
        const uint32_t s = (uint32_t)((x >> (16 + 3 * t)) & 7);
        float f;
        if (a6) {
            if (s >= 6) {
                absolute = true;
                f = 0.0f;
            } else {
                f = c_d5_linear6[s];
            }
        } else {
            f = c_d5_linear[s];
        }
        v[(size_t)b * 16 + t] = f;
    }
    const uint32_t d = (uint32_t)(((int)lo - (int)hi) * ((int)lo - (int)hi)) / 8;
    w[b] = d < 1 ? 1 : (d > 2048 ? 2048 : d);
    kind[b] = absolute ? 2 : (a6 ? 1 : 0);
}

// optimize_selectors_task of qdxt5 (crn_qdxt5.cpp:578-700), one workgroup
// per selector cluster: 16 block slots x 16 texels, each thread summing the
// squared errors of all eight values of its texel over its slot's blocks, by
// block type, then the slots' sums added (exact integers).
__global__ void __launch_bounds__(256) alpha_selector_opt_kernel(const uint32_t* __restrict__ blocks,
                                                                 const uint32_t* __restrict__ idx,
                                                                 const uint32_t* __restrict__ ofs,
                                                                 unsigned long long* __restrict__ blk, uint32_t comp)
{
    __shared__ unsigned long long part[kSelSlots][2][16][8];
    __shared__ uint32_t pcnt[kSelSlots][2];
    __shared__ unsigned long long words[2];
    __shared__ uint32_t counts[2];
    const uint32_t c = blockIdx.x, tid = threadIdx.x;
    const uint32_t b0 = ofs[c], m = ofs[c + 1] - b0;
    if (m <= 1) return;
    const uint32_t slot = tid >> 4, t = tid & 15;
    unsigned long long err[2][8];
#pragma unroll
    for (int v = 0; v < 8; v++) err[0][v] = err[1][v] = 0;
    uint32_t cnt1 = 0, cnt = 0;
#pragma unroll 2
    for (uint32_t j = slot; j < m; j += kSelSlots) {
        const uint32_t b = idx[b0 + j];
        const unsigned long long x = blk[b];
        const int a = ch(blocks[(size_t)b * 16 + t], (int)comp);
        const uint32_t lo = (uint32_t)(x & 255), hi = (uint32_t)((x >> 8) & 255);
        const bool typ = lo <= hi;
        uint32_t v[8];
        alpha_values(lo, hi, v);
#pragma unroll
        for (int k = 0; k < 8; k++) {
            const int e = a - (int)v[k];
            const unsigned long long e2 = (unsigned long long)(e * e);
            err[0][k] += typ ? 0ull : e2;
            err[1][k] += typ ? e2 : 0ull;
        }
        cnt1 += typ ? 1u : 0u;
        cnt++;
    }
#pragma unroll
    for (int k = 0; k < 8; k++) {
        part[slot][0][t][k] = err[0][k];
        part[slot][1][t][k] = err[1][k];
    }
    if (t == 0) {
        pcnt[slot][0] = cnt - cnt1;
        pcnt[slot][1] = cnt1;
    }
    __syncthreads();
    if (tid < 32) {
        const uint32_t typ = tid >> 4, tt = tid & 15;
        uint32_t best_v = 0;
        unsigned long long best_e = 0xFFFFFFFFFFull;
        for (uint32_t k = 0; k < 8; k++) {
            unsigned long long e = 0;
            for (uint32_t q = 0; q < kSelSlots; q++) e += part[q][typ][tt][k];
            if (e < best_e) {
                best_e = e;
                best_v = k;
            }
        }
        uint32_t n = 0;
        for (uint32_t q = 0; q < kSelSlots; q++) n += pcnt[q][typ];
        unsigned long long word = (unsigned long long)best_v << (3 * tt);
        for (uint32_t o = 1; o < 16; o <<= 1) word |= __shfl_xor(word, (int)o);
        if (tt == 0) {
            words[typ] = word;
            counts[typ] = n;
        }
    }
    __syncthreads();
    for (uint32_t j = tid; j < m; j += blockDim.x) {
        const uint32_t b = idx[b0 + j];
        const unsigned long long x = blk[b];
        const uint32_t typ = (uint32_t)(x & 255) <= (uint32_t)((x >> 8) & 255) ? 1u : 0u;
        if (counts[typ] > 1) blk[b] = (x & 0xFFFFull) | (words[typ] << 16);
    }
}

// ---------------------------------------------------------------- host side
std::mutex g_q_mu;
bool g_q_tables[64];

int q_tables(int dev)
{
    std::lock_guard<std::mutex> lk(g_q_mu);
    if (dev < 0 || dev >= 64) return CRN2_EINVAL;
    if (g_q_tables[dev]) return CRN2_OK;
    uint8_t t5[256][2], t6[256][2];
    build_omatch4(t5, t6);
    if (hipMemcpyToSymbol(HIP_SYMBOL(q_om5), t5, 512) != hipSuccess ||
        hipMemcpyToSymbol(HIP_SYMBOL(q_om6), t6, 512) != hipSuccess)
        return CRN2_EHIP;
    g_q_tables[dev] = true;
    return CRN2_OK;
}

uint32_t grid(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }

// Stage timing of the tree splits with HIP events on the launch stream
// (crn2_qdxt_profile): per dimension the summed launch times, launch count
// and vector passes (one vector read by one pass of a split).
struct QProf {
    bool on = false;
    double ms[3] = {0, 0, 0};
    uint64_t launches[3] = {0, 0, 0};
    unsigned long long* passes = nullptr;   // device counters, one per dimension
};
QProf g_qprof;
std::mutex g_qprof_mu;

int dim_slot(uint32_t d) { return d == 2 ? 0 : (d == 6 ? 1 : 2); }

// Temporaries of one call live in stream order on the call's stream:
// freeing one waits for nothing, where hipFree waits for the whole device and
// would stall the work that other streams (other textures, the other
// channels of one texture) have in flight.
template <class T>
hipError_t tmp_alloc(T** p, size_t bytes, hipStream_t st)
{
    return hipMallocAsync(reinterpret_cast<void**>(p), bytes, st);
}

inline void tmp_free(void* p, hipStream_t st)
{
    if (p) (void)hipFreeAsync(p, st);
}

// Device buffers that grow as the trees do, stream-ordered on the stream of
// the call that grows them.  A handle's buffers outlive that call: every call
// on a handle has finished its work on its stream when it returns, and the
// handle's destroy waits for the stream of its last call.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int grow(size_t need, hipStream_t st, bool keep)
    {
        if (need <= bytes) return CRN2_OK;
        size_t nb = bytes ? bytes : 4096;
        while (nb < need) nb *= 2;
        void* q = nullptr;
        if (tmp_alloc(&q, nb, st) != hipSuccess) return CRN2_ENOMEM;
        if (keep && bytes && hipMemcpyAsync(q, p, bytes, hipMemcpyDeviceToDevice, st) != hipSuccess) {
            tmp_free(q, st);
            return CRN2_EHIP;
        }
        tmp_free(p, st);
        p = q;
        bytes = nb;
        return CRN2_OK;
    }
    void release(hipStream_t st)
    {
        tmp_free(p, st);
        p = nullptr;
        bytes = 0;
    }
};

struct HNode {
    uint32_t slot, seg, n;
    float var;
    int left = -1, right = -1, cbi = -1;
    bool uns = false;
    int state = 0;   // 0 not split yet, 1 split result known, 2 split queued
    VqResult res;
    uint32_t child_seg = 0, lslot = 0, rslot = 0;
};

// One best-first tree replayed on the host (crn_clusterizer.h:65-176).
struct Tree {
    std::vector<HNode> nodes;
    std::vector<uint32_t> heap;
    uint32_t heap_size = 0, total_leaves = 1, max_size = 0, leaves = 1;
    int pending = -1, split_index = 0;
    bool done = false;

    void insert_heap(uint32_t node)
    {
        const float var = nodes[node].var;
        uint32_t pos = ++heap_size;
        if (heap_size >= heap.size()) heap.resize(heap_size + 1);
        for (;;) {
            const uint32_t parent = pos >> 1;
            if (!parent) break;
            if (nodes[heap[parent]].var > var) break;
            heap[pos] = heap[parent];
            pos = parent;
        }
        heap[pos] = node;
    }
    void down_heap(uint32_t pos)
    {
        uint32_t child;
        const uint32_t orig = heap[pos];
        const float ov = nodes[orig].var;
        while ((child = pos << 1) <= heap_size) {
            if (child < heap_size && nodes[heap[child]].var < nodes[heap[child + 1]].var) child++;
            if (ov > nodes[heap[child]].var) break;
            heap[pos] = heap[child];
            pos = child;
        }
        heap[pos] = orig;
    }
    // Replays as far as the known splits allow; returns false when the
    // popped node still needs its split.
    bool advance()
    {
        for (;;) {
            if (pending >= 0) {
                HNode& p = nodes[(uint32_t)pending];
                if (p.n == 1) {
                    // split_node returns at once for a single vector
                } else {
                    if (p.state != 1) return false;
                    if (!p.res.split) {
                        p.uns = true;
                    } else {
                        HNode l, r;
                        l.slot = p.lslot;
                        l.seg = p.child_seg;
                        l.n = p.res.nl;
                        l.var = p.res.lvar;
                        r.slot = p.rslot;
                        r.seg = p.child_seg + p.res.nl;
                        r.n = p.res.nr;
                        r.var = p.res.rvar;
                        const uint32_t li = (uint32_t)nodes.size();
                        nodes.push_back(l);
                        nodes.push_back(r);
                        HNode& pp = nodes[(uint32_t)pending];
                        pp.left = (int)li;
                        pp.right = (int)li + 1;
                        pp.cbi = split_index++;
                        leaves++;
                        if (nodes[li].n > 1 && nodes[li].var > 0.0f) insert_heap(li);
                        if (nodes[li + 1].n > 1 && nodes[li + 1].var > 0.0f) insert_heap(li + 1);
                    }
                }
                pending = -1;
                total_leaves++;
            }
            if (!(total_leaves < max_size && heap_size)) {
                done = true;
                return true;
            }
            const uint32_t worst = heap[1];
            heap[1] = heap[heap_size];
            heap_size--;
            if (heap_size) down_heap(1);
            pending = (int)worst;
        }
    }
    // The popped node and the heap nodes that may be popped within the
    // remaining budget, not yet split.
    void candidates(std::vector<uint32_t>& out) const
    {
        out.clear();
        if (pending >= 0 && nodes[(uint32_t)pending].state == 0) out.push_back((uint32_t)pending);
        const uint32_t budget = max_size > total_leaves ? max_size - total_leaves : 0;
        std::vector<std::pair<float, uint32_t>> h;
        h.reserve(heap_size);
        for (uint32_t i = 1; i <= heap_size; i++) h.push_back({nodes[heap[i]].var, heap[i]});
        std::sort(h.begin(), h.end(), [](const std::pair<float, uint32_t>& a, const std::pair<float, uint32_t>& b) {
            return a.first > b.first;
        });
        if (!budget || h.empty()) return;
        const float edge = h[std::min<size_t>(budget, h.size()) - 1].first;
        for (size_t i = 0; i < h.size(); i++) {
            if (i >= budget && h[i].first < edge) break;
            if (nodes[h[i].second].state == 0) out.push_back(h[i].second);
        }
    }
    // retrieve_clusters (crn_clusterizer.h:321-352): (segment, count) per cluster
    void retrieve(uint32_t max_clusters, std::vector<std::pair<uint32_t, uint32_t>>& out) const
    {
        out.clear();
        std::vector<uint32_t> stack;
        uint32_t cur = 0;
        for (;;) {
            const HNode& nd = nodes[cur];
            if (nd.left < 0 || (nd.cbi + 2) > (int)max_clusters) {
                out.push_back({nd.seg, nd.n});
                if (stack.empty()) break;
                cur = stack.back();
                stack.pop_back();
                continue;
            }
            cur = (uint32_t)nd.left;
            stack.push_back((uint32_t)nd.right);
        }
    }
};

// The tree machinery of one dimension: vectors, pool, slots.
struct Forest {
    uint32_t d = 0;
    const float* v = nullptr;
    const uint32_t* w = nullptr;
    float exact_max = 0.0f;   // > 0: the vectors' components are integers in [0, exact_max]
    DevBuf pool, cen, tw, jobs, res, varbuf;
    uint32_t pool_used = 0, slots = 0;
    hipStream_t st = nullptr;

    void release()
    {
        for (DevBuf* b : {&pool, &cen, &tw, &jobs, &res, &varbuf}) b->release(st);
    }
    VqArgs args()
    {
        VqArgs a;
        a.v = v;
        a.w = w;
        a.pool = (uint32_t*)pool.p;
        a.cen = (float*)cen.p;
        a.tw = (unsigned long long*)tw.p;
        a.passes = g_qprof.on && g_qprof.passes ? g_qprof.passes + dim_slot(d) : nullptr;
        a.exact_max = exact_max;
        return a;
    }
    int reserve(uint32_t more_pool, uint32_t more_slots)
    {
        int rc = pool.grow(((size_t)pool_used + more_pool) * 4, st, true);
        if (rc == CRN2_OK) rc = cen.grow(((size_t)slots + more_slots) * d * 4, st, true);
        if (rc == CRN2_OK) rc = tw.grow(((size_t)slots + more_slots) * 8, st, true);
        return rc;
    }
    // A new tree over pool[seg, seg + n): root computed on the device.
    int root(Tree& t, uint32_t seg, uint32_t n, uint32_t max_size)
    {
        int rc = reserve(0, 1);
        if (rc == CRN2_OK) rc = varbuf.grow(4, st, false);
        if (rc != CRN2_OK) return rc;
        const uint32_t slot = slots++;
        VqArgs a = args();
        float* vb = (float*)varbuf.p;
        if (d == 2) hipLaunchKernelGGL(vq_root_kernel<2>, dim3(1), dim3(64), 0, st, a, seg, n, slot, vb);
        else if (d == 6) hipLaunchKernelGGL(vq_root_kernel<6>, dim3(1), dim3(64), 0, st, a, seg, n, slot, vb);
        else hipLaunchKernelGGL(vq_root_kernel<16>, dim3(1), dim3(64), 0, st, a, seg, n, slot, vb);
        float var = 0;
        if (hipMemcpyAsync(&var, vb, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return CRN2_EHIP;
        t = Tree();
        HNode r;
        r.slot = slot;
        r.seg = seg;
        r.n = n;
        r.var = var;
        t.nodes.push_back(r);
        t.max_size = max_size;
        t.heap.assign((size_t)max_size + 2, 0);
        t.heap[1] = 0;
        t.heap_size = 1;
        return CRN2_OK;
    }
    // Grows every tree to its end, one device round per step.
    int grow(std::vector<Tree*>& trees)
    {
        std::vector<VqSplit> sp;
        std::vector<std::pair<Tree*, uint32_t>> who;
        std::vector<uint32_t> cand;
        std::vector<VqResult> rs;
        for (;;) {
            sp.clear();
            who.clear();
            for (Tree* t : trees) {
                if (t->done) continue;
                if (t->advance()) continue;
                t->candidates(cand);
                for (uint32_t ni : cand) {
                    HNode& nd = t->nodes[ni];
                    if (nd.n < 2) continue;
                    VqSplit s;
                    s.node = nd.slot;
                    s.seg = nd.seg;
                    s.n = nd.n;
                    sp.push_back(s);
                    who.push_back({t, ni});
                }
            }
            if (sp.empty()) {
                bool all = true;
                for (Tree* t : trees) all = all && t->done;
                if (all) return CRN2_OK;
                return CRN2_EINVAL;   // a pending node without a split to run: cannot happen
            }
            uint64_t need = 0;
            for (const VqSplit& s : sp) need += s.n;
            if (need > 0xFFFFFFFFull - pool_used) return CRN2_ENOMEM;
            int rc = reserve((uint32_t)need, 2 * (uint32_t)sp.size());
            if (rc == CRN2_OK) rc = jobs.grow(sp.size() * sizeof(VqSplit), st, false);
            if (rc == CRN2_OK) rc = res.grow(sp.size() * sizeof(VqResult), st, false);
            if (rc != CRN2_OK) return rc;
            for (size_t i = 0; i < sp.size(); i++) {
                sp[i].child_seg = pool_used;
                pool_used += sp[i].n;
                sp[i].left = slots++;
                sp[i].right = slots++;
                HNode& nd = who[i].first->nodes[who[i].second];
                nd.child_seg = sp[i].child_seg;
                nd.lslot = sp[i].left;
                nd.rslot = sp[i].right;
                nd.state = 2;
            }
            if (hipMemcpyAsync(jobs.p, sp.data(), sp.size() * sizeof(VqSplit), hipMemcpyHostToDevice, st) !=
                hipSuccess)
                return CRN2_EHIP;
            VqArgs a = args();
            const dim3 g((uint32_t)sp.size());
            const VqSplit* js = (const VqSplit*)jobs.p;
            VqResult* rr = (VqResult*)res.p;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            const bool prof = g_qprof.on && hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess;
            if (prof) (void)hipEventRecord(e0, st);
            if (d == 2) hipLaunchKernelGGL(vq_split_kernel<2>, g, dim3(64 * kVqWaves<2>), 0, st, a, js, rr);
            else if (d == 6) hipLaunchKernelGGL(vq_split_kernel<6>, g, dim3(64 * kVqWaves<6>), 0, st, a, js, rr);
            else hipLaunchKernelGGL(vq_split_kernel<16>, g, dim3(64 * kVqWaves<16>), 0, st, a, js, rr);
            if (prof) (void)hipEventRecord(e1, st);
            rs.resize(sp.size());
            if (hipMemcpyAsync(rs.data(), res.p, sp.size() * sizeof(VqResult), hipMemcpyDeviceToHost, st) !=
                    hipSuccess ||
                hipStreamSynchronize(st) != hipSuccess)
                return CRN2_EHIP;
            if (prof) {
                float ms = 0;
                if (hipEventElapsedTime(&ms, e0, e1) == hipSuccess) {
                    std::lock_guard<std::mutex> lk(g_qprof_mu);
                    g_qprof.ms[dim_slot(d)] += ms;
                    g_qprof.launches[dim_slot(d)]++;
                }
            }
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            for (size_t i = 0; i < sp.size(); i++) {
                HNode& nd = who[i].first->nodes[who[i].second];
                nd.res = rs[i];
                nd.state = 1;
            }
        }
    }
    // Threaded division jobs (one launch): left part count per job.
    int divide(const std::vector<DivJob>& js, std::vector<uint32_t>& nl)
    {
        int rc = jobs.grow(js.size() * sizeof(DivJob), st, false);
        if (rc == CRN2_OK) rc = res.grow(js.size() * 4, st, false);
        if (rc != CRN2_OK) return rc;
        if (hipMemcpyAsync(jobs.p, js.data(), js.size() * sizeof(DivJob), hipMemcpyHostToDevice, st) != hipSuccess)
            return CRN2_EHIP;
        VqArgs a = args();
        const DivJob* jp = (const DivJob*)jobs.p;
        uint32_t* o = (uint32_t*)res.p;
        const dim3 g((uint32_t)js.size());
        if (d == 2) hipLaunchKernelGGL(vq_division_kernel<2>, g, dim3(64 * kVqWaves<2>), 0, st, a, jp, o);
        else if (d == 6) hipLaunchKernelGGL(vq_division_kernel<6>, g, dim3(64 * kVqWaves<6>), 0, st, a, jp, o);
        else hipLaunchKernelGGL(vq_division_kernel<16>, g, dim3(64 * kVqWaves<16>), 0, st, a, jp, o);
        nl.resize(js.size());
        if (hipMemcpyAsync(nl.data(), res.p, js.size() * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            return CRN2_EHIP;
        return CRN2_OK;
    }
    // A host list of vector indices at a fresh pool segment.
    int upload(const std::vector<uint32_t>& list, uint32_t& seg)
    {
        const uint32_t n = (uint32_t)list.size();
        int rc = reserve(n, 0);
        if (rc != CRN2_OK) return rc;
        seg = pool_used;
        pool_used += n;
        if (n && (hipMemcpyAsync((uint32_t*)pool.p + seg, list.data(), (size_t)n * 4, hipMemcpyHostToDevice, st) !=
                      hipSuccess ||
                  hipStreamSynchronize(st) != hipSuccess))
            return CRN2_EHIP;
        return CRN2_OK;
    }
    // Identity list 0..n-1 at a fresh pool segment.
    int identity(uint32_t n, uint32_t& seg)
    {
        std::vector<uint32_t> id(n);
        for (uint32_t i = 0; i < n; i++) id[i] = i;
        return upload(id, seg);
    }
    // Clusters to contiguous device lists: idx (n entries) and host offsets.
    int lists(const std::vector<std::pair<uint32_t, uint32_t>>& cl, uint32_t* idx, std::vector<uint32_t>& ofs)
    {
        std::vector<SegCopy> sc(cl.size());
        ofs.assign(cl.size() + 1, 0);
        for (size_t k = 0; k < cl.size(); k++) {
            sc[k].seg = cl[k].first;
            sc[k].n = cl[k].second;
            sc[k].dst = ofs[k];
            ofs[k + 1] = ofs[k] + cl[k].second;
        }
        if (cl.empty()) return CRN2_OK;
        int rc = jobs.grow(sc.size() * sizeof(SegCopy), st, false);
        if (rc != CRN2_OK) return rc;
        if (hipMemcpyAsync(jobs.p, sc.data(), sc.size() * sizeof(SegCopy), hipMemcpyHostToDevice, st) != hipSuccess)
            return CRN2_EHIP;
        hipLaunchKernelGGL(seg_copy_kernel, dim3((uint32_t)sc.size()), dim3(256), 0, st, (const uint32_t*)pool.p,
                           (const SegCopy*)jobs.p, idx);
        return hipStreamSynchronize(st) == hipSuccess ? CRN2_OK : CRN2_EHIP;
    }
};

// threaded_clusterizer::create_clusters (crn_threaded_clusterizer.h:71-181)
// over the vectors listed (ascending) at pool[seg, seg + n): clusters as pool
// segments, partitions in order.
int threaded_clusters_seg(Forest& F, uint32_t seg, uint32_t n, uint32_t max_clusters,
                          std::vector<std::pair<uint32_t, uint32_t>>& cl)
{
    cl.clear();
    int rc = CRN2_OK;
    if (max_clusters < 128) {
        Tree t;
        rc = F.root(t, seg, n, max_clusters);
        if (rc != CRN2_OK) return rc;
        std::vector<Tree*> ts = {&t};
        rc = F.grow(ts);
        if (rc != CRN2_OK) return rc;
        t.retrieve(t.leaves, cl);
        return CRN2_OK;
    }
    rc = F.reserve(2 * n, 0);
    if (rc != CRN2_OK) return rc;
    const uint32_t s1 = F.pool_used;
    F.pool_used += n;
    std::vector<uint32_t> nl;
    rc = F.divide({{seg, n, s1}}, nl);
    if (rc != CRN2_OK) return rc;
    const uint32_t s2 = F.pool_used;
    F.pool_used += n;
    const uint32_t n0 = nl[0], n1 = n - nl[0];
    rc = F.divide({{s1, n0, s2}, {s1 + n0, n1, s2 + n0}}, nl);
    if (rc != CRN2_OK) return rc;
    const uint32_t ps[4] = {s2, s2 + nl[0], s2 + n0, s2 + n0 + nl[1]};
    const uint32_t pn[4] = {nl[0], n0 - nl[0], nl[1], n1 - nl[1]};
    uint32_t parts = 0;
    for (int k = 0; k < 4; k++) parts += pn[k] != 0;
    const uint32_t target = (max_clusters + (parts / 2)) / parts;
    std::vector<Tree> trees(4);
    std::vector<Tree*> tp;
    for (int k = 0; k < 4; k++) {
        if (!pn[k]) continue;
        rc = F.root(trees[k], ps[k], pn[k], target);
        if (rc != CRN2_OK) return rc;
        tp.push_back(&trees[k]);
    }
    rc = F.grow(tp);
    if (rc != CRN2_OK) return rc;
    std::vector<std::pair<uint32_t, uint32_t>> part;
    for (int k = 0; k < 4; k++) {
        if (!pn[k]) continue;
        trees[k].retrieve(trees[k].leaves, part);
        cl.insert(cl.end(), part.begin(), part.end());
    }
    return CRN2_OK;
}

int threaded_clusters(Forest& F, uint32_t n, uint32_t max_clusters, std::vector<std::pair<uint32_t, uint32_t>>& cl)
{
    uint32_t seg = 0;
    const int rc = F.identity(n, seg);
    return rc != CRN2_OK ? rc : threaded_clusters_seg(F, seg, n, max_clusters, cl);
}

}  // namespace

struct crn2_qdxt1 {
    uint32_t n = 0;
    const uint32_t* blocks = nullptr;
    hipStream_t st = nullptr;
    Forest F;            // 6-D endpoint tree
    Tree tree;
    uint32_t max_sel = 0;
    float* vecs = nullptr;
    uint32_t* weights = nullptr;
    uint8_t* enc = nullptr;
    uint32_t endpoint_clusters = 0, selector_clusters = 0;
};

extern "C" int crn2_qdxt1_init(const void* blocks, uint32_t nblocks, uint32_t nlevels, const uint32_t* level_desc,
                               int hierarchical, crn2_qdxt1** out, void* stream)
{
    if (!blocks || !nblocks || !out || (nlevels && !level_desc)) return CRN2_EINVAL;
    *out = nullptr;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return CRN2_EHIP;
    int rc = q_tables(dev);
    if (rc != CRN2_OK) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    // chunks of every level (crn_qdxt1.cpp:106-125)
    std::vector<TChunk> chunks;
    const bool hier = hierarchical && nlevels;
    if (hier) {
        for (uint32_t l = 0; l < nlevels; l++) {
            const uint32_t first = level_desc[3 * l], bw = level_desc[3 * l + 1], bh = level_desc[3 * l + 2];
            if (!bw || !bh || (uint64_t)first + (uint64_t)bw * bh > nblocks) return CRN2_EINVAL;
            float der = 1.5f;
            if (l && der > .25f) {
                const float t = der / powf(3.1f, (float)l);
                der = .25f > t ? .25f : t;
            }
            for (uint32_t cy = 0; cy < (bh + 1) / 2; cy++)
                for (uint32_t cx = 0; cx < (bw + 1) / 2; cx++) chunks.push_back({first, bw, bh, cx, cy, der});
        }
    }
    crn2_qdxt1* q = new crn2_qdxt1();
    q->n = nblocks;
    q->blocks = static_cast<const uint32_t*>(blocks);
    q->st = st;
    q->F.d = 6;
    q->F.st = st;
    q->F.exact_max = 255.0f;   // colours
    TChunk* dch = nullptr;
    uint32_t* lerr = nullptr;
    unsigned long long* table = nullptr;
    uint32_t* count = nullptr;
    uint32_t tsize = 1;
    while (tsize < 2 * nblocks) tsize <<= 1;
    auto fail = [&](int code) {
        if (dch) tmp_free(dch, st);
        if (lerr) tmp_free(lerr, st);
        if (table) tmp_free(table, st);
        if (count) tmp_free(count, st);
        crn2_qdxt1_destroy(q);
        return code;
    };
    if (tmp_alloc(&q->vecs, (size_t)nblocks * 6 * 4, st) != hipSuccess ||
        tmp_alloc(&q->weights, (size_t)nblocks * 4, st) != hipSuccess || tmp_alloc(&q->enc, nblocks, st) != hipSuccess)
        return fail(CRN2_ENOMEM);
    if (hier) {
        const uint32_t nc = (uint32_t)chunks.size();
        if (tmp_alloc(&dch, nc * sizeof(TChunk), st) != hipSuccess ||
            tmp_alloc(&lerr, (size_t)nc * 9 * 4, st) != hipSuccess)
            return fail(CRN2_ENOMEM);
        if (hipMemcpyAsync(dch, chunks.data(), nc * sizeof(TChunk), hipMemcpyHostToDevice, st) != hipSuccess)
            return fail(CRN2_EHIP);
        hipLaunchKernelGGL(qdxt_layout_kernel, dim3(grid((uint64_t)nc * 9, 256)), dim3(256), 0, st, q->blocks, dch,
                           nc, lerr);
        hipLaunchKernelGGL(qdxt_choose_kernel, dim3(grid(nc, 256)), dim3(256), 0, st, q->blocks, dch, nc, lerr,
                           q->vecs, q->weights, q->enc);
    } else {
        hipLaunchKernelGGL(qdxt_block_train_kernel, dim3(grid(nblocks, 256)), dim3(256), 0, st, q->blocks, nblocks,
                           q->vecs, q->weights, q->enc);
    }
    // distinct dxt_fast selector words
    if (tmp_alloc(&table, (size_t)tsize * 8, st) != hipSuccess || tmp_alloc(&count, 4, st) != hipSuccess)
        return fail(CRN2_ENOMEM);
    if (hipMemsetAsync(table, 0, (size_t)tsize * 8, st) != hipSuccess || hipMemsetAsync(count, 0, 4, st) != hipSuccess)
        return fail(CRN2_EHIP);
    hipLaunchKernelGGL(qdxt_selword_kernel, dim3(grid(nblocks, 256)), dim3(256), 0, st, q->blocks, nblocks, table,
                       tsize - 1, count);
    uint32_t distinct = 0;
    if (hipMemcpyAsync(&distinct, count, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail(CRN2_EHIP);
    q->max_sel = distinct + 128;
    // the endpoint tree, grown to 65,535 leaves (crn_qdxt1.cpp:402-407)
    q->F.v = q->vecs;
    q->F.w = q->weights;
    uint32_t seg = 0;
    rc = q->F.identity(nblocks, seg);
    if (rc == CRN2_OK) rc = q->F.root(q->tree, seg, nblocks, 65535);
    std::vector<Tree*> ts = {&q->tree};
    if (rc == CRN2_OK) rc = q->F.grow(ts);
    if (rc != CRN2_OK) return fail(rc);
    tmp_free(dch, st);
    tmp_free(lerr, st);
    tmp_free(table, st);
    tmp_free(count, st);
    dch = nullptr;
    lerr = nullptr;
    table = nullptr;
    count = nullptr;
    *out = q;
    return CRN2_OK;
}

extern "C" int crn2_qdxt1_training(const crn2_qdxt1* q, float* vecs, uint32_t* weights, uint8_t* encodings)
{
    if (!q) return CRN2_EINVAL;
    if ((vecs && hipMemcpy(vecs, q->vecs, (size_t)q->n * 24, hipMemcpyDeviceToHost) != hipSuccess) ||
        (weights && hipMemcpy(weights, q->weights, (size_t)q->n * 4, hipMemcpyDeviceToHost) != hipSuccess) ||
        (encodings && hipMemcpy(encodings, q->enc, q->n, hipMemcpyDeviceToHost) != hipSuccess))
        return CRN2_EHIP;
    return CRN2_OK;
}

extern "C" int crn2_qdxt1_info(const crn2_qdxt1* q, uint32_t* codebook_size, uint32_t* max_selector_clusters,
                               uint32_t* endpoint_clusters, uint32_t* selector_clusters)
{
    if (!q) return CRN2_EINVAL;
    if (codebook_size) *codebook_size = q->tree.leaves;
    if (max_selector_clusters) *max_selector_clusters = q->max_sel;
    if (endpoint_clusters) *endpoint_clusters = q->endpoint_clusters;
    if (selector_clusters) *selector_clusters = q->selector_clusters;
    return CRN2_OK;
}

extern "C" int crn2_qdxt1_pack(crn2_qdxt1* q, const crn2_qdxt1_params* p, float quality_power_mul,
                               void* out_blocks, void* stream)
{
    if (!q || !p || !out_blocks || p->quality_level > 255 || p->dxt_quality < 0 || p->dxt_quality > 4 ||
        (p->flags & ~7))
        return CRN2_EINVAL;
    const int perceptual = p->flags & 1, use_alpha = (p->flags >> 1) & 1;
    const uint32_t thresh = use_alpha ? (p->dxt1a_alpha_threshold > 255 ? 255 : p->dxt1a_alpha_threshold) : 0;
    const float quality = p->quality_level / (float)255;
    // multi-block clusters go through the cluster endpoint search at uber
    // quality: the hc evaluators without 3-colour blocks (DXT5's colour part,
    // DXT1 without both block types), the uber evaluator with them (and
    // DXT1A transparency in a cluster, crn_qdxt1.cpp:590-626)
    if (quality < 1.0f && p->dxt_quality != 4) return CRN2_EUNSUPPORTED;
    if (quality >= 1.0f && p->dxt_quality == 0) return CRN2_EUNSUPPORTED;
    hipStream_t st = static_cast<hipStream_t>(stream);
    q->st = st;
    q->F.st = st;
    const float eq = powf(quality, 1.8f * quality_power_mul), sq = powf(quality, 1.65f * quality_power_mul);
    const uint32_t cbs = q->tree.leaves;
    uint32_t mec = (uint32_t)(cbs * eq);
    mec = mec < 96 ? 96 : (mec > cbs ? cbs : mec);
    uint32_t msc = (uint32_t)(q->max_sel * sq);
    msc = msc < 128 ? 128 : (msc > q->max_sel ? q->max_sel : msc);
    unsigned long long* out = static_cast<unsigned long long*>(out_blocks);
    const uint32_t n = q->n;
    int rc = CRN2_OK;
    if (quality >= 1.0f) {
        // every block alone: the block search with dxt_image's DXT1A rule
        Dxt1Rec* rec = nullptr;
        if (hipMallocAsync((void**)&rec, (size_t)n * sizeof(Dxt1Rec), st) != hipSuccess) return CRN2_ENOMEM;
        const int fl = perceptual | (use_alpha << 1) | (thresh ? 64 : 0) | (int)(thresh << 8);
        rc = crn2_dxt1_blocks_device(q->blocks, n, p->dxt_quality, fl, rec, st);
        if (rc == CRN2_OK)
            hipLaunchKernelGGL(single_blocks_kernel, dim3(grid(n, 256)), dim3(256), 0, st, rec, n, out);
        (void)hipFreeAsync(rec, st);
        q->endpoint_clusters = n;
        q->selector_clusters = 0;
        if (rc == CRN2_OK && hipStreamSynchronize(st) != hipSuccess) rc = CRN2_EHIP;
        return rc;
    }
    std::vector<std::pair<uint32_t, uint32_t>> cl;
    q->tree.retrieve(mec, cl);
    const uint32_t nc = (uint32_t)cl.size();
    q->endpoint_clusters = nc;
    uint32_t *idx = nullptr, *px = nullptr, *tofs = nullptr, *bcl = nullptr, *sofs_d = nullptr;
    Dxt1Rec* rec = nullptr;
    uint8_t* sel = nullptr;
    float* sv = nullptr;
    uint32_t* sw = nullptr;
    auto cleanup = [&]() {
        for (void* x : {(void*)idx, (void*)px, (void*)tofs, (void*)bcl, (void*)rec, (void*)sel, (void*)sv, (void*)sw,
                        (void*)sofs_d})
            tmp_free(x, st);
    };
    if (tmp_alloc(&idx, (size_t)n * 4, st) != hipSuccess || tmp_alloc(&px, (size_t)n * 64, st) != hipSuccess ||
        tmp_alloc(&tofs, ((size_t)nc + 1) * 4, st) != hipSuccess ||
        tmp_alloc(&bcl, (size_t)n * 4, st) != hipSuccess ||
        tmp_alloc(&rec, (size_t)nc * sizeof(Dxt1Rec), st) != hipSuccess ||
        tmp_alloc(&sel, (size_t)n * 16, st) != hipSuccess) {
        cleanup();
        return CRN2_ENOMEM;
    }
    std::vector<uint32_t> ofs;
    rc = q->F.lists(cl, idx, ofs);
    if (rc != CRN2_OK) {
        cleanup();
        return rc;
    }
    std::vector<uint32_t> texofs(nc + 1), blockcl(n);
    for (uint32_t k = 0; k <= nc; k++) texofs[k] = ofs[k] * 16;
    for (uint32_t k = 0; k < nc; k++)
        for (uint32_t j = ofs[k]; j < ofs[k + 1]; j++) blockcl[j] = k;
    if (hipMemcpyAsync(tofs, texofs.data(), ((size_t)nc + 1) * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(bcl, blockcl.data(), (size_t)n * 4, hipMemcpyHostToDevice, st) != hipSuccess) {
        cleanup();
        return CRN2_EHIP;
    }
    hipLaunchKernelGGL(gather_texels_kernel, dim3(grid((uint64_t)n * 16, 256)), dim3(256), 0, st, q->blocks, idx, n,
                       px);
    rc = crn2_dxt1_clusters_device(px, tofs, nc, perceptual | (use_alpha << 1) | (int)(thresh << 8), rec, sel, st);
    if (rc != CRN2_OK) {
        cleanup();
        return rc;
    }
    hipLaunchKernelGGL(cluster_blocks_kernel, dim3(grid(n, 256)), dim3(256), 0, st, idx, bcl, n, rec, sel, out);
    // selector clusters (crn_qdxt1.cpp:843-874, 1006-1020)
    if (tmp_alloc(&sv, (size_t)n * 64, st) != hipSuccess || tmp_alloc(&sw, (size_t)n * 4, st) != hipSuccess) {
        cleanup();
        return CRN2_ENOMEM;
    }
    hipLaunchKernelGGL(selector_vec_kernel, dim3(grid(n, 256)), dim3(256), 0, st, out, n, perceptual, sv, sw);
    Forest S;
    S.d = 16;
    S.v = sv;
    S.w = sw;
    S.st = st;
    S.exact_max = 3.0f;   // g_dxt1_to_linear values
    std::vector<std::pair<uint32_t, uint32_t>> scl;
    rc = threaded_clusters(S, n, msc, scl);
    if (rc == CRN2_OK) {
        q->selector_clusters = (uint32_t)scl.size();
        std::vector<uint32_t> sofs;
        rc = S.lists(scl, idx, sofs);
        if (rc == CRN2_OK && tmp_alloc(&sofs_d, sofs.size() * 4, st) != hipSuccess) rc = CRN2_ENOMEM;
        if (rc == CRN2_OK &&
            hipMemcpyAsync(sofs_d, sofs.data(), sofs.size() * 4, hipMemcpyHostToDevice, st) != hipSuccess)
            rc = CRN2_EHIP;
        if (rc == CRN2_OK && !scl.empty())
            hipLaunchKernelGGL(selector_opt_kernel, dim3((uint32_t)scl.size()), dim3(256), 0, st, q->blocks, idx,
                               sofs_d, out, perceptual, thresh);
    }
    if (rc == CRN2_OK && hipStreamSynchronize(st) != hipSuccess) rc = CRN2_EHIP;
    S.release();
    cleanup();
    return rc;
}

extern "C" int crn2_qdxt1_destroy(crn2_qdxt1* q)
{
    if (!q) return CRN2_OK;
    // the handle's work is done once the stream of its last call is idle;
    // its buffers then go back in order on the null stream
    if (q->st) (void)hipStreamSynchronize(q->st);
    q->F.st = nullptr;
    q->F.release();
    tmp_free(q->vecs, nullptr);
    tmp_free(q->weights, nullptr);
    tmp_free(q->enc, nullptr);
    delete q;
    return CRN2_OK;
}

// The clusterizer on its own, for any vectors (dimension 2, 6 or 16) in
// device memory: grown to max_size leaves and cut at max_clusters (0: every
// leaf).  idx / ofs (host) receive the clusters.
extern "C" int crn2_vq_clusters_device(const float* vecs, const uint32_t* weights, uint32_t n, uint32_t dim,
                                       uint32_t max_size, uint32_t max_clusters, uint32_t* idx, uint32_t* ofs,
                                       uint32_t* nclusters, uint32_t* codebook_size, void* stream)
{
    if (!vecs || !weights || !n || (dim != 2 && dim != 6 && dim != 16) || !idx || !ofs || !nclusters) return CRN2_EINVAL;
    Forest F;
    F.d = dim;
    F.v = vecs;
    F.w = weights;
    F.st = static_cast<hipStream_t>(stream);
    Tree t;
    uint32_t seg = 0;
    int rc = F.identity(n, seg);
    if (rc == CRN2_OK) rc = F.root(t, seg, n, max_size);
    std::vector<Tree*> ts = {&t};
    if (rc == CRN2_OK) rc = F.grow(ts);
    std::vector<std::pair<uint32_t, uint32_t>> cl;
    uint32_t* didx = nullptr;
    std::vector<uint32_t> o;
    if (rc == CRN2_OK) {
        t.retrieve(max_clusters ? max_clusters : t.leaves, cl);
        if (tmp_alloc(&didx, (size_t)n * 4, F.st) != hipSuccess) rc = CRN2_ENOMEM;
    }
    if (rc == CRN2_OK) rc = F.lists(cl, didx, o);
    if (rc == CRN2_OK && hipMemcpy(idx, didx, (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = CRN2_EHIP;
    if (rc == CRN2_OK) {
        memcpy(ofs, o.data(), o.size() * 4);
        *nclusters = (uint32_t)cl.size();
        if (codebook_size) *codebook_size = t.leaves;
    }
    if (didx) tmp_free(didx, F.st);
    F.release();
    return rc;
}

// threaded_clusterizer::create_clusters for vectors in device memory.
extern "C" int crn2_vq_threaded_clusters_device(const float* vecs, const uint32_t* weights, uint32_t n,
                                                uint32_t dim, uint32_t max_clusters, uint32_t* idx, uint32_t* ofs,
                                                uint32_t* nclusters, void* stream)
{
    if (!vecs || !weights || !n || (dim != 2 && dim != 6 && dim != 16) || !idx || !ofs || !nclusters) return CRN2_EINVAL;
    Forest F;
    F.d = dim;
    F.v = vecs;
    F.w = weights;
    F.st = static_cast<hipStream_t>(stream);
    std::vector<std::pair<uint32_t, uint32_t>> cl;
    int rc = threaded_clusters(F, n, max_clusters, cl);
    uint32_t* didx = nullptr;
    std::vector<uint32_t> o;
    if (rc == CRN2_OK && tmp_alloc(&didx, (size_t)n * 4, F.st) != hipSuccess) rc = CRN2_ENOMEM;
    if (rc == CRN2_OK) rc = F.lists(cl, didx, o);
    if (rc == CRN2_OK && hipMemcpy(idx, didx, (size_t)n * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = CRN2_EHIP;
    if (rc == CRN2_OK) {
        memcpy(ofs, o.data(), o.size() * 4);
        *nclusters = (uint32_t)cl.size();
    }
    if (didx) tmp_free(didx, F.st);
    F.release();
    return rc;
}

// ------------------------------------------------------------------- qdxt5
struct crn2_qdxt5 {
    uint32_t n = 0, comp = 3;
    const uint32_t* blocks = nullptr;
    hipStream_t st = nullptr;
    Forest F;            // 2-D endpoint tree
    Tree tree;
    uint32_t max_sel = 0;
    float* vecs = nullptr;
    uint32_t* weights = nullptr;
    uint8_t* enc = nullptr;
    uint32_t endpoint_clusters = 0, selector_clusters = 0;
};

extern "C" int crn2_qdxt5_init(const void* blocks, uint32_t nblocks, uint32_t nlevels, const uint32_t* level_desc,
                               int hierarchical, uint32_t comp_index, crn2_qdxt5** out, void* stream)
{
    if (!blocks || !nblocks || !out || (nlevels && !level_desc) || comp_index > 3) return CRN2_EINVAL;
    *out = nullptr;
    hipStream_t st = static_cast<hipStream_t>(stream);
    std::vector<TChunk> chunks;
    const bool hier = hierarchical && nlevels;
    if (hier) {
        for (uint32_t l = 0; l < nlevels; l++) {
            const uint32_t first = level_desc[3 * l], bw = level_desc[3 * l + 1], bh = level_desc[3 * l + 2];
            if (!bw || !bh || (uint64_t)first + (uint64_t)bw * bh > nblocks) return CRN2_EINVAL;
            // the alpha PSNR derating (crn_qdxt5.cpp:223-229)
            float der = 2.4f;
            if (l && der > .25f) {
                const float t = der / powf(3.0f, (float)l);
                der = .25f > t ? .25f : t;
            }
            for (uint32_t cy = 0; cy < (bh + 1) / 2; cy++)
                for (uint32_t cx = 0; cx < (bw + 1) / 2; cx++) chunks.push_back({first, bw, bh, cx, cy, der});
        }
    }
    crn2_qdxt5* q = new crn2_qdxt5();
    q->n = nblocks;
    q->comp = comp_index;
    q->blocks = static_cast<const uint32_t*>(blocks);
    q->st = st;
    q->F.d = 2;
    q->F.st = st;
    q->F.exact_max = 255.0f;   // alpha values
    TChunk* dch = nullptr;
    uint32_t* lerr = nullptr;
    unsigned long long* table = nullptr;
    uint32_t* count = nullptr;
    uint32_t tsize = 1;
    while (tsize < 2 * nblocks) tsize <<= 1;
    auto release = [&]() {
        if (dch) tmp_free(dch, st);
        if (lerr) tmp_free(lerr, st);
        if (table) tmp_free(table, st);
        if (count) tmp_free(count, st);
    };
    auto fail = [&](int code) {
        release();
        crn2_qdxt5_destroy(q);
        return code;
    };
    if (tmp_alloc(&q->vecs, (size_t)nblocks * 2 * 4, st) != hipSuccess ||
        tmp_alloc(&q->weights, (size_t)nblocks * 4, st) != hipSuccess || tmp_alloc(&q->enc, nblocks, st) != hipSuccess)
        return fail(CRN2_ENOMEM);
    if (hier) {
        const uint32_t nc = (uint32_t)chunks.size();
        if (tmp_alloc(&dch, nc * sizeof(TChunk), st) != hipSuccess ||
            tmp_alloc(&lerr, (size_t)nc * 9 * 4, st) != hipSuccess)
            return fail(CRN2_ENOMEM);
        if (hipMemcpyAsync(dch, chunks.data(), nc * sizeof(TChunk), hipMemcpyHostToDevice, st) != hipSuccess)
            return fail(CRN2_EHIP);
        hipLaunchKernelGGL(qdxt5_layout_kernel, dim3(grid((uint64_t)nc * 9, 256)), dim3(256), 0, st, q->blocks, dch,
                           nc, comp_index, lerr);
        hipLaunchKernelGGL(qdxt5_choose_kernel, dim3(grid(nc, 256)), dim3(256), 0, st, q->blocks, dch, nc,
                           comp_index, lerr, q->vecs, q->weights, q->enc);
    } else {
        hipLaunchKernelGGL(qdxt5_block_train_kernel, dim3(grid(nblocks, 256)), dim3(256), 0, st, q->blocks, nblocks,
                           comp_index, q->vecs, q->weights, q->enc);
    }
    if (tmp_alloc(&table, (size_t)tsize * 8, st) != hipSuccess || tmp_alloc(&count, 4, st) != hipSuccess)
        return fail(CRN2_ENOMEM);
    if (hipMemsetAsync(table, 0, (size_t)tsize * 8, st) != hipSuccess || hipMemsetAsync(count, 0, 4, st) != hipSuccess)
        return fail(CRN2_EHIP);
    hipLaunchKernelGGL(qdxt5_selword_kernel, dim3(grid(nblocks, 256)), dim3(256), 0, st, q->blocks, nblocks,
                       comp_index, table, tsize - 1, count);
    uint32_t distinct = 0;
    if (hipMemcpyAsync(&distinct, count, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return fail(CRN2_EHIP);
    q->max_sel = distinct + 128;
    q->F.v = q->vecs;
    q->F.w = q->weights;
    uint32_t seg = 0;
    int rc = q->F.identity(nblocks, seg);
    if (rc == CRN2_OK) rc = q->F.root(q->tree, seg, nblocks, 65535);
    std::vector<Tree*> ts = {&q->tree};
    if (rc == CRN2_OK) rc = q->F.grow(ts);
    if (rc != CRN2_OK) return fail(rc);
    release();
    *out = q;
    return CRN2_OK;
}

extern "C" int crn2_qdxt5_training(const crn2_qdxt5* q, float* vecs, uint32_t* weights, uint8_t* encodings)
{
    if (!q) return CRN2_EINVAL;
    if ((vecs && hipMemcpy(vecs, q->vecs, (size_t)q->n * 8, hipMemcpyDeviceToHost) != hipSuccess) ||
        (weights && hipMemcpy(weights, q->weights, (size_t)q->n * 4, hipMemcpyDeviceToHost) != hipSuccess) ||
        (encodings && hipMemcpy(encodings, q->enc, q->n, hipMemcpyDeviceToHost) != hipSuccess))
        return CRN2_EHIP;
    return CRN2_OK;
}

extern "C" int crn2_qdxt5_info(const crn2_qdxt5* q, uint32_t* codebook_size, uint32_t* max_selector_clusters,
                               uint32_t* endpoint_clusters, uint32_t* selector_clusters)
{
    if (!q) return CRN2_EINVAL;
    if (codebook_size) *codebook_size = q->tree.leaves;
    if (max_selector_clusters) *max_selector_clusters = q->max_sel;
    if (endpoint_clusters) *endpoint_clusters = q->endpoint_clusters;
    if (selector_clusters) *selector_clusters = q->selector_clusters;
    return CRN2_OK;
}

extern "C" int crn2_qdxt5_pack(crn2_qdxt5* q, const crn2_qdxt5_params* p, void* out_blocks, void* stream)
{
    if (!q || !p || !out_blocks || p->quality_level > 255 || p->dxt_quality < 0 || p->dxt_quality > 4 ||
        (p->flags & ~1))
        return CRN2_EINVAL;
    const int use_both = p->flags & 1;
    hipStream_t st = static_cast<hipStream_t>(stream);
    q->st = st;
    q->F.st = st;
    const float quality = p->quality_level / (float)255;
    const float eq = powf(quality, 2.1f), sq = powf(quality, 1.65f);
    const uint32_t cbs = q->tree.leaves;
    uint32_t mec = (uint32_t)(cbs * eq);
    mec = mec < 16 ? 16 : (mec > cbs ? cbs : mec);
    uint32_t msc = (uint32_t)(q->max_sel * sq);
    msc = msc < 32 ? 32 : (msc > q->max_sel ? q->max_sel : msc);
    unsigned long long* out = static_cast<unsigned long long*>(out_blocks);
    const uint32_t n = q->n;
    std::vector<std::pair<uint32_t, uint32_t>> cl;
    uint32_t seg0 = 0;
    int rc = CRN2_OK;
    if (quality >= 1.0f) {
        // every block alone, in block order
        rc = q->F.identity(n, seg0);
        if (rc != CRN2_OK) return rc;
        cl.resize(n);
        for (uint32_t b = 0; b < n; b++) cl[b] = {seg0 + b, 1};
    } else {
        q->tree.retrieve(mec, cl);
    }
    const uint32_t nc = (uint32_t)cl.size();
    q->endpoint_clusters = nc;
    q->selector_clusters = 0;
    uint32_t *idx = nullptr, *vofs = nullptr, *bcl = nullptr, *res = nullptr, *sw = nullptr, *sofs_d = nullptr;
    uint8_t *vals = nullptr, *sel = nullptr, *kind = nullptr;
    float* sv = nullptr;
    auto cleanup = [&]() {
        for (void* x : {(void*)idx, (void*)vofs, (void*)bcl, (void*)res, (void*)sw, (void*)sofs_d, (void*)vals,
                        (void*)sel, (void*)kind, (void*)sv})
            tmp_free(x, st);
    };
    if (tmp_alloc(&idx, (size_t)n * 4, st) != hipSuccess ||
        tmp_alloc(&vofs, ((size_t)nc + 1) * 4, st) != hipSuccess ||
        tmp_alloc(&bcl, (size_t)n * 4, st) != hipSuccess || tmp_alloc(&res, (size_t)nc * 4, st) != hipSuccess ||
        tmp_alloc(&vals, (size_t)n * 16, st) != hipSuccess || tmp_alloc(&sel, (size_t)n * 16, st) != hipSuccess) {
        cleanup();
        return CRN2_ENOMEM;
    }
    std::vector<uint32_t> ofs;
    rc = q->F.lists(cl, idx, ofs);
    if (rc != CRN2_OK) {
        cleanup();
        return rc;
    }
    std::vector<uint32_t> vofs_h(nc + 1), blockcl(n);
    for (uint32_t k = 0; k <= nc; k++) vofs_h[k] = ofs[k] * 16;
    for (uint32_t k = 0; k < nc; k++)
        for (uint32_t j = ofs[k]; j < ofs[k + 1]; j++) blockcl[j] = k;
    if (hipMemcpyAsync(vofs, vofs_h.data(), ((size_t)nc + 1) * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(bcl, blockcl.data(), (size_t)n * 4, hipMemcpyHostToDevice, st) != hipSuccess) {
        cleanup();
        return CRN2_EHIP;
    }
    hipLaunchKernelGGL(gather_values_kernel, dim3(grid((uint64_t)n * 16, 256)), dim3(256), 0, st, q->blocks, idx, n,
                       q->comp, vals);
    if (p->dxt_quality == 0)
        hipLaunchKernelGGL(alpha_fast_kernel, dim3(nc), dim3(64), 0, st, vals, vofs, res, sel);
    else
        hipLaunchKernelGGL(qdxt5_search_kernel, dim3(nc), dim3(64), 0, st, vals, vofs, p->dxt_quality, use_both, res,
                           sel);
    hipLaunchKernelGGL(alpha_blocks_kernel, dim3(grid(n, 256)), dim3(256), 0, st, idx, bcl, n, res, sel, out);
    if (quality >= 1.0f) {
        rc = hipStreamSynchronize(st) == hipSuccess ? CRN2_OK : CRN2_EHIP;
        cleanup();
        return rc;
    }
    // selector clusters per block type (crn_qdxt5.cpp:700-841)
    if (tmp_alloc(&sv, (size_t)n * 64, st) != hipSuccess || tmp_alloc(&sw, (size_t)n * 4, st) != hipSuccess ||
        tmp_alloc(&kind, n, st) != hipSuccess) {
        cleanup();
        return CRN2_ENOMEM;
    }
    hipLaunchKernelGGL(alpha_selector_vec_kernel, dim3(grid(n, 256)), dim3(256), 0, st, out, n, sv, sw, kind);
    std::vector<uint8_t> kh(n);
    if (hipMemcpyAsync(kh.data(), kind, n, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        cleanup();
        return CRN2_EHIP;
    }
    Forest S;
    S.d = 16;
    S.v = sv;
    S.w = sw;
    S.st = st;
    S.exact_max = 7.0f;   // g_dxt5_to_linear / g_dxt5_alpha6_to_linear values
    std::vector<std::pair<uint32_t, uint32_t>> scl, part;
    for (uint32_t type = 0; type < 2 && rc == CRN2_OK; type++) {
        std::vector<uint32_t> members;
        for (uint32_t b = 0; b < n; b++)
            if (kh[b] == type) members.push_back(b);
        const uint32_t cnt = (uint32_t)members.size();
        if (!cnt) continue;
        if ((cnt / (float)n) < .01f) continue;
        uint32_t mc = (uint32_t)(((uint64_t)cnt * msc + (n - 1)) / n);
        mc = mc < 64 ? 64 : mc;
        mc = mc < cnt ? mc : cnt;
        if (mc >= cnt) continue;
        uint32_t seg = 0;
        rc = S.upload(members, seg);
        if (rc == CRN2_OK) rc = threaded_clusters_seg(S, seg, cnt, mc, part);
        if (rc == CRN2_OK) scl.insert(scl.end(), part.begin(), part.end());
    }
    if (rc == CRN2_OK && !scl.empty()) {
        q->selector_clusters = (uint32_t)scl.size();
        std::vector<uint32_t> sofs;
        rc = S.lists(scl, idx, sofs);
        if (rc == CRN2_OK && tmp_alloc(&sofs_d, sofs.size() * 4, st) != hipSuccess) rc = CRN2_ENOMEM;
        if (rc == CRN2_OK &&
            hipMemcpyAsync(sofs_d, sofs.data(), sofs.size() * 4, hipMemcpyHostToDevice, st) != hipSuccess)
            rc = CRN2_EHIP;
        if (rc == CRN2_OK)
            hipLaunchKernelGGL(alpha_selector_opt_kernel, dim3((uint32_t)scl.size()), dim3(256), 0, st, q->blocks,
                               idx, sofs_d, out, q->comp);
    }
    if (rc == CRN2_OK && hipStreamSynchronize(st) != hipSuccess) rc = CRN2_EHIP;
    S.release();
    cleanup();
    return rc;
}

extern "C" int crn2_qdxt5_destroy(crn2_qdxt5* q)
{
    if (!q) return CRN2_OK;
    // the handle's work is done once the stream of its last call is idle;
    // its buffers then go back in order on the null stream
    if (q->st) (void)hipStreamSynchronize(q->st);
    q->F.st = nullptr;
    q->F.release();
    tmp_free(q->vecs, nullptr);
    tmp_free(q->weights, nullptr);
    tmp_free(q->enc, nullptr);
    delete q;
    return CRN2_OK;
}

// Split timing: on != 0 resets and starts it; ms[3], launches[3] and
// passes[3] (2-, 6- and 16-D trees) receive the totals since then.
extern "C" int crn2_qdxt_profile(int on, double* ms, uint64_t* launches, uint64_t* passes)
{
    std::lock_guard<std::mutex> lk(g_qprof_mu);
    if (!g_qprof.passes && hipMalloc((void**)&g_qprof.passes, 3 * sizeof(unsigned long long)) != hipSuccess) {
        g_qprof.passes = nullptr;
        return CRN2_ENOMEM;
    }
    if (on) {
        for (int i = 0; i < 3; i++) {
            g_qprof.ms[i] = 0;
            g_qprof.launches[i] = 0;
        }
        if (hipMemset(g_qprof.passes, 0, 3 * sizeof(unsigned long long)) != hipSuccess) return CRN2_EHIP;
        g_qprof.on = true;
        return CRN2_OK;
    }
    g_qprof.on = false;
    unsigned long long p[3] = {0, 0, 0};
    if (hipMemcpy(p, g_qprof.passes, sizeof(p), hipMemcpyDeviceToHost) != hipSuccess) return CRN2_EHIP;
    for (int i = 0; i < 3; i++) {
        if (ms) ms[i] = g_qprof.ms[i];
        if (launches) launches[i] = g_qprof.launches[i];
        if (passes) passes[i] = p[i];
    }
    return CRN2_OK;
}

// ------------------------------------------------- clustered DDS (crn_compress)
// dds_comp::convert_to_dxt below the top quality (crnlib/crn_dds_comp.cpp:
// 148-205) through mipmapped_texture::qdxt_pack_init / qdxt_pack
// (crn_mipmapped_texture.cpp:2310-2560): the pixel blocks of every face and
// level (faces outer, texels past the edge clamped) packed by qdxt1 (DXT1,
// DXT5's colour) and qdxt5 (DXT5's alpha, DXT5A, the two DXN channels), the
// elements interleaved as qdxt_pack does and written as a DDS file.
extern "C" int crn2_compress_dds_clustered(const uint8_t* const* levels, uint32_t width, uint32_t height,
                                           uint32_t nlevels, uint32_t faces, uint32_t format,
                                           uint32_t quality_level, int dxt_quality, int flags,
                                           uint32_t alpha_threshold, uint8_t** out, uint32_t* out_size)
{
    if (!out || !out_size) return CRN2_EINVAL;
    *out = nullptr;
    *out_size = 0;
    if (!levels || !width || !height || !nlevels || nlevels > 16 || (faces != 1 && faces != 6) ||
        quality_level > 255 || dxt_quality < 0 || dxt_quality > 4 || (flags & ~7))
        return CRN2_EINVAL;
    const bool dxt1 = format == CRN2_FMT_DXT1, dxt5 = format == CRN2_FMT_DXT5, dxt5a = format == CRN2_FMT_DXT5A;
    const bool dxn = format == CRN2_FMT_DXN_XY || format == CRN2_FMT_DXN_YX;
    if (!dxt1 && !dxt5 && !dxt5a && !dxn) return CRN2_EUNSUPPORTED;
    const int perceptual = flags & 1, both = (flags >> 1) & 1, hier = (flags >> 2) & 1;
    // blocks and level descriptors (qdxt_pack_init, :2410-2469)
    std::vector<uint32_t> blocks;
    std::vector<uint32_t> desc;
    std::vector<uint32_t> lcount;
    for (uint32_t f = 0; f < faces; f++)
        for (uint32_t l = 0; l < nlevels; l++) {
            const uint8_t* img = levels[f * nlevels + l];
            if (!img) return CRN2_EINVAL;
            const uint32_t w = std::max<uint32_t>(1, width >> l), h = std::max<uint32_t>(1, height >> l);
            const uint32_t bw = (w + 3) / 4, bh = (h + 3) / 4;
            desc.push_back((uint32_t)(blocks.size() / 16));
            desc.push_back(bw);
            desc.push_back(bh);
            lcount.push_back(bw * bh);
            for (uint32_t by = 0; by < bh; by++)
                for (uint32_t bx = 0; bx < bw; bx++)
                    for (uint32_t y = 0; y < 4; y++)
                        for (uint32_t x = 0; x < 4; x++) {
                            const uint32_t py = std::min(by * 4 + y, h - 1), px = std::min(bx * 4 + x, w - 1);
                            uint32_t t;
                            memcpy(&t, img + 4 * ((size_t)py * w + px), 4);
                            blocks.push_back(t);
                        }
        }
    const uint32_t n = (uint32_t)(blocks.size() / 16);
    const uint32_t nl = (uint32_t)lcount.size();
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return CRN2_EHIP;
    void* dblocks = nullptr;
    unsigned long long *c = nullptr, *a0 = nullptr, *a1 = nullptr;
    auto release = [&]() {
        for (void* x : {(void*)dblocks, (void*)c, (void*)a0, (void*)a1})
            if (x) (void)hipFree(x);
    };
    if (hipMalloc(&dblocks, blocks.size() * 4) != hipSuccess ||
        hipMemcpy(dblocks, blocks.data(), blocks.size() * 4, hipMemcpyHostToDevice) != hipSuccess) {
        release();
        return CRN2_ENOMEM;
    }
    // the colour (qdxt_pack, :2528-2544: codebook curve 0.75 for DXT5's
    // colour) and each alpha channel (qdxt_pack_init :2318-2365) are
    // independent: one job each, on its own stream and host thread, so that
    // their serial tree rounds overlap
    const uint32_t comps[2] = {dxn ? (format == CRN2_FMT_DXN_XY ? 0u : 1u) : 3u,
                               format == CRN2_FMT_DXN_XY ? 1u : 0u};
    const int nalpha = dxn ? 2 : ((dxt5 || dxt5a) ? 1 : 0);
    struct Job {
        int alpha;            // -1: colour; else the alpha channel's index
        unsigned long long* dst;
        int rc;
    };
    std::vector<Job> jobs;
    int rc = CRN2_OK;
    if (dxt1 || dxt5) {
        rc = hipMalloc((void**)&c, (size_t)n * 8) == hipSuccess ? CRN2_OK : CRN2_ENOMEM;
        jobs.push_back({-1, c, CRN2_OK});
    }
    for (int k = 0; k < nalpha && rc == CRN2_OK; k++) {
        unsigned long long*& dst = k ? a1 : a0;
        rc = hipMalloc((void**)&dst, (size_t)n * 8) == hipSuccess ? CRN2_OK : CRN2_ENOMEM;
        jobs.push_back({k, dst, CRN2_OK});
    }
    if (rc != CRN2_OK) {
        release();
        return rc;
    }
    auto run = [&](Job& jb) {
        hipStream_t st = nullptr;
        if (hipSetDevice(dev) != hipSuccess || hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            jb.rc = CRN2_EHIP;
            return;
        }
        if (jb.alpha < 0) {
            crn2_qdxt1* q = nullptr;
            jb.rc = crn2_qdxt1_init(dblocks, n, nl, desc.data(), hier, &q, st);
            if (jb.rc == CRN2_OK) {
                crn2_qdxt1_params p;
                p.quality_level = quality_level;
                p.dxt_quality = dxt_quality;
                p.flags = perceptual | ((dxt1 && both) ? CRN2_QDXT_ALPHA_BLOCKS : 0) |
                          (hier ? CRN2_QDXT_HIERARCHICAL : 0);
                p.dxt1a_alpha_threshold = alpha_threshold;
                jb.rc = crn2_qdxt1_pack(q, &p, dxt5 ? 0.75f : 1.0f, jb.dst, st);
            }
            if (q) crn2_qdxt1_destroy(q);
        } else {
            crn2_qdxt5* q = nullptr;
            jb.rc = crn2_qdxt5_init(dblocks, n, nl, desc.data(), hier, comps[jb.alpha], &q, st);
            if (jb.rc == CRN2_OK) {
                crn2_qdxt5_params p;
                p.quality_level = quality_level;
                p.dxt_quality = dxt_quality;
                p.flags = both ? CRN2_QDXT5_BOTH_BLOCK_TYPES : 0;
                jb.rc = crn2_qdxt5_pack(q, &p, jb.dst, st);
            }
            if (q) crn2_qdxt5_destroy(q);
        }
        if (hipStreamSynchronize(st) != hipSuccess && jb.rc == CRN2_OK) jb.rc = CRN2_EHIP;
        (void)hipStreamDestroy(st);
    };
    {
        // a helper thread that cannot start runs its job here instead
        std::vector<std::thread> helpers;
        std::vector<size_t> inline_jobs;
        for (size_t k = 1; k < jobs.size(); k++) {
            try {
                helpers.emplace_back(run, std::ref(jobs[k]));
            } catch (...) {
                inline_jobs.push_back(k);
            }
        }
        if (!jobs.empty()) run(jobs[0]);
        for (size_t k : inline_jobs) run(jobs[k]);
        for (std::thread& t : helpers) t.join();
    }
    for (const Job& jb : jobs)
        if (jb.rc != CRN2_OK && rc == CRN2_OK) rc = jb.rc;
    std::vector<unsigned long long> hc, ha0, ha1;
    if (rc == CRN2_OK && c) {
        hc.resize(n);
        if (hipMemcpy(hc.data(), c, (size_t)n * 8, hipMemcpyDeviceToHost) != hipSuccess) rc = CRN2_EHIP;
    }
    if (rc == CRN2_OK && a0) {
        ha0.resize(n);
        if (hipMemcpy(ha0.data(), a0, (size_t)n * 8, hipMemcpyDeviceToHost) != hipSuccess) rc = CRN2_EHIP;
    }
    if (rc == CRN2_OK && a1) {
        ha1.resize(n);
        if (hipMemcpy(ha1.data(), a1, (size_t)n * 8, hipMemcpyDeviceToHost) != hipSuccess) rc = CRN2_EHIP;
    }
    release();
    if (rc != CRN2_OK) return rc;
    // elements per block as qdxt_pack interleaves them (:2536-2553)
    const uint32_t per = (uint32_t)(c ? 1 : 0) + (uint32_t)nalpha;
    std::vector<std::vector<unsigned long long>> data(nl);
    std::vector<const void*> ptrs(nl);
    for (uint32_t i = 0, b0 = 0; i < nl; i++) {
        data[i].resize((size_t)lcount[i] * per);
        for (uint32_t b = 0; b < lcount[i]; b++) {
            unsigned long long* e = &data[i][(size_t)b * per];
            uint32_t k = 0;
            if (nalpha) e[k++] = ha0[b0 + b];
            if (nalpha > 1) e[k++] = ha1[b0 + b];
            if (c) e[k++] = hc[b0 + b];
        }
        b0 += lcount[i];
        ptrs[i] = data[i].data();
    }
    return crn2_write_dds(width, height, nlevels, faces, format, ptrs.data(), out, out_size);
}
