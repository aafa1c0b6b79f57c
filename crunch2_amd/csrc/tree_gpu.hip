// tree_gpu.hip -- crnlib's top-down VQ codebook generator
// (tree_clusterizer<vecNF>::generate_codebook, crnlib/crn_tree_clusterizer.h:
// 89-171, split_node :265-570) on gfx950.
//
// The reference splits the highest-variance node of a priority queue (ties to
// the lower node index) max_splits - 1 times; each split seeds two children
// from the furthest vector and the vector furthest from it, refines the seeds
// by the sign of the projection on the principal axis (ten power iterations
// of the weighted covariance), runs up to 1024 Lloyd 2-means passes and
// partitions the node's vectors stably (left first).  The codebook is every
// leaf's centroid in node index order.
//
// On the GPU:
//  * One workgroup of 256 threads splits one node.  Every float or double sum
//    over a node's vectors uses one fixed order: chunks of 4096 positions, in
//    a chunk thread t sums positions chunk + t + 256 j in order, the 256
//    thread sums are combined by a pairwise tree (t += t + s, s = 128 .. 1)
//    in LDS, and chunk sums are added in chunk order to totals that start at
//    zero.  Integer sums (weights, counts) use LDS atomics: exact.  The CPU
//    restatement uses the same order, so codebooks agree bit for bit.
//  * Which nodes get split depends only on the nodes' own data, and the pop
//    order is a strict total order, so the host replays the priority queue
//    exactly while the GPU computes splits ahead of it: whenever the queue's
//    top has no split result yet, every node among the first R of the queue
//    (R = splits still allowed) without one is split in one launch, one
//    workgroup each.  A node split ahead of time that is never popped stays a
//    leaf: its vectors were only reordered within its own range, and its
//    centroid is unchanged.
//  * Vectors stay in place (n x dim floats, with their weighted copies and
//    weighted dot products); a permutation of positions -> vector indices is
//    what the splits reorder.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <vector>

#include "crn2_internal.h"

namespace {

constexpr uint32_t kLanes = 256;
constexpr uint32_t kChunk = 4096;
constexpr uint32_t kMaxDim = 16;

struct TcData {
    const float* vec;       // n x D, by vector index
    const float* wvec;      // n x D: vec * (float)weight
    const double* wdot;     // n: (double)(dot(vec, vec) * (float)weight)
    const uint32_t* weight;
    uint32_t* perm;         // position -> vector index
    uint32_t* tmp;          // scratch for the partition, n
};

struct TcJob {
    uint32_t begin, end;
    uint64_t w;
    float c[kMaxDim];
};

struct TcOut {
    uint32_t status;        // 0 unsplittable, 1 split
    uint32_t mid;
    uint64_t lw, rw;
    float lvar, rvar;
    float lc[kMaxDim], rc[kMaxDim];
};

struct TcRoot {
    uint64_t w;
    double tt;
    float sum[kMaxDim];
};

template <int D>
struct CovTable {
    static constexpr int kEntries = D * (D + 1) / 2;
    uint8_t x[kEntries > 0 ? kEntries : 1], y[kEntries > 0 ? kEntries : 1];
    constexpr CovTable() : x(), y()
    {
        int q = 0;
        for (int a = 0; a < D; a++)
            for (int b = a; b < D; b++, q++) {
                x[q] = (uint8_t)a;
                y[q] = (uint8_t)b;
            }
    }
};

template <int D>
__device__ inline float sqdist(const float* a, const float* b)
{
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < D; i++) {
        const float t = a[i] - b[i];
        s += t * t;
    }
    return s;
}

template <int D>
__device__ inline float dotf(const float* a, const float* b)
{
    float s = a[0] * b[0];
#pragma unroll
    for (int i = 1; i < D; i++) s += a[i] * b[i];
    return s;
}

template <int D>
__device__ inline void load_vec(const float* __restrict__ p, float (&v)[D])
{
#pragma unroll
    for (int k = 0; k < D; k++) v[k] = p[k];
}

// Shared state of a split workgroup of GR lane groups (256 lanes each; a
// group handles one 4096-position chunk at a time).
template <int D, int GR>
struct TcShared {
    static constexpr int kCovCap = GR == 1 ? 34 : 12;     // covariance entries reduced at once
    static constexpr int kRedF = (CovTable<D>::kEntries < kCovCap ? CovTable<D>::kEntries : kCovCap) > 2 * D
                                     ? (CovTable<D>::kEntries < kCovCap ? CovTable<D>::kEntries : kCovCap)
                                     : 2 * D;
    float red[kRedF][kLanes * GR];
    double redd[2][kLanes * GR];
    float totf[kRedF > 2 * D ? kRedF : 2 * D];
    double totd[2];
    unsigned long long totu[4];
    float c[D], lc[D], rc[D], cl[D], cr[D], axis[D], seedf[D], seedo[D];
    float covar[D][D];
    float best_d[kLanes * GR / 64];
    uint32_t best_p[kLanes * GR / 64];
    uint32_t wl[kLanes * GR / 64], wr[kLanes * GR / 64];
    uint32_t pos;
    int go;
};

// Tree-combine, in every lane group, the K values its lanes left in
// red[k][group * 256 + lane] (and KD doubles), then add the groups' chunk sums
// to totf / totd in group order (= chunk order); groups >= nv had no chunk.
template <int D, int GR>
__device__ inline void tree_combine(TcShared<D, GR>& s, int K, int KD, uint32_t nv)
{
    const uint32_t tid = threadIdx.x;
    for (uint32_t sh = kLanes / 2; sh; sh >>= 1) {
        __syncthreads();
        for (uint32_t idx = tid; idx < GR * sh * (uint32_t)K; idx += kLanes * GR) {
            const uint32_t g = idx / (sh * K), r = idx % (sh * K), k = r / sh, t = r % sh + g * kLanes;
            s.red[k][t] = s.red[k][t] + s.red[k][t + sh];
        }
        for (uint32_t idx = tid; idx < GR * sh * (uint32_t)KD; idx += kLanes * GR) {
            const uint32_t g = idx / (sh * KD), r = idx % (sh * KD), k = r / sh, t = r % sh + g * kLanes;
            s.redd[k][t] = s.redd[k][t] + s.redd[k][t + sh];
        }
    }
    __syncthreads();
    if (tid < (uint32_t)K)
        for (uint32_t g = 0; g < nv; g++) s.totf[tid] = s.totf[tid] + s.red[tid][g * kLanes];
    if (tid < (uint32_t)KD)
        for (uint32_t g = 0; g < nv; g++) s.totd[tid] = s.totd[tid] + s.redd[tid][g * kLanes];
    __syncthreads();
}

// First position (lowest on ties) of the largest sqdist(vec, ref) in [b, e).
template <int D, int GR>
__device__ inline uint32_t furthest(const TcData& a, TcShared<D, GR>& s, uint32_t b, uint32_t e, const float* ref)
{
    constexpr uint32_t kT = kLanes * GR;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    float r[D];
#pragma unroll
    for (int k = 0; k < D; k++) r[k] = ref[k];
    float bd = -1.0f;
    uint32_t bp = 0xFFFFFFFFu;
    for (uint32_t i = b + tid; i < e; i += kT) {
        float v[D];
        load_vec<D>(a.vec + (size_t)a.perm[i] * D, v);
        const float dd = sqdist<D>(v, r);
        if (dd > bd) {
            bd = dd;
            bp = i;
        }
    }
    for (int off = 32; off; off >>= 1) {
        const float od = __shfl_xor(bd, off);
        const uint32_t op = (uint32_t)__shfl_xor((int)bp, off);
        if (od > bd || (od == bd && op < bp)) {
            bd = od;
            bp = op;
        }
    }
    if (lane == 0) {
        s.best_d[wave] = bd;
        s.best_p[wave] = bp;
    }
    __syncthreads();
    if (tid == 0) {
        float d0 = s.best_d[0];
        uint32_t p0 = s.best_p[0];
        for (uint32_t w = 1; w < kT / 64; w++)
            if (s.best_d[w] > d0 || (s.best_d[w] == d0 && s.best_p[w] < p0)) {
                d0 = s.best_d[w];
                p0 = s.best_p[w];
            }
        s.pos = p0;
    }
    __syncthreads();
    return s.pos;
}

// One node split per workgroup of GR lane groups: group g takes chunks
// it * GR + g, it = 0, 1, ...; within a chunk lane l sums positions
// chunk + l + 256 j, exactly as a single group would.
template <int D, int GR>
__global__ void __launch_bounds__(kLanes * GR) tc_split_kernel(TcData a, const TcJob* __restrict__ jobs,
                                                               TcOut* __restrict__ outs)
{
    constexpr uint32_t kT = kLanes * GR;
    __shared__ TcShared<D, GR> s;
    constexpr CovTable<D> ct;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t gl = tid & (kLanes - 1), grp = tid / kLanes;   // lane within its group, group
    const TcJob& job = jobs[blockIdx.x];
    const uint32_t b = job.begin, e = job.end;
    const uint32_t nchunks = (e - b + kChunk - 1) / kChunk, iters = (nchunks + GR - 1) / GR;
    if (tid < (uint32_t)D) s.c[tid] = job.c[tid];
    __syncthreads();

    // seeds
    {
        const uint32_t pf = furthest<D, GR>(a, s, b, e, s.c);
        if (tid < (uint32_t)D) s.seedf[tid] = a.vec[(size_t)a.perm[pf] * D + tid];
        __syncthreads();
        const uint32_t po = furthest<D, GR>(a, s, b, e, s.seedf);
        if (tid < (uint32_t)D) {
            s.seedo[tid] = a.vec[(size_t)a.perm[po] * D + tid];
            s.lc[tid] = (s.seedf[tid] + s.c[tid]) * .5f;
            s.rc[tid] = (s.seedo[tid] + s.c[tid]) * .5f;
        }
        __syncthreads();
    }

    float c[D];
#pragma unroll
    for (int k = 0; k < D; k++) c[k] = s.c[k];

    if (b + 2 < e) {
        // weighted covariance, upper triangle, in groups of at most kRedF entries
        constexpr int NE = CovTable<D>::kEntries;
        constexpr int G = TcShared<D, GR>::kRedF;
        constexpr int NG = (NE + G - 1) / G;
#pragma unroll
        for (int g = 0; g < NG; g++) {
            const int g0 = g * G;
            const int gn = NE - g0 < G ? NE - g0 : G;
            if (tid < (uint32_t)G) s.totf[tid] = 0.0f;
            for (uint32_t it = 0; it < iters; it++) {
                const uint32_t c0 = b + (it * GR + grp) * kChunk, c1 = min(e, c0 + kChunk);
                float acc[G];
#pragma unroll
                for (int q = 0; q < G; q++) acc[q] = 0.0f;
                for (uint32_t i = c0 + gl; i < c1; i += kLanes) {
                    const uint32_t vi = a.perm[i];
                    float v[D], w[D];
                    load_vec<D>(a.vec + (size_t)vi * D, v);
                    const float fw = (float)a.weight[vi];
#pragma unroll
                    for (int k = 0; k < D; k++) {
                        v[k] = v[k] - c[k];
                        w[k] = v[k] * fw;
                    }
#pragma unroll
                    for (int q = 0; q < G; q++)
                        if (q < gn) acc[q] = acc[q] + v[ct.x[(g0 + q) % NE]] * w[ct.y[(g0 + q) % NE]];
                }
#pragma unroll
                for (int q = 0; q < G; q++)
                    if (q < gn) s.red[q][tid] = acc[q];
                tree_combine<D, GR>(s, gn, 0, min(GR, nchunks - it * GR));
            }
            if (tid == 0) {
                const float divider = (float)job.w;
                for (int q = 0; q < gn; q++) {
                    const int x = ct.x[g0 + q], y = ct.y[g0 + q];
                    s.covar[x][y] = s.totf[q] / divider;
                    s.covar[y][x] = s.covar[x][y];
                }
            }
            __syncthreads();
        }
        if (tid == 0) {
            float axis[D];
            for (int i = 0; i < D; i++) axis[i] = 1.0f;
            for (int iter = 0; iter < 10; iter++) {
                float nx[D];
                double max_sum = 0;
                for (int i = 0; i < D; i++) {
                    double sum = 0;
                    for (int j = 0; j < D; j++) sum += axis[j] * s.covar[i][j];
                    nx[i] = (float)sum;
                    max_sum = i ? (sum > max_sum ? sum : max_sum) : sum;
                }
                if (max_sum != 0.0f) {
                    const float sc = (float)(1.0f / max_sum);
                    for (int i = 0; i < D; i++) nx[i] *= sc;
                }
                for (int i = 0; i < D; i++) axis[i] = nx[i];
            }
            double n2 = axis[0] * axis[0];
            for (int i = 1; i < D; i++) n2 += axis[i] * axis[i];
            if (n2 != 0) {
                const float sc = (float)(1.0f / sqrt(n2));
                for (int i = 0; i < D; i++) axis[i] *= sc;
            }
            for (int i = 0; i < D; i++) s.axis[i] = axis[i];
        }
        if (tid < (uint32_t)(2 * D)) s.totf[tid] = 0.0f;
        if (tid < 2) s.totd[tid] = 0.0;
        __syncthreads();
        float ax[D];
#pragma unroll
        for (int k = 0; k < D; k++) ax[k] = s.axis[k];
        for (uint32_t it = 0; it < iters; it++) {
            const uint32_t c0 = b + (it * GR + grp) * kChunk, c1 = min(e, c0 + kChunk);
            float fl[D], fr[D];
            double dl = 0.0, dr = 0.0;
#pragma unroll
            for (int k = 0; k < D; k++) fl[k] = fr[k] = 0.0f;
            for (uint32_t i = c0 + gl; i < c1; i += kLanes) {
                const uint32_t vi = a.perm[i];
                float v[D], wv[D];
                load_vec<D>(a.vec + (size_t)vi * D, v);
                load_vec<D>(a.wvec + (size_t)vi * D, wv);
#pragma unroll
                for (int k = 0; k < D; k++) v[k] = v[k] - c[k];
                const float fw = (float)a.weight[vi];
                if (dotf<D>(v, ax) < 0.0f) {
#pragma unroll
                    for (int k = 0; k < D; k++) fl[k] = fl[k] + wv[k];
                    dl += fw;
                } else {
#pragma unroll
                    for (int k = 0; k < D; k++) fr[k] = fr[k] + wv[k];
                    dr += fw;
                }
            }
#pragma unroll
            for (int k = 0; k < D; k++) {
                s.red[k][tid] = fl[k];
                s.red[D + k][tid] = fr[k];
            }
            s.redd[0][tid] = dl;
            s.redd[1][tid] = dr;
            tree_combine<D, GR>(s, 2 * D, 2, min(GR, nchunks - it * GR));
        }
        if (tid == 0 && s.totd[0] > 0.0f && s.totd[1] > 0.0f) {
            const float sl = (float)(1.0f / s.totd[0]), sr = (float)(1.0f / s.totd[1]);
            for (int k = 0; k < D; k++) {
                s.lc[k] = s.totf[k] * sl;
                s.rc[k] = s.totf[D + k] * sr;
            }
        }
        __syncthreads();
    }

    // Lloyd passes
    float prev_total = 1e+10f, lvar = 0.0f, rvar = 0.0f;   // meaningful in thread 0
    for (int loop = 0; loop < 1024; loop++) {
        if (tid < (uint32_t)D) {
            s.cl[tid] = s.lc[tid];
            s.cr[tid] = s.rc[tid];
        }
        if (tid < (uint32_t)(2 * D)) s.totf[tid] = 0.0f;
        if (tid < 2) s.totd[tid] = 0.0;
        if (tid < 4) s.totu[tid] = 0;
        __syncthreads();
        float cl[D], cr[D];
#pragma unroll
        for (int k = 0; k < D; k++) {
            cl[k] = s.cl[k];
            cr[k] = s.cr[k];
        }
        unsigned long long ul = 0, ur = 0, nl = 0, nr = 0;
        for (uint32_t it = 0; it < iters; it++) {
            const uint32_t c0 = b + (it * GR + grp) * kChunk, c1 = min(e, c0 + kChunk);
            float fl[D], fr[D];
            double dl = 0.0, dr = 0.0;
#pragma unroll
            for (int k = 0; k < D; k++) fl[k] = fr[k] = 0.0f;
            for (uint32_t i = c0 + gl; i < c1; i += kLanes) {
                const uint32_t vi = a.perm[i];
                float v[D], wv[D];
                load_vec<D>(a.vec + (size_t)vi * D, v);
                load_vec<D>(a.wvec + (size_t)vi * D, wv);
                const double wd = a.wdot[vi];
                const uint32_t w = a.weight[vi];
                if (sqdist<D>(cl, v) < sqdist<D>(cr, v)) {
#pragma unroll
                    for (int k = 0; k < D; k++) fl[k] = fl[k] + wv[k];
                    dl += wd;
                    ul += w;
                    nl++;
                } else {
#pragma unroll
                    for (int k = 0; k < D; k++) fr[k] = fr[k] + wv[k];
                    dr += wd;
                    ur += w;
                    nr++;
                }
            }
#pragma unroll
            for (int k = 0; k < D; k++) {
                s.red[k][tid] = fl[k];
                s.red[D + k][tid] = fr[k];
            }
            s.redd[0][tid] = dl;
            s.redd[1][tid] = dr;
            tree_combine<D, GR>(s, 2 * D, 2, min(GR, nchunks - it * GR));
        }
        for (int off = 32; off; off >>= 1) {
            ul += (unsigned long long)__shfl_xor((long long)ul, off);
            ur += (unsigned long long)__shfl_xor((long long)ur, off);
            nl += (unsigned long long)__shfl_xor((long long)nl, off);
            nr += (unsigned long long)__shfl_xor((long long)nr, off);
        }
        if (lane == 0) {
            atomicAdd(&s.totu[0], ul);
            atomicAdd(&s.totu[1], ur);
            atomicAdd(&s.totu[2], nl);
            atomicAdd(&s.totu[3], nr);
        }
        __syncthreads();
        if (tid == 0) {
            const uint64_t lw = s.totu[0], rw = s.totu[1];
            int go = 0;
            if (!lw || !rw) {
                go = -1;
            } else {
                lvar = (float)(s.totd[0] - (dotf<D>(s.totf, s.totf) / (float)lw));
                rvar = (float)(s.totd[1] - (dotf<D>(s.totf + D, s.totf + D) / (float)rw));
                const float il = 1.0f / (float)lw, ir = 1.0f / (float)rw;
                for (int k = 0; k < D; k++) {
                    s.lc[k] = s.totf[k] * il;
                    s.rc[k] = s.totf[D + k] * ir;
                }
                const float total = lvar + rvar;
                if (total < .00001f) go = 0;
                else if (((prev_total - total) / total) < .00001f) go = 0;
                else {
                    prev_total = total;
                    go = 1;
                }
            }
            s.go = go;
        }
        __syncthreads();
        if (s.go != 1) break;
    }

    TcOut& o = outs[blockIdx.x];
    if (s.go < 0) {
        if (tid == 0) o.status = 0;
        return;
    }
    const uint32_t mid = b + (uint32_t)s.totu[2];

    // stable partition by the last pass's comparison
    float cl[D], cr[D];
#pragma unroll
    for (int k = 0; k < D; k++) {
        cl[k] = s.cl[k];
        cr[k] = s.cr[k];
    }
    uint32_t run_l = 0, run_r = 0;
    const uint64_t lt_mask = lane ? (~0ull >> (64 - lane)) : 0ull;
    for (uint32_t base = b; base < e; base += kT) {
        const uint32_t pos = base + tid;
        const bool valid = pos < e;
        uint32_t vi = 0;
        bool left = false;
        if (valid) {
            vi = a.perm[pos];
            float v[D];
            load_vec<D>(a.vec + (size_t)vi * D, v);
            left = sqdist<D>(cl, v) < sqdist<D>(cr, v);
        }
        const uint64_t ml = __ballot(valid && left), mr = __ballot(valid && !left);
        if (lane == 0) {
            s.wl[wave] = (uint32_t)__popcll(ml);
            s.wr[wave] = (uint32_t)__popcll(mr);
        }
        __syncthreads();
        uint32_t ol = run_l, or_ = run_r, tl = 0, tr = 0;
        for (uint32_t w = 0; w < kT / 64; w++) {
            if (w < wave) {
                ol += s.wl[w];
                or_ += s.wr[w];
            }
            tl += s.wl[w];
            tr += s.wr[w];
        }
        if (valid) {
            if (left) a.tmp[b + ol + (uint32_t)__popcll(ml & lt_mask)] = vi;
            else a.tmp[mid + or_ + (uint32_t)__popcll(mr & lt_mask)] = vi;
        }
        run_l += tl;
        run_r += tr;
        __syncthreads();
    }
    __threadfence_block();
    __syncthreads();
    for (uint32_t pos = b + tid; pos < e; pos += kT)
        a.perm[pos] = __hip_atomic_load(&a.tmp[pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);

    if (tid == 0) {
        o.status = 1;
        o.mid = mid;
        o.lw = s.totu[0];
        o.rw = s.totu[1];
        o.lvar = lvar;
        o.rvar = rvar;
        for (int k = 0; k < D; k++) {
            o.lc[k] = s.lc[k];
            o.rc[k] = s.rc[k];
        }
    }
}

template <int D>
__global__ void __launch_bounds__(256) tc_prep_kernel(const float* __restrict__ vec, const uint32_t* __restrict__ weight,
                                                     uint32_t n, float* __restrict__ wvec, double* __restrict__ wdot,
                                                     uint32_t* __restrict__ perm)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float v[D];
    load_vec<D>(vec + (size_t)i * D, v);
    const float fw = (float)weight[i];
#pragma unroll
    for (int k = 0; k < D; k++) wvec[(size_t)i * D + k] = v[k] * fw;
    wdot[i] = (double)(dotf<D>(v, v) * fw);
    perm[i] = i;
}

// Root sums over all positions (identity order) in the fixed chunk order.
template <int D>
__global__ void __launch_bounds__(kLanes) tc_root_kernel(TcData a, uint32_t n, TcRoot* __restrict__ out)
{
    __shared__ TcShared<D, 1> s;
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    if (tid < (uint32_t)D) s.totf[tid] = 0.0f;
    if (tid == 0) {
        s.totd[0] = 0.0;
        s.totu[0] = 0;
    }
    __syncthreads();
    unsigned long long uw = 0;
    for (uint32_t c0 = 0; c0 < n; c0 += kChunk) {
        const uint32_t c1 = min(n, c0 + kChunk);
        float f[D];
        double d = 0.0;
#pragma unroll
        for (int k = 0; k < D; k++) f[k] = 0.0f;
        for (uint32_t i = c0 + tid; i < c1; i += kLanes) {
            const uint32_t vi = a.perm[i];
#pragma unroll
            for (int k = 0; k < D; k++) f[k] = f[k] + a.wvec[(size_t)vi * D + k];
            d += a.wdot[vi];
            uw += a.weight[vi];
        }
#pragma unroll
        for (int k = 0; k < D; k++) s.red[k][tid] = f[k];
        s.redd[0][tid] = d;
        tree_combine<D, 1>(s, D, 1, 1);
    }
    for (int off = 32; off; off >>= 1) uw += (unsigned long long)__shfl_xor((long long)uw, off);
    if (lane == 0) atomicAdd(&s.totu[0], uw);
    __syncthreads();
    if (tid == 0) {
        out->w = s.totu[0];
        out->tt = s.totd[0];
        for (int k = 0; k < D; k++) out->sum[k] = s.totf[k];
    }
}

__global__ void __launch_bounds__(256) tc_leaf_kernel(const uint32_t* __restrict__ perm,
                                                     const uint32_t* __restrict__ pos_leaf, uint32_t n,
                                                     uint32_t* __restrict__ vector_leaf)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) vector_leaf[perm[i]] = pos_leaf[i];
}

struct HostNode {
    uint32_t begin, end;
    uint64_t w;
    float var;
    float c[kMaxDim];
    int left = -1, right = -1;
    int state = 0;          // 0 not split yet, 1 split result ready, 2 unsplittable
    TcOut out;
};

// Heap order of the reference's NodeInfo: larger variance first, then lower index.
struct Before {
    const std::vector<HostNode>* nodes;
    bool operator()(int a, int b) const     // true when a comes after b (std heap "less")
    {
        const float va = (*nodes)[a].var, vb = (*nodes)[b].var;
        if (va != vb) return va < vb;
        return a > b;
    }
};

// Dims whose split kernel also comes in a four-group (1024-thread) variant
// for large nodes: 2 and 6 (16 floats per lane need too much LDS for it).
template <int D>
constexpr bool kWide = D <= 6;

template <int D>
void launch_wide(const TcData& a, const TcJob* jobs, TcOut* outs, uint32_t n, hipStream_t st)
{
    if constexpr (D <= 6)
        hipLaunchKernelGGL((tc_split_kernel<D, 4>), dim3(n), dim3(kLanes * 4), 0, st, a, jobs, outs);
}

template <int D>
int tree_codebook(const float* vectors, const uint32_t* weights, uint32_t n, uint32_t max_splits, float* codebook,
                  uint32_t* codebook_size, uint32_t* vector_leaf, hipStream_t st)
{
    float* wvec = nullptr;
    double* wdot = nullptr;
    uint32_t* perm = nullptr;
    uint32_t* tmp = nullptr;
    TcJob* djobs = nullptr;
    TcOut* douts = nullptr;
    TcRoot* droot = nullptr;
    int rc = CRN2_OK;
    auto fail = [&](int code) {
        if (rc == CRN2_OK) rc = code;
        return rc;
    };
    const size_t nn = n;
    if (hipMallocAsync((void**)&wvec, nn * D * sizeof(float), st) != hipSuccess ||
        hipMallocAsync((void**)&wdot, nn * sizeof(double), st) != hipSuccess ||
        hipMallocAsync((void**)&perm, nn * sizeof(uint32_t), st) != hipSuccess ||
        hipMallocAsync((void**)&tmp, nn * sizeof(uint32_t), st) != hipSuccess ||
        hipMallocAsync((void**)&djobs, (size_t)max_splits * sizeof(TcJob), st) != hipSuccess ||
        hipMallocAsync((void**)&douts, (size_t)max_splits * sizeof(TcOut), st) != hipSuccess ||
        hipMallocAsync((void**)&droot, sizeof(TcRoot), st) != hipSuccess)
        fail(CRN2_ENOMEM);
    TcData a{vectors, wvec, wdot, weights, perm, tmp};
    std::vector<HostNode> nodes;
    std::vector<uint32_t> pos_leaf;
    TcRoot root;
    if (rc == CRN2_OK) {
        hipLaunchKernelGGL(tc_prep_kernel<D>, dim3((n + 255) / 256), dim3(256), 0, st, vectors, weights, n, wvec, wdot,
                           perm);
        hipLaunchKernelGGL(tc_root_kernel<D>, dim3(1), dim3(kLanes), 0, st, a, n, droot);
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(&root, droot, sizeof(root), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            fail(CRN2_EHIP);
        else if (!root.w)
            fail(CRN2_EINVAL);
    }
    if (rc == CRN2_OK) {
        nodes.reserve(2 * (size_t)max_splits + 1);
        HostNode r;
        r.begin = 0;
        r.end = n;
        r.w = root.w;
        float dd = root.sum[0] * root.sum[0];
        for (int k = 1; k < D; k++) dd += root.sum[k] * root.sum[k];
        r.var = (float)(root.tt - (dd / (float)root.w));
        const float inv = 1.0f / (float)root.w;
        for (int k = 0; k < D; k++) r.c[k] = root.sum[k] * inv;
        nodes.push_back(r);
        std::vector<int> heap{0};
        Before before{&nodes};
        std::vector<TcJob> jobs;
        std::vector<int> job_node;
        std::vector<TcOut> outs;
        uint32_t splits = 1;
        while (splits < max_splits && !heap.empty()) {
            const int top = heap.front();
            if (nodes[top].var <= 0.0f || nodes[top].begin + 1 == nodes[top].end) break;
            if (nodes[top].state == 0) {
                // split, in one launch, every node among the first R of the queue that has no result yet
                std::vector<int> order(heap);
                std::sort(order.begin(), order.end(), [&](int x, int y) { return before(y, x); });
                const size_t R = std::min<size_t>(order.size(), max_splits - splits);
                jobs.clear();
                job_node.clear();
                for (size_t i = 0; i < R; i++) {
                    const HostNode& h = nodes[order[i]];
                    if (h.state != 0 || h.var <= 0.0f || h.begin + 1 == h.end) continue;
                    TcJob j;
                    j.begin = h.begin;
                    j.end = h.end;
                    j.w = h.w;
                    for (int k = 0; k < (int)kMaxDim; k++) j.c[k] = k < D ? h.c[k] : 0.0f;
                    jobs.push_back(j);
                    job_node.push_back(order[i]);
                }
                // nodes of more than two chunks first, split by workgroups of
                // four lane groups (same sums, same order); the rest one group each
                {
                    std::vector<size_t> idx(jobs.size());
                    for (size_t i = 0; i < idx.size(); i++) idx[i] = i;
                    std::stable_partition(idx.begin(), idx.end(),
                                          [&](size_t i) { return jobs[i].end - jobs[i].begin > 2 * kChunk; });
                    std::vector<TcJob> j2(jobs.size());
                    std::vector<int> n2(jobs.size());
                    for (size_t i = 0; i < idx.size(); i++) {
                        j2[i] = jobs[idx[i]];
                        n2[i] = job_node[idx[i]];
                    }
                    jobs.swap(j2);
                    job_node.swap(n2);
                }
                size_t nbig = 0;
                while (nbig < jobs.size() && jobs[nbig].end - jobs[nbig].begin > 2 * kChunk) nbig++;
                if (!kWide<D>) nbig = 0;
                outs.resize(jobs.size());
                if (hipMemcpyAsync(djobs, jobs.data(), jobs.size() * sizeof(TcJob), hipMemcpyHostToDevice, st) !=
                    hipSuccess) {
                    fail(CRN2_EHIP);
                    break;
                }
                if (nbig) launch_wide<D>(a, djobs, douts, (uint32_t)nbig, st);
                if (jobs.size() > nbig)
                    hipLaunchKernelGGL((tc_split_kernel<D, 1>), dim3((uint32_t)(jobs.size() - nbig)), dim3(kLanes), 0,
                                       st, a, djobs + nbig, douts + nbig);
                if (hipGetLastError() != hipSuccess ||
                    hipMemcpyAsync(outs.data(), douts, outs.size() * sizeof(TcOut), hipMemcpyDeviceToHost, st) !=
                        hipSuccess ||
                    hipStreamSynchronize(st) != hipSuccess) {
                    fail(CRN2_EHIP);
                    break;
                }
                for (size_t i = 0; i < jobs.size(); i++) {
                    HostNode& h = nodes[job_node[i]];
                    h.out = outs[i];
                    h.state = outs[i].status ? 1 : 2;
                }
            }
            std::pop_heap(heap.begin(), heap.end(), before);
            heap.pop_back();
            splits++;
            if (nodes[top].state == 2) continue;
            const TcOut o = nodes[top].out;
            HostNode l, r2;
            l.begin = nodes[top].begin;
            l.end = o.mid;
            l.w = o.lw;
            l.var = o.lvar;
            r2.begin = o.mid;
            r2.end = nodes[top].end;
            r2.w = o.rw;
            r2.var = o.rvar;
            for (int k = 0; k < D; k++) {
                l.c[k] = o.lc[k];
                r2.c[k] = o.rc[k];
            }
            const int li = (int)nodes.size();
            nodes[top].left = li;
            nodes[top].right = li + 1;
            nodes.push_back(l);
            nodes.push_back(r2);
            heap.push_back(li);
            std::push_heap(heap.begin(), heap.end(), before);
            heap.push_back(li + 1);
            std::push_heap(heap.begin(), heap.end(), before);
        }
    }
    if (rc == CRN2_OK) {
        std::vector<float> cb;
        if (vector_leaf) pos_leaf.assign(n, 0);
        uint32_t m = 0;
        for (const HostNode& h : nodes) {
            if (h.left != -1) continue;
            cb.insert(cb.end(), h.c, h.c + D);
            if (vector_leaf) std::fill(pos_leaf.begin() + h.begin, pos_leaf.begin() + h.end, m);
            m++;
        }
        *codebook_size = m;
        if (hipMemcpyAsync(codebook, cb.data(), cb.size() * sizeof(float), hipMemcpyHostToDevice, st) != hipSuccess)
            fail(CRN2_EHIP);
        if (rc == CRN2_OK && vector_leaf) {
            if (hipMemcpyAsync(tmp, pos_leaf.data(), nn * sizeof(uint32_t), hipMemcpyHostToDevice, st) != hipSuccess)
                fail(CRN2_EHIP);
            else
                hipLaunchKernelGGL(tc_leaf_kernel, dim3((n + 255) / 256), dim3(256), 0, st, perm, tmp, n, vector_leaf);
        }
        if (rc == CRN2_OK && (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess))
            fail(CRN2_EHIP);
    }
    for (void* p : {(void*)wvec, (void*)wdot, (void*)perm, (void*)tmp, (void*)djobs, (void*)douts, (void*)droot})
        if (p) (void)hipFreeAsync(p, st);
    (void)hipStreamSynchronize(st);
    return rc;
}

}  // namespace

extern "C" int crn2_tree_codebook_device(const float* vectors, const uint32_t* weights, uint32_t n, uint32_t dim,
                                         uint32_t max_splits, float* codebook, uint32_t* codebook_size,
                                         uint32_t* vector_leaf, void* stream)
{
    if (!vectors || !weights || !n || !max_splits || !codebook || !codebook_size) return CRN2_EINVAL;
    hipStream_t st = static_cast<hipStream_t>(stream);
    switch (dim) {
    case 2: return tree_codebook<2>(vectors, weights, n, max_splits, codebook, codebook_size, vector_leaf, st);
    case 6: return tree_codebook<6>(vectors, weights, n, max_splits, codebook, codebook_size, vector_leaf, st);
    case 16: return tree_codebook<16>(vectors, weights, n, max_splits, codebook, codebook_size, vector_leaf, st);
    default: return CRN2_EUNSUPPORTED;
    }
}
