// mlp_gpu.hip -- learned DXT1 endpoint proposal: features, training kernels,
// BF16 matrix-core inference and the integer refinement that judges it.
//
// No reference counterpart: the learned proposal sits beside crnlib's
// dxt1_endpoint_optimizer (crnlib/crn_dxt1.cpp:2234).  The features restate the
// block statistics of compute_vectors / compute_pca (crn_dxt1.cpp:155-256); the
// refinement is the +-1 endpoint lattice of refine_solution / the 26-neighbour
// probe (crn_dxt1.cpp:601-643, 905-1016) judged by the exact DXT1 block error
// (crn_dxt1.cpp:1370-1592, weights 8/25/1 from crn_color.h:724-726).
//
// Model: 57 -> 128 -> 128 -> 7, tanh-form GELU, weights and layer inputs in
// bfloat16, sums in float32.  Training runs layer by layer through a strided
// FP32 GEMM (bf16 rounding of the operands on load), bias/GELU kernels, the
// loss kernel and Adam; inference is one fused kernel on
// v_mfma_f32_32x32x16_bf16 (a wave owns 32 blocks, weights staged in LDS).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "crn2_internal.h"

namespace {

constexpr uint32_t kIn = 57, kHid = 128, kOut = 7;
constexpr float kGeluC = 0.7978845608028654f;
constexpr float kGeluA = 0.044715f;
constexpr float kConfWeight = 0.05f;

__device__ inline float bf16_round(float x)
{
    uint32_t u = __float_as_uint(x);
    u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
    return __uint_as_float(u);
}

__device__ inline float gelu(float z)
{
    const float t = tanhf(kGeluC * (z + kGeluA * z * z * z));
    return 0.5f * z * (1.0f + t);
}

__device__ inline float gelu_grad(float z)
{
    const float t = tanhf(kGeluC * (z + kGeluA * z * z * z));
    const float du = kGeluC * (1.0f + 3.0f * kGeluA * z * z);
    return 0.5f * (1.0f + t) + 0.5f * z * (1.0f - t * t) * du;
}

// ------------------------------------------------------------- features
__device__ inline void block_features(const uint8_t* px, float* f)
{
    float c[16][3];
    float mean[3] = {0, 0, 0};
    for (int p = 0; p < 16; p++)
        for (int k = 0; k < 3; k++) {
            c[p][k] = (float)px[4 * p + k] / 255.0f;
            mean[k] += c[p][k];
        }
    for (int k = 0; k < 3; k++) mean[k] /= 16.0f;
    for (int p = 0; p < 16; p++)
        for (int k = 0; k < 3; k++) f[3 * p + k] = c[p][k] - 0.5f;
    for (int i = 0; i < 3; i++)
        for (int jj = 0; jj < 3; jj++) {
            float s = 0;
            for (int p = 0; p < 16; p++) s += (c[p][i] - mean[i]) * (c[p][jj] - mean[jj]);
            f[48 + 3 * i + jj] = s / 16.0f * 4.0f;
        }
}

__global__ void __launch_bounds__(256) mlp_features_kernel(const uint8_t* __restrict__ rgba, uint32_t n,
                                                           float* __restrict__ feat)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    block_features(rgba + (size_t)b * 64, feat + (size_t)b * kIn);
}

// ---------------------------------------------------------------- GEMM
// C(m, n) [+]= sum_k A(m, k) B(k, n), A(m, k) = A[m sam + k sak],
// B(k, n) = B[k sbk + n sbn], C row-major with leading dimension ldc.  64 x 64
// tiles, 16-deep K slices in LDS, 4 x 4 outputs per thread; grid.z splits K
// and adds its partial sums atomically (C must then start cleared).
struct GemmArgs {
    const float* A;
    const float* B;
    float* C;
    uint32_t M, N, K;
    int64_t sam, sak, sbk, sbn, ldc;
    int accumulate;      // add into C instead of overwriting it
    int round_a, round_b;
};

__global__ void __launch_bounds__(256) gemm_f32_kernel(GemmArgs g)
{
    __shared__ float As[16][64 + 1];
    __shared__ float Bs[16][64 + 1];
    const uint32_t tid = threadIdx.x;
    const uint32_t tx = tid & 15u, ty = tid >> 4;
    const uint32_t m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    const uint32_t ksplit = (g.K + gridDim.z - 1) / gridDim.z;
    const uint32_t kb = blockIdx.z * ksplit, ke = min(g.K, kb + ksplit);
    float acc[4][4] = {};
    for (uint32_t k0 = kb; k0 < ke; k0 += 16) {
        for (uint32_t e = tid; e < 1024; e += 256) {
            const uint32_t kk = e >> 6, mm = e & 63u;
            const uint32_t m = m0 + mm, k = k0 + kk;
            float a = 0, b = 0;
            if (m < g.M && k < ke) {
                a = g.A[(int64_t)m * g.sam + (int64_t)k * g.sak];
                if (g.round_a) a = bf16_round(a);
            }
            const uint32_t nn = n0 + mm;
            if (nn < g.N && k < ke) {
                b = g.B[(int64_t)k * g.sbk + (int64_t)nn * g.sbn];
                if (g.round_b) b = bf16_round(b);
            }
            As[kk][mm] = a;
            Bs[kk][mm] = b;
        }
        __syncthreads();
#pragma unroll
        for (uint32_t kk = 0; kk < 16; kk++) {
            float a[4], b[4];
            for (int i = 0; i < 4; i++) a[i] = As[kk][ty * 4 + i];
            for (int i = 0; i < 4; i++) b[i] = Bs[kk][tx * 4 + i];
            for (int i = 0; i < 4; i++)
                for (int jj = 0; jj < 4; jj++) acc[i][jj] = fmaf(a[i], b[jj], acc[i][jj]);
        }
        __syncthreads();
    }
    for (int i = 0; i < 4; i++) {
        const uint32_t m = m0 + ty * 4 + i;
        if (m >= g.M) continue;
        for (int jj = 0; jj < 4; jj++) {
            const uint32_t nn = n0 + tx * 4 + jj;
            if (nn >= g.N) continue;
            float* c = g.C + (int64_t)m * g.ldc + nn;
            if (gridDim.z > 1) atomicAdd(c, acc[i][jj]);
            else *c = g.accumulate ? *c + acc[i][jj] : acc[i][jj];
        }
    }
}

// ------------------------------------------------- bias, activation, loss
// Z = C + bias (row broadcast); with gelu, A = bf16(gelu(Z)), else A = Z.
__global__ void __launch_bounds__(256) bias_act_kernel(float* __restrict__ Z, float* __restrict__ A,
                                                       const float* __restrict__ bias, uint32_t rows, uint32_t cols,
                                                       int use_gelu)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (uint64_t)rows * cols) return;
    const float z = Z[i] + bias[i % cols];
    Z[i] = z;
    A[i] = use_gelu ? bf16_round(gelu(z)) : z;
}

__global__ void __launch_bounds__(256) gelu_back_kernel(const float* __restrict__ dA, const float* __restrict__ Z,
                                                        float* __restrict__ dZ, uint64_t n)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dZ[i] = dA[i] * gelu_grad(Z[i]);
}

// Y: rows x 7 outputs, T: rows x 6 targets.  dY = d loss / dY for the batch
// mean of squared error (outputs 0-5) + kConfWeight * logistic loss of
// output 6 against "all six rounded components within one step".
__global__ void __launch_bounds__(256) mlp_loss_kernel(const float* __restrict__ Y, const float* __restrict__ T,
                                                       const float* __restrict__ labels, uint32_t rows,
                                                       float* __restrict__ dY, float* __restrict__ loss)
{
    __shared__ float red[256];
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0;
    if (r < rows) {
        const float* y = Y + (size_t)r * kOut;
        const float* t = T + (size_t)r * 6;
        const float scale[6] = {31, 63, 31, 31, 63, 31};
        bool ok = true;
        const float inv = 1.0f / (float)rows;
        for (int k = 0; k < 6; k++) {
            const float d = y[k] - t[k];
            l += d * d;
            dY[(size_t)r * kOut + k] = 2.0f * inv * d;
            const float q = fminf(fmaxf(rintf(y[k] * scale[k]), 0.0f), scale[k]);
            ok &= fabsf(q - rintf(t[k] * scale[k])) <= 1.0f;
        }
        const float lab = labels ? labels[r] : (ok ? 1.0f : 0.0f);
        const float z = y[6];
        const float s = 1.0f / (1.0f + expf(-z));
        l += kConfWeight * (fmaxf(z, 0.0f) + log1pf(expf(-fabsf(z))) - lab * z);
        dY[(size_t)r * kOut + 6] = kConfWeight * inv * (s - lab);
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (uint32_t s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0] / (float)rows);
}

__global__ void __launch_bounds__(256) adam_kernel(float* __restrict__ w, const float* __restrict__ g,
                                                   float* __restrict__ m, float* __restrict__ v, uint32_t n, float lr,
                                                   float b1, float b2, float eps, float c1, float c2)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// ------------------------------------------------- matrix-core inference
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr uint32_t kInPad = 64;               // K of layer 1 (57 features + zeros)
constexpr uint32_t kOutPad = 32;              // N of layer 3 (7 outputs + zeros)
constexpr uint32_t kXs = kInPad + 8;          // LDS row strides (bf16 elements), padded against bank conflicts
constexpr uint32_t kHs = kHid + 8;
constexpr uint32_t kInferWaves = 4;
// LDS layout (bytes)
constexpr uint32_t kW1Ofs = 0;                                    // [128][kXs]  W1 transposed (n-major, k contiguous)
constexpr uint32_t kW2Ofs = kW1Ofs + kHid * kXs * 2;              // [128][kHs]
constexpr uint32_t kW3Ofs = kW2Ofs + kHid * kHs * 2;              // [32][kHs]
constexpr uint32_t kBOfs = kW3Ofs + kOutPad * kHs * 2;            // biases: 128 + 128 + 32 floats
constexpr uint32_t kWaveOfs = kBOfs + (2 * kHid + kOutPad) * 4;   // per wave: two [32][kHs] buffers
constexpr uint32_t kWaveBytes = 2 * 32 * kHs * 2;
constexpr uint32_t kInferLds = kWaveOfs + kInferWaves * kWaveBytes;

__device__ inline bf16x8 lds_bf16x8(const uint8_t* lds, uint32_t byte_ofs)
{
    return *reinterpret_cast<const bf16x8*>(lds + byte_ofs);
}

__device__ inline __bf16 to_bf16(float x)
{
    const uint32_t u = __float_as_uint(bf16_round(x)) >> 16;
    return __builtin_bit_cast(__bf16, (uint16_t)u);
}

// One layer for the wave's 32 rows: in [32][ks] bf16 (row stride `is`), out
// written as bf16(gelu(z)) into [32][os] or, for the last layer, z to `y`.
template <uint32_t KSTEPS, uint32_t NTILES, bool LAST>
__device__ inline void mfma_layer(const uint8_t* lds, uint32_t in_ofs, uint32_t is, uint32_t w_ofs, uint32_t ws,
                                  const float* bias, uint8_t* out_lds, uint32_t out_ofs, uint32_t os, float* y,
                                  uint32_t row0, uint32_t n)
{
    const uint32_t lane = threadIdx.x & 63u, g = lane >> 5, col = lane & 31u;
    for (uint32_t nt = 0; nt < NTILES; nt++) {
        f32x16 acc = {0};
#pragma unroll
        for (uint32_t ks = 0; ks < KSTEPS; ks++) {
            const uint32_t k = 16 * ks + 8 * g;
            const bf16x8 a = lds_bf16x8(lds, in_ofs + (col * is + k) * 2);
            const bf16x8 b = lds_bf16x8(lds, w_ofs + ((32 * nt + col) * ws + k) * 2);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        const uint32_t c = 32 * nt + col;
        const float bc = bias[c];
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const uint32_t r = 8 * (i >> 2) + 4 * g + (i & 3);
            const float z = acc[i] + bc;
            if (LAST) {
                if (c < kOut && row0 + r < n) y[(size_t)(row0 + r) * 8 + c] = z;
            } else {
                *reinterpret_cast<__bf16*>(out_lds + out_ofs + (r * os + c) * 2) = to_bf16(gelu(z));
            }
        }
    }
}

__global__ void __launch_bounds__(256) mlp_infer_kernel(const uint8_t* __restrict__ rgba, uint32_t n,
                                                        const float* __restrict__ theta, float* __restrict__ y)
{
    extern __shared__ __align__(16) uint8_t lds[];
    const uint32_t tid = threadIdx.x;
    // stage weights (bf16, transposed so a column of W is contiguous in K)
    const float* W1 = theta;
    const float* b1 = W1 + kIn * kHid;
    const float* W2 = b1 + kHid;
    const float* b2 = W2 + kHid * kHid;
    const float* W3 = b2 + kHid;
    const float* b3 = W3 + kHid * kOut;
    for (uint32_t e = tid; e < kHid * kXs; e += 256) {
        const uint32_t nn = e / kXs, k = e % kXs;
        reinterpret_cast<__bf16*>(lds + kW1Ofs)[e] = to_bf16(k < kIn ? W1[k * kHid + nn] : 0.0f);
    }
    for (uint32_t e = tid; e < kHid * kHs; e += 256) {
        const uint32_t nn = e / kHs, k = e % kHs;
        reinterpret_cast<__bf16*>(lds + kW2Ofs)[e] = to_bf16(k < kHid ? W2[k * kHid + nn] : 0.0f);
    }
    for (uint32_t e = tid; e < kOutPad * kHs; e += 256) {
        const uint32_t nn = e / kHs, k = e % kHs;
        reinterpret_cast<__bf16*>(lds + kW3Ofs)[e] = to_bf16(k < kHid && nn < kOut ? W3[k * kOut + nn] : 0.0f);
    }
    float* bl = reinterpret_cast<float*>(lds + kBOfs);
    for (uint32_t e = tid; e < 2 * kHid + kOutPad; e += 256)
        bl[e] = e < kHid ? b1[e] : e < 2 * kHid ? b2[e - kHid] : (e - 2 * kHid < kOut ? b3[e - 2 * kHid] : 0.0f);
    __syncthreads();

    const uint32_t wave = tid >> 6, lane = tid & 63u;
    const uint32_t bufA = kWaveOfs + wave * kWaveBytes, bufB = bufA + 32 * kHs * 2;
    const uint32_t tiles = (n + 31) / 32;
    for (uint32_t t = blockIdx.x * kInferWaves + wave; t < tiles; t += gridDim.x * kInferWaves) {
        const uint32_t row0 = t * 32;
        // features of the wave's 32 blocks into bufA as [32][kXs] bf16 (lanes 0-31, one block each)
        if (lane < 32) {
            float f[kIn];
            const uint32_t b = row0 + lane;
            if (b < n) block_features(rgba + (size_t)b * 64, f);
            else
                for (uint32_t k = 0; k < kIn; k++) f[k] = 0;
            __bf16* xr = reinterpret_cast<__bf16*>(lds + bufA) + lane * kXs;
            for (uint32_t k = 0; k < kInPad; k++) xr[k] = to_bf16(k < kIn ? f[k] : 0.0f);
        }
        __builtin_amdgcn_wave_barrier();
        mfma_layer<kInPad / 16, kHid / 32, false>(lds, bufA, kXs, kW1Ofs, kXs, bl, lds, bufB, kHs, nullptr, row0, n);
        __builtin_amdgcn_wave_barrier();
        mfma_layer<kHid / 16, kHid / 32, false>(lds, bufB, kHs, kW2Ofs, kHs, bl + kHid, lds, bufA, kHs, nullptr, row0,
                                                n);
        __builtin_amdgcn_wave_barrier();
        mfma_layer<kHid / 16, 1, true>(lds, bufA, kHs, kW3Ofs, kHs, bl + 2 * kHid, nullptr, 0, 0, y, row0, n);
        __builtin_amdgcn_wave_barrier();
    }
}

// --------------------------------------------------------- refinement
__device__ inline void expand565(uint32_t c, int* o)
{
    const int r = (c >> 11) & 31, g = (c >> 5) & 63, b = c & 31;
    o[0] = (r << 3) | (r >> 2);
    o[1] = (g << 2) | (g >> 4);
    o[2] = (b << 3) | (b >> 2);
}

// Exact perceptual error of the DXT1 block (c0, c1) with best selectors under
// the standard DXT1 decode: c0 > c1 four colours, c0 <= c1 three + black.
__device__ inline uint32_t dxt1_error(const int (*px)[3], uint32_t c0, uint32_t c1, uint32_t* sel_out)
{
    int p[4][3];
    expand565(c0, p[0]);
    expand565(c1, p[1]);
    const bool four = c0 > c1;
    for (int k = 0; k < 3; k++) {
        if (four) {
            p[2][k] = (2 * p[0][k] + p[1][k]) / 3;
            p[3][k] = (p[0][k] + 2 * p[1][k]) / 3;
        } else {
            p[2][k] = (p[0][k] + p[1][k]) / 2;
            p[3][k] = 0;
        }
    }
    uint32_t err = 0, sel = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        uint32_t best = 0xFFFFFFFFu, bs = 0;
        for (int s = 0; s < 4; s++) {
            const int dr = px[i][0] - p[s][0], dg = px[i][1] - p[s][1], db = px[i][2] - p[s][2];
            const uint32_t e = (uint32_t)(8 * dr * dr + 25 * dg * dg + db * db);
            if (e < best) {
                best = e;
                bs = (uint32_t)s;
            }
        }
        err += best;
        sel |= bs << (2 * i);
    }
    if (sel_out) *sel_out = sel;
    return err;
}

__device__ inline uint32_t pack565(const int* c) { return ((uint32_t)c[0] << 11) | ((uint32_t)c[1] << 5) | (uint32_t)c[2]; }

// The same error, but stops summing once it reaches `bound` (returns a value
// >= bound then): a candidate that cannot win costs only a few texels.
__device__ inline uint32_t dxt1_error_below(const int (*px)[3], uint32_t c0, uint32_t c1, uint32_t bound)
{
    int p[4][3];
    expand565(c0, p[0]);
    expand565(c1, p[1]);
    const bool four = c0 > c1;
    for (int k = 0; k < 3; k++) {
        if (four) {
            p[2][k] = (2 * p[0][k] + p[1][k]) / 3;
            p[3][k] = (p[0][k] + 2 * p[1][k]) / 3;
        } else {
            p[2][k] = (p[0][k] + p[1][k]) / 2;
            p[3][k] = 0;
        }
    }
    uint32_t err = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        uint32_t best = 0xFFFFFFFFu;
        for (int s = 0; s < 4; s++) {
            const int dr = px[i][0] - p[s][0], dg = px[i][1] - p[s][1], db = px[i][2] - p[s][2];
            best = min(best, (uint32_t)(8 * dr * dr + 25 * dg * dg + db * db));
        }
        err += best;
        if (err >= bound) break;
    }
    return err;
}

// Ordered-pair error of (a, b) as pair_error, bounded as dxt1_error_below.
__device__ inline uint32_t pair_error_below(const int (*px)[3], const int* a, const int* b, uint32_t bound)
{
    const uint32_t ca = pack565(a), cb = pack565(b);
    return dxt1_error_below(px, ca > cb ? ca : cb, ca > cb ? cb : ca, bound);
}

// Four-colour error of the unordered endpoint pair (a, b) (equal pairs use the
// three-colour decode); returns the ordered pair in (lo, hi).
__device__ inline uint32_t pair_error(const int (*px)[3], const int* a, const int* b, uint32_t& lo, uint32_t& hi)
{
    const uint32_t ca = pack565(a), cb = pack565(b);
    lo = ca > cb ? ca : cb;
    hi = ca > cb ? cb : ca;
    return dxt1_error(px, lo, hi, nullptr);
}

// +-1 single-channel steps from endpoints e while the exact error drops.
// Returns true if it stopped at a point no single step improves.
__device__ inline bool descend(const int (*px)[3], int (*e)[3], uint32_t& best, uint32_t max_iters, uint32_t& steps)
{
    const int lim[3] = {31, 63, 31};
    for (;;) {
        if (steps >= max_iters) return false;
        uint32_t cand = best;
        int bs = -1, bk = 0, bd = 0;
        for (int s = 0; s < 2; s++)
            for (int k = 0; k < 3; k++)
                for (int d = -1; d <= 1; d += 2) {
                    const int v = e[s][k] + d;
                    if (v < 0 || v > lim[k]) continue;
                    e[s][k] = v;
                    const uint32_t er = pair_error_below(px, e[0], e[1], cand);
                    e[s][k] = v - d;
                    if (er < cand) {
                        cand = er;
                        bs = s;
                        bk = k;
                        bd = d;
                    }
                }
        if (bs < 0) return true;
        for (int s = 0; s < 2; s++)      // static indices: e stays in registers
            for (int k = 0; k < 3; k++)
                if (s == bs && k == bk) e[s][k] += bd;
        best = cand;
        steps++;
    }
}

// Least-squares endpoints for the selectors of the four-colour block (lo, hi)
// (colour i = lo + w_i (hi - lo), w in {0, 1/3, 2/3, 1}), per channel in
// 8-bit space, rounded to 5:6:5.  False if the selectors use one weight only.
__device__ inline bool lsq_endpoints(const int (*px)[3], uint32_t lo, uint32_t hi, int (*e)[3])
{
    uint32_t sel;
    dxt1_error(px, lo, hi, &sel);
    const int u_of[4] = {0, 3, 1, 2};   // 3 w by selector
    int a11 = 0, a12 = 0, a22 = 0, b1[3] = {0, 0, 0}, b2[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const int u = u_of[(sel >> (2 * i)) & 3u], v = 3 - u;
        a11 += v * v;
        a12 += u * v;
        a22 += u * u;
        for (int k = 0; k < 3; k++) {
            b1[k] += v * px[i][k];
            b2[k] += u * px[i][k];
        }
    }
    const float det = (float)a11 * (float)a22 - (float)a12 * (float)a12;
    if (det <= 0.0f) return false;
    const float lim[3] = {31.0f, 63.0f, 31.0f};
    for (int k = 0; k < 3; k++) {
        const float A = 3.0f * ((float)a22 * (float)b1[k] - (float)a12 * (float)b2[k]) / det;
        const float B = 3.0f * ((float)a11 * (float)b2[k] - (float)a12 * (float)b1[k]) / det;
        e[0][k] = (int)fminf(fmaxf(rintf(A * lim[k] / 255.0f), 0.0f), lim[k]);
        e[1][k] = (int)fminf(fmaxf(rintf(B * lim[k] / 255.0f), 0.0f), lim[k]);
    }
    return true;
}

// Descent, then while the least-squares refit of the result's selectors
// lowers the error, descent again from the refit.
__device__ inline uint32_t refine_from(const int (*px)[3], int (*e)[3], uint32_t max_iters, bool refit,
                                       uint32_t& steps)
{
    uint32_t lo, hi;
    uint32_t best = pair_error(px, e[0], e[1], lo, hi);
    uint32_t used = 0;
    for (int round = 0;; round++) {
        if (!descend(px, e, best, max_iters, used) || !refit || round == 3) break;
        pair_error(px, e[0], e[1], lo, hi);
        if (lo == hi) break;
        int f[2][3];
        if (!lsq_endpoints(px, lo, hi, f)) break;
        const uint32_t ef = pair_error(px, f[0], f[1], lo, hi);
        if (ef >= best) break;
        best = ef;
        for (int s = 0; s < 2; s++)
            for (int k = 0; k < 3; k++) e[s][k] = f[s][k];
    }
    steps += used;
    return best;
}

// Best improving move among all 728 combinations of -1 / 0 / +1 on the six
// endpoint components (refine_solution's both-ends lattice,
// crnlib/crn_dxt1.cpp:601-643); false if none improves.
__device__ inline bool box_step(const int (*px)[3], int (*e)[3], uint32_t& best)
{
    const int lim[3] = {31, 63, 31};
    uint32_t cand = best;
    int bm = -1;
    for (int m = 0; m < 729; m++) {
        if (m == 364) continue;   // all zero
        int f[2][3];
        int r = m;
        bool ok = true;
        for (int s = 0; s < 2; s++)
            for (int k = 0; k < 3; k++) {
                const int v = e[s][k] + (r % 3) - 1;
                r /= 3;
                ok &= v >= 0 && v <= lim[k];
                f[s][k] = v;
            }
        if (!ok) continue;
        const uint32_t er = pair_error_below(px, f[0], f[1], cand);
        if (er < cand) {
            cand = er;
            bm = m;
        }
    }
    if (bm < 0) return false;
    for (int s = 0; s < 2; s++)
        for (int k = 0; k < 3; k++) {
            e[s][k] += (bm % 3) - 1;
            bm /= 3;
        }
    best = cand;
    return true;
}

// Principal-axis start: the extreme projections of the texels on the
// dominant eigenvector of their covariance (power iteration).
__device__ inline void pca_start(const int (*px)[3], int (*e)[3])
{
    float m[3] = {0, 0, 0};
#pragma unroll
    for (int i = 0; i < 16; i++)
        for (int k = 0; k < 3; k++) m[k] += (float)px[i][k];
    for (int k = 0; k < 3; k++) m[k] *= 1.0f / 16.0f;
    float c[6] = {0, 0, 0, 0, 0, 0};   // xx xy xz yy yz zz
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const float x = px[i][0] - m[0], y = px[i][1] - m[1], z = px[i][2] - m[2];
        c[0] += x * x;
        c[1] += x * y;
        c[2] += x * z;
        c[3] += y * y;
        c[4] += y * z;
        c[5] += z * z;
    }
    float v[3] = {1.0f, 1.0f, 1.0f};
    for (int it = 0; it < 8; it++) {
        const float a = c[0] * v[0] + c[1] * v[1] + c[2] * v[2];
        const float b = c[1] * v[0] + c[3] * v[1] + c[4] * v[2];
        const float d = c[2] * v[0] + c[4] * v[1] + c[5] * v[2];
        const float mx = fmaxf(fabsf(a), fmaxf(fabsf(b), fabsf(d)));
        if (mx <= 0.0f) break;
        v[0] = a / mx;
        v[1] = b / mx;
        v[2] = d / mx;
    }
    const float n2 = v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
    float tmin = 0.0f, tmax = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const float t = ((px[i][0] - m[0]) * v[0] + (px[i][1] - m[1]) * v[1] + (px[i][2] - m[2]) * v[2]) / n2;
        tmin = fminf(tmin, t);
        tmax = fmaxf(tmax, t);
    }
    const float lim[3] = {31.0f, 63.0f, 31.0f};
    for (int k = 0; k < 3; k++) {
        e[0][k] = (int)fminf(fmaxf(rintf((m[k] + tmax * v[k]) * lim[k] / 255.0f), 0.0f), lim[k]);
        e[1][k] = (int)fminf(fmaxf(rintf((m[k] + tmin * v[k]) * lim[k] / 255.0f), 0.0f), lim[k]);
    }
}

// Sorts the 16 texels by falling error under endpoints e (a fixed sorting
// network: the texels stay in registers); nibble i of *perm becomes the
// original index of texel i.
__device__ inline void order_texels(int (*px)[3], const int (*e)[3], uint64_t* perm)
{
    *perm = 0xFEDCBA9876543210ull;
    uint32_t lo, hi;
    {
        const uint32_t ca = pack565(e[0]), cb = pack565(e[1]);
        lo = ca > cb ? ca : cb;
        hi = ca > cb ? cb : ca;
    }
    int p[4][3];
    expand565(lo, p[0]);
    expand565(hi, p[1]);
    const bool four = lo > hi;
    for (int k = 0; k < 3; k++) {
        p[2][k] = four ? (2 * p[0][k] + p[1][k]) / 3 : (p[0][k] + p[1][k]) / 2;
        p[3][k] = four ? (p[0][k] + 2 * p[1][k]) / 3 : 0;
    }
    uint32_t key[16];
#pragma unroll
    for (int i = 0; i < 16; i++) {
        uint32_t best = 0xFFFFFFFFu;
        for (int s = 0; s < 4; s++) {
            const int dr = px[i][0] - p[s][0], dg = px[i][1] - p[s][1], db = px[i][2] - p[s][2];
            best = min(best, (uint32_t)(8 * dr * dr + 25 * dg * dg + db * db));
        }
        key[i] = best;
    }
    // Batcher's odd-even merge sort network for 16 elements, descending
#pragma unroll
    for (int pp = 1; pp < 16; pp <<= 1)
#pragma unroll
        for (int k = pp; k >= 1; k >>= 1)
#pragma unroll
            for (int j = k % pp; j + k < 16; j += 2 * k)
#pragma unroll
                for (int i = 0; i < k; i++)
                    if (i + j + k < 16 && (i + j) / (2 * pp) == (i + j + k) / (2 * pp)) {
                        const int a = i + j, b = i + j + k;
                        if (key[a] < key[b]) {
                            const uint32_t tk = key[a];
                            key[a] = key[b];
                            key[b] = tk;
                            for (int c = 0; c < 3; c++) {
                                const int tv = px[a][c];
                                px[a][c] = px[b][c];
                                px[b][c] = tv;
                            }
                            const uint64_t na = (*perm >> (4 * a)) & 15u, nb = (*perm >> (4 * b)) & 15u;
                            *perm ^= ((na ^ nb) << (4 * a)) | ((na ^ nb) << (4 * b));
                        }
                    }
}

// Start points, each refined, the lowest final error kept (earlier start on
// ties): bit 0 of `starts` = the network's proposal y (8 floats per block,
// components scaled to [0, 1]); bit 1 = the block's brightest and darkest
// texels (luma 2r + 4g + b, first extreme); bit 2 = the principal axis.
// Bit 3 adds the least-squares refit after each descent; bit 4 polishes the
// kept result on the both-ends lattice.
// out per block (16 bytes): u32 c0 | c1 << 16, u32 selectors, u32 error,
// u32 single steps taken (max_iters per start = stopped early)
__global__ void __launch_bounds__(256) refine_kernel(const uint8_t* __restrict__ rgba, uint32_t n,
                                                     const float* __restrict__ y, uint32_t max_iters, uint32_t starts,
                                                     uint32_t* __restrict__ out)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    int px[16][3];
#pragma unroll
    for (int i = 0; i < 16; i++)
        for (int k = 0; k < 3; k++) px[i][k] = rgba[(size_t)b * 64 + 4 * i + k];
    const int lim[3] = {31, 63, 31};
    const bool refit = (starts & 8u) != 0;
    uint32_t best = 0xFFFFFFFFu, steps = 0;
    int keep[2][3] = {{0, 0, 0}, {0, 0, 0}};
    auto consider = [&](int (*e)[3]) {
        uint32_t st = 0;
        const uint32_t eb = refine_from(px, e, max_iters, refit, st);
        steps = max(steps, st);
        if (eb < best) {
            best = eb;
            for (int s = 0; s < 2; s++)
                for (int k = 0; k < 3; k++) keep[s][k] = e[s][k];
        }
    };
    if (starts & 1u) {
        const float* yy = y + (size_t)b * 8;
        int e[2][3];
        for (int s = 0; s < 2; s++)
            for (int k = 0; k < 3; k++)
                e[s][k] = (int)fminf(fmaxf(rintf(yy[3 * s + k] * (float)lim[k]), 0.0f), (float)lim[k]);
        consider(e);
    }
    if (starts & 2u) {
        // brightest and darkest texels (first extreme), kept by value so that
        // the texels are only ever indexed statically
        int bl = -1, dl = 1 << 30, bc[3] = {0, 0, 0}, dc[3] = {0, 0, 0};
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const int l = 2 * px[i][0] + 4 * px[i][1] + px[i][2];
            if (l > bl) {
                bl = l;
                for (int k = 0; k < 3; k++) bc[k] = px[i][k];
            }
            if (l < dl) {
                dl = l;
                for (int k = 0; k < 3; k++) dc[k] = px[i][k];
            }
        }
        int e[2][3];
        for (int k = 0; k < 3; k++) {
            e[0][k] = (int)rintf((float)bc[k] * (float)lim[k] / 255.0f);
            e[1][k] = (int)rintf((float)dc[k] * (float)lim[k] / 255.0f);
        }
        consider(e);
    }
    if (starts & 4u) {
        int e[2][3];
        pca_start(px, e);
        consider(e);
    }
    uint64_t perm = 0;   // 0: texels in their own order
    if ((starts & 16u) && best) {
        // The lattice scans 728 candidates whose sums stop at the current
        // best: visit the texels in order of falling error under the kept
        // endpoints so that losing candidates stop early.  Only the order of
        // the integer sums changes (selectors are mapped back below), so
        // every result is the same.
        order_texels(px, keep, &perm);
    }
    if (starts & 16u) {
        // polish the best result on the both-ends lattice, each improvement
        // followed by the single-step descent (and refit)
        for (int it = 0; it < 4 && best; it++) {
            if (!box_step(px, keep, best)) break;
            uint32_t st = 0;
            best = refine_from(px, keep, max_iters, refit, st);
            steps = max(steps, st);
        }
    }
    uint32_t lo, hi, sel;
    pair_error(px, keep[0], keep[1], lo, hi);
    best = dxt1_error(px, lo, hi, &sel);
    if (perm) {
        uint32_t os = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) os |= ((sel >> (2 * i)) & 3u) << (2 * ((perm >> (4 * i)) & 15u));
        sel = os;
    }
    out[(size_t)b * 4 + 0] = lo | (hi << 16);
    out[(size_t)b * 4 + 1] = sel;
    out[(size_t)b * 4 + 2] = best;
    out[(size_t)b * 4 + 3] = steps;
}

// Error of given DXT1 blocks (c0 | c1 << 16 per block) under the same decode.
__global__ void __launch_bounds__(256) block_error_kernel(const uint8_t* __restrict__ rgba, uint32_t n,
                                                          const uint32_t* __restrict__ ends, uint32_t stride,
                                                          uint32_t* __restrict__ err, uint32_t* __restrict__ sel)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= n) return;
    int px[16][3];
    for (int i = 0; i < 16; i++)
        for (int k = 0; k < 3; k++) px[i][k] = rgba[(size_t)b * 64 + 4 * i + k];
    const uint32_t e = ends[(size_t)b * stride];
    uint32_t s;
    err[b] = dxt1_error(px, e & 0xFFFFu, e >> 16, &s);
    if (sel) sel[b] = s;
}

inline uint32_t grid1(uint64_t n, uint32_t tb) { return (uint32_t)((n + tb - 1) / tb); }

}  // namespace

extern "C" {

int crn2_mlp_features_device(const void* rgba, uint32_t n, float* feat, void* stream)
{
    if (!n) return CRN2_OK;
    if (!rgba || !feat) return CRN2_EINVAL;
    hipLaunchKernelGGL(mlp_features_kernel, dim3(grid1(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const uint8_t*>(rgba), n, feat);
    return hipGetLastError() == hipSuccess ? CRN2_OK : CRN2_EHIP;
}

int crn2_gemm_f32_device(uint32_t M, uint32_t N, uint32_t K, const float* A, int64_t sam, int64_t sak, const float* B,
                         int64_t sbk, int64_t sbn, float* C, int64_t ldc, int accumulate, int round_a, int round_b,
                         uint32_t ksplit, void* stream)
{
    if (!M || !N) return CRN2_OK;
    if (!A || !B || !C || !ksplit || ksplit > 1024) return CRN2_EINVAL;
    GemmArgs g{A, B, C, M, N, K, sam, sak, sbk, sbn, ldc, accumulate, round_a, round_b};
    hipLaunchKernelGGL(gemm_f32_kernel, dim3((N + 63) / 64, (M + 63) / 64, ksplit), dim3(256), 0,
                       static_cast<hipStream_t>(stream), g);
    return hipGetLastError() == hipSuccess ? CRN2_OK : CRN2_EHIP;
}

int crn2_mlp_bias_act_device(float* Z, float* A, const float* bias, uint32_t rows, uint32_t cols, int use_gelu,
                             void* stream)
{
    const uint64_t n = (uint64_t)rows * cols;
    if (!n) return CRN2_OK;
    hipLaunchKernelGGL(bias_act_kernel, dim3(grid1(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), Z, A,
                       bias, rows, cols, use_gelu);
    return hipGetLastError() == hipSuccess ? CRN2_OK : CRN2_EHIP;
}

int crn2_mlp_gelu_back_device(const float* dA, const float* Z, float* dZ, uint64_t n, void* stream)
{
    if (!n) return CRN2_OK;
    hipLaunchKernelGGL(gelu_back_kernel, dim3(grid1(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), dA, Z,
                       dZ, n);
    return hipGetLastError() == hipSuccess ? CRN2_OK : CRN2_EHIP;
}

int crn2_mlp_loss_device(const float* Y, const float* T, const float* labels, uint32_t rows, float* dY, float* loss,
                         void* stream)
{
    if (!rows) return CRN2_OK;
    hipLaunchKernelGGL(mlp_loss_kernel, dim3(grid1(rows, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), Y, T,
                       labels, rows, dY, loss);
    return hipGetLastError() == hipSuccess ? CRN2_OK : CRN2_EHIP;
}

int crn2_adam_device(float* w, const float* g, float* m, float* v, uint32_t n, float lr, float b1, float b2, float eps,
                     uint32_t step, void* stream)
{
    if (!n) return CRN2_OK;
    const float c1 = 1.0f - powf(b1, (float)step), c2 = 1.0f - powf(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(grid1(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream), w, g, m, v, n,
                       lr, b1, b2, eps, c1, c2);
    return hipGetLastError() == hipSuccess ? CRN2_OK : CRN2_EHIP;
}

int crn2_mlp_infer_device(const void* rgba, uint32_t n, const float* theta, float* y, void* stream)
{
    if (!n) return CRN2_OK;
    if (!rgba || !theta || !y) return CRN2_EINVAL;
    // the LDS limit is a property of the kernel on each device: set it once per device
    static bool attr[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return CRN2_EHIP;
    if (!attr[dev]) {
        if (hipFuncSetAttribute(reinterpret_cast<const void*>(&mlp_infer_kernel),
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)kInferLds) != hipSuccess)
            return CRN2_EHIP;
        attr[dev] = true;
    }
    const uint32_t tiles = (n + 31) / 32;
    const uint32_t wgs = std::min<uint32_t>((tiles + kInferWaves - 1) / kInferWaves, 1024);
    hipLaunchKernelGGL(mlp_infer_kernel, dim3(wgs), dim3(256), kInferLds, static_cast<hipStream_t>(stream),
                       static_cast<const uint8_t*>(rgba), n, theta, y);
    return hipGetLastError() == hipSuccess ? CRN2_OK : CRN2_EHIP;
}

int crn2_dxt1_refine_device(const void* rgba, uint32_t n, const float* y, uint32_t max_iters, uint32_t starts,
                            uint32_t* out, void* stream)
{
    if (!n) return CRN2_OK;
    if (!rgba || !out || !(starts & 7u) || (starts & ~31u) || ((starts & 1u) && !y)) return CRN2_EINVAL;
    hipLaunchKernelGGL(refine_kernel, dim3(grid1(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const uint8_t*>(rgba), n, y, max_iters, starts, out);
    return hipGetLastError() == hipSuccess ? CRN2_OK : CRN2_EHIP;
}

int crn2_dxt1_block_error_device(const void* rgba, uint32_t n, const uint32_t* endpoints, uint32_t stride_u32,
                                 uint32_t* err, uint32_t* selectors, void* stream)
{
    if (!n) return CRN2_OK;
    if (!rgba || !endpoints || !err || !stride_u32) return CRN2_EINVAL;
    hipLaunchKernelGGL(block_error_kernel, dim3(grid1(n, 256)), dim3(256), 0, static_cast<hipStream_t>(stream),
                       static_cast<const uint8_t*>(rgba), n, endpoints, stride_u32, err, selectors);
    return hipGetLastError() == hipSuccess ? CRN2_OK : CRN2_EHIP;
}

}  // extern "C"
