// dxt1c_gpu.hip -- DXT1 endpoint search for pixel clusters of any size, one
// wave per cluster (the dxt_hc endpoint-cluster case).
//
// Reference: dxt1_endpoint_optimizer::compute (crnlib/crn_dxt1.cpp:2081-2246)
// at uber quality.  Without 3-colour blocks (the path dxt_hc uses for its
// endpoint clusters, determine_color_endpoint_codebook_task,
// crn_dxt_hc.cpp:663-754) the hc evaluators (evaluate_solution_hc_*,
// compute_selectors_hc, crn_dxt1.cpp:1759-1869) judge every candidate; with
// them (qdxt1's clusters, crn_qdxt1.cpp:616-626) evaluate_solution_uber
// (:1370-1592) scores both block types, and the solid-block tries, the
// one-colour shortcut and return_solution take their 3-colour branches
// (:119, :143, :2220, :263-365).  Unique colours (first-appearance order), the error planes,
// PCA, the probe sweeps, refine_solution, try_median4,
// optimize_endpoint_comps and the special cases follow crn_dxt1.cpp as in the
// per-block kernel (dxt1_gpu.hip).
//
// Layout: one workgroup of one wave per cluster; the cluster's unique colours,
// weights, selectors, colour hash and per-channel error tables live in LDS.
// Control flow is wave-uniform (every lane holds the same search state); the
// O(colours) integer loops (candidate errors, selectors, planes, the LSQ and
// per-channel sums) are split across the lanes and reduced with wave
// shuffles, which is exact for integer sums.  Float sums (means, covariance,
// 4-means) keep the reference's sequential order so the float trajectory
// matches the CPU restatement.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <mutex>
#include <type_traits>

#include "crn2_internal.h"

namespace {

// Largest cluster: 4096 unique colours.  The kernel comes in four LDS sizes
// (512, 1024, 2048 and 4096 unique colours: 18, 28, 49 and 89 KiB) so that
// small clusters, the common case, run eight, five or three waves per CU
// instead of one.
constexpr uint32_t kMaxU = 4096;

__device__ __constant__ int16_t c_uber_probe[15] = {0, 1, 2, 3, 5, 7, 9, 10, 13, 15, 19, 27, 43, 59, 91};
__device__ __constant__ int8_t c_adj[26][3] = {
    {-1, -1, -1}, {0, -1, -1}, {1, -1, -1}, {-1, 0, -1}, {0, 0, -1}, {1, 0, -1}, {-1, 1, -1}, {0, 1, -1},
    {1, 1, -1}, {-1, -1, 0}, {0, -1, 0}, {1, -1, 0}, {-1, 0, 0}, {1, 0, 0}, {-1, 1, 0}, {0, 1, 0},
    {1, 1, 0}, {-1, -1, 1}, {0, -1, 1}, {1, -1, 1}, {-1, 0, 1}, {0, 0, 1}, {1, 0, 1}, {-1, 1, 1},
    {0, 1, 1}, {1, 1, 1}};
__device__ uint8_t g_c_om5[256][2], g_c_om6[256][2], g_c_om5_3[256][2], g_c_om6_3[256][2];

template <uint32_t MaxU>
struct LdsT {
    static constexpr uint32_t kCapU = MaxU;        // unique colours
    static constexpr uint32_t kCapH = 2 * MaxU;    // colour hash slots (power of two)
    uint32_t col[MaxU];      // r | g << 8 | b << 16
    uint32_t w[MaxU];
    uint8_t sel[MaxU];
    uint32_t hkey[kCapH];    // colour | 0xFF000000, 0 = empty
    uint16_t hval[kCapH];
    unsigned long long pl[3][64][3];   // error planes: count, sum, sum of squares per quantised value
    uint64_t rlo[32], rhi[32], glo[64], ghi[64], blo[32], bhi[32];
};

// The store of the size classes whose texel count never exceeds the colour
// capacity (512, 1024, 2048): no colour hash.  The unique-colour pass sorts
// (colour, texel index) pairs in a buffer that overlays the search's tables
// (dead until the search starts), and a per-texel unique index replaces the
// hash lookup of the final selectors: 11, 16 and 29 KiB instead of 18, 28
// and 49, so more clusters run per CU.
template <uint32_t MaxU>
struct LdsS {
    static constexpr uint32_t kCapU = MaxU;
    static constexpr uint32_t kSearch = MaxU + 8 * (3 * 64 * 3) + 8 * (32 + 32 + 64 + 64 + 32 + 32);
    static constexpr uint32_t kSort = 6 * MaxU;   // u32 colours + u16 texel indices
    uint32_t col[MaxU];
    uint16_t w[MaxU];
    uint16_t tmap[MaxU];                 // texel -> unique colour
    uint32_t fbits[MaxU / 32];           // first texel of each colour
    uint32_t fbase[MaxU / 32];           // first texels before each word of fbits
    uint8_t sel[MaxU];                   // from here on: the sort buffers' space
    unsigned long long pl[3][64][3];
    uint64_t rlo[32], rhi[32], glo[64], ghi[64], blo[32], bhi[32];
    uint8_t pad[kSort > kSearch ? kSort - kSearch : 4];
};

// A colour store tagged with the evaluator: kUber -- evaluate_solution_uber
// with both block types (3-colour blocks allowed); else the hc evaluators.
template <class Base, bool U>
struct LdsU : Base {
    static constexpr bool kUber = U;
};

struct Best {
    uint32_t lo, hi;
    uint64_t err;
    int alt, enforce, enforced, alpha;
};

struct St {
    int perceptual;
    int have_alpha;      // DXT1A: texels below the alpha threshold are transparent (3-colour blocks only)
    uint32_t nu, total_w;
    float mean[3], meanw[3], pw[3], axis[3];
    Best best;
    bool sel_valid;      // L.sel holds the selectors of the current best solution
    bool sel_nonempty;   // selectors have been computed at least once (results::m_selectors not empty)
};

__device__ inline uint32_t lane_id() { return threadIdx.x & 63u; }

__device__ inline uint64_t wave_sum(uint64_t v)
{
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, off);
        const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), off);
        v += ((uint64_t)hi << 32) | lo;
    }
    return v;
}

__device__ inline uint32_t hash32(uint32_t x)
{
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return x;
}

__device__ inline void unpack565(uint32_t p, int scaled, int& r, int& g, int& b)
{
    uint32_t bb = p & 31u, gg = (p >> 5) & 63u, rr = (p >> 11) & 31u;
    if (scaled) {
        bb = (bb << 3) | (bb >> 2);
        gg = (gg << 2) | (gg >> 4);
        rr = (rr << 3) | (rr >> 2);
    }
    r = (int)rr;
    g = (int)gg;
    b = (int)bb;
}

__device__ inline uint32_t pack565(int r, int g, int b, int scaled)
{
    uint32_t R = (uint32_t)min(max(r, 0), 255), G = (uint32_t)min(max(g, 0), 255), B = (uint32_t)min(max(b, 0), 255);
    if (scaled) {
        R = (R * 31u + 127u) / 255u;
        G = (G * 63u + 127u) / 255u;
        B = (B * 31u + 127u) / 255u;
    }
    R = min(R, 31u);
    G = min(G, 63u);
    B = min(B, 31u);
    return B | (G << 5) | (R << 11);
}

__device__ inline void canon(uint32_t& lo, uint32_t& hi)
{
    if (lo < hi) {
        const uint32_t t = lo;
        lo = hi;
        hi = t;
    }
}

__device__ inline uint32_t cdist(int perceptual, uint32_t px, int r, int g, int b)
{
    const int dr = (int)(px & 255u) - r, dg = (int)((px >> 8) & 255u) - g, db = (int)((px >> 16) & 255u) - b;
    return perceptual ? (uint32_t)(dr * dr * 8 + dg * dg * 25 + db * db) : (uint32_t)(dr * dr + dg * dg + db * db);
}

struct Pal {
    int r[4], g[4], b[4];
};

__device__ inline Pal palette4(uint32_t lo, uint32_t hi, int alt)
{
    Pal p;
    unpack565(lo, 1, p.r[0], p.g[0], p.b[0]);
    unpack565(hi, 1, p.r[1], p.g[1], p.b[1]);
    p.r[2] = (p.r[0] * 2 + p.r[1] + alt) / 3;
    p.g[2] = (p.g[0] * 2 + p.g[1] + alt) / 3;
    p.b[2] = (p.b[0] * 2 + p.b[1] + alt) / 3;
    p.r[3] = (p.r[1] * 2 + p.r[0] + alt) / 3;
    p.g[3] = (p.g[1] * 2 + p.g[0] + alt) / 3;
    p.b[3] = (p.b[1] * 2 + p.b[0] + alt) / 3;
    return p;
}

// get_block_colors3 with the alternate rounding (crn_dxt1.cpp:1432-1440)
__device__ inline Pal palette3(uint32_t lo, uint32_t hi, int alt)
{
    Pal p;
    unpack565(lo, 1, p.r[0], p.g[0], p.b[0]);
    unpack565(hi, 1, p.r[1], p.g[1], p.b[1]);
    p.r[2] = (p.r[0] + p.r[1] + alt) >> 1;
    p.g[2] = (p.g[0] + p.g[1] + alt) >> 1;
    p.b[2] = (p.b[0] + p.b[1] + alt) >> 1;
    p.r[3] = p.g[3] = p.b[3] = 0;
    return p;
}

// evaluate_solution_hc (crn_dxt1.cpp:1759-1835): sum over colours of the
// weighted distance to the nearest of the four colours.
template <class Lds>
__device__ bool eval_hc(St& s, const Lds& L, uint32_t lo, uint32_t hi, int alt)
{
    const Pal p = palette4(lo, hi, alt);
    uint64_t err = 0;
    for (uint32_t i = lane_id(); i < s.nu; i += 64) {
        const uint32_t px = L.col[i];
        const uint32_t e0 = cdist(s.perceptual, px, p.r[0], p.g[0], p.b[0]);
        const uint32_t e1 = cdist(s.perceptual, px, p.r[1], p.g[1], p.b[1]);
        const uint32_t e2 = cdist(s.perceptual, px, p.r[2], p.g[2], p.b[2]);
        const uint32_t e3 = cdist(s.perceptual, px, p.r[3], p.g[3], p.b[3]);
        err += (uint64_t)min(min(e0, e1), min(e2, e3)) * L.w[i];
    }
    err = wave_sum(err);
    if (err >= s.best.err) return false;
    s.best.lo = lo;
    s.best.hi = hi;
    s.best.err = err;
    s.best.alt = alt;
    s.best.enforce = lo == hi;
    if (s.best.enforce) {
        if ((s.best.lo & 31) != 31) {
            s.best.lo++;
            s.best.enforced = 1;
        } else {
            s.best.hi--;
            s.best.enforced = 0;
        }
    }
    s.sel_valid = false;
    return true;
}

// evaluate_solution_uber (crn_dxt1.cpp:1370-1592), both block types: the
// 4-colour error, then (alpha blocks allowed) the 3-colour error, each
// kept when below the best so far (exact integer sums over the colours);
// a 4-colour pair with equal endpoints enforces one selector.  The selectors
// of the kept type are recomputed on demand (compute_selectors).
template <class Lds>
__device__ bool eval_uber(St& s, const Lds& L, uint32_t lo, uint32_t hi, int alt)
{
    const Pal p4 = palette4(lo, hi, alt), p3 = palette3(lo, hi, alt);
    uint64_t e4 = 0, e3 = 0;
    for (uint32_t i = lane_id(); i < s.nu; i += 64) {
        const uint32_t px = L.col[i];
        const uint32_t d0 = cdist(s.perceptual, px, p4.r[0], p4.g[0], p4.b[0]);
        const uint32_t d1 = cdist(s.perceptual, px, p4.r[1], p4.g[1], p4.b[1]);
        const uint32_t d2 = cdist(s.perceptual, px, p4.r[2], p4.g[2], p4.b[2]);
        const uint32_t d3 = cdist(s.perceptual, px, p4.r[3], p4.g[3], p4.b[3]);
        e4 += (uint64_t)min(min(d0, d1), min(d2, d3)) * L.w[i];
        {
            const uint32_t t2 = cdist(s.perceptual, px, p3.r[2], p3.g[2], p3.b[2]);
            e3 += (uint64_t)min(min(d0, d1), t2) * L.w[i];
        }
    }
    e4 = wave_sum(e4);
    e3 = wave_sum(e3);
    uint64_t terr = s.best.err;
    int talpha = 0;
    bool took = false;
    if (!s.have_alpha && e4 < terr) {
        terr = e4;
        took = true;
    }
    if (e3 < terr) {
        terr = e3;
        talpha = 1;
        took = true;
    }
    if (!took || terr >= s.best.err) return false;
    s.best.lo = lo;
    s.best.hi = hi;
    s.best.err = terr;
    s.best.alpha = talpha;
    s.best.alt = alt;
    s.best.enforce = !talpha && lo == hi;
    if (s.best.enforce) {
        if ((s.best.lo & 31) != 31) {
            s.best.lo++;
            s.best.enforced = 1;
        } else {
            s.best.hi--;
            s.best.enforced = 0;
        }
    }
    s.sel_valid = false;
    s.sel_nonempty = true;
    return true;
}

// compute_selectors_hc (crn_dxt1.cpp:1837-1869)
template <class Lds>
__device__ void compute_selectors(St& s, Lds& L)
{
    if (s.sel_valid) return;
    if (s.best.enforce) {
        for (uint32_t i = lane_id(); i < s.nu; i += 64) L.sel[i] = (uint8_t)s.best.enforced;
    } else if (Lds::kUber && s.best.alpha) {
        // the 3-colour palette's first nearest entry (evaluate_solution_uber)
        const Pal p = palette3(s.best.lo, s.best.hi, s.best.alt);
        for (uint32_t i = lane_id(); i < s.nu; i += 64) {
            const uint32_t px = L.col[i];
            const uint32_t e0 = cdist(s.perceptual, px, p.r[0], p.g[0], p.b[0]);
            const uint32_t e1 = cdist(s.perceptual, px, p.r[1], p.g[1], p.b[1]);
            const uint32_t e2 = cdist(s.perceptual, px, p.r[2], p.g[2], p.b[2]);
            L.sel[i] = (uint8_t)(e1 < e0 ? (e2 < e1 ? 2 : 1) : (e2 < e0 ? 2 : 0));
        }
    } else {
        const Pal p = palette4(s.best.lo, s.best.hi, s.best.alt);
        for (uint32_t i = lane_id(); i < s.nu; i += 64) {
            const uint32_t px = L.col[i];
            const uint32_t e0 = cdist(s.perceptual, px, p.r[0], p.g[0], p.b[0]);
            const uint32_t e1 = cdist(s.perceptual, px, p.r[1], p.g[1], p.b[1]);
            const uint32_t e2 = cdist(s.perceptual, px, p.r[2], p.g[2], p.b[2]);
            const uint32_t e3 = cdist(s.perceptual, px, p.r[3], p.g[3], p.b[3]);
            const uint32_t e01 = min(e0, e1), e23 = min(e2, e3);
            L.sel[i] = (uint8_t)(e01 <= e23 ? (e01 == e0 ? 0 : 1) : (e23 == e2 ? 2 : 3));
        }
    }
    __builtin_amdgcn_wave_barrier();
    s.sel_valid = true;
    s.sel_nonempty = true;
}

// evaluate_solution with the plane lower bound (crn_dxt1.cpp:1312-1340); the
// bound only skips candidates that cannot win.
template <class Lds>
__device__ bool evaluate(St& s, Lds& L, uint32_t lo, uint32_t hi, int alt)
{
    int ar, ag, ab, br, bg, bb;
    unpack565(lo, 0, ar, ag, ab);
    unpack565(hi, 0, br, bg, bb);
    const uint64_t re = ar < br ? L.rlo[ar] + L.rhi[br] : L.rhi[ar] + L.rlo[br];
    const uint64_t ge = ag < bg ? L.glo[ag] + L.ghi[bg] : L.ghi[ag] + L.glo[bg];
    const uint64_t be = ab < bb ? L.blo[ab] + L.bhi[bb] : L.bhi[ab] + L.blo[bb];
    if (re + ge + be >= s.best.err) return false;
    if constexpr (Lds::kUber) return eval_uber(s, L, lo, hi, alt);
    else return eval_hc(s, L, lo, hi, alt);
}

template <class Lds>
__device__ inline bool eval_canon(St& s, Lds& L, uint32_t lo, uint32_t hi)
{
    canon(lo, hi);
    return evaluate(s, L, lo, hi, 0);
}

__device__ inline int clampi(int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

// normalised (and perceptually weighted) colour i minus the mean, as
// compute_vectors stores it (crn_dxt1.cpp:155-189)
template <class Lds>
__device__ inline void norm_of(const St& s, const Lds& L, uint32_t i, float* n, float* nw)
{
    const uint32_t c = L.col[i];
    for (int k = 0; k < 3; k++) {
        const float v = (float)((c >> (8 * k)) & 255u) * 1.0f / 255.0f;
        n[k] = v - s.mean[k];
        nw[k] = s.pw[k] * v - s.meanw[k];
    }
}

template <class Lds>
__device__ void compute_vectors(St& s, const Lds& L)
{
    for (int k = 0; k < 3; k++) s.mean[k] = s.meanw[k] = 0.0f;
    for (uint32_t i = 0; i < s.nu; i++) {
        const uint32_t c = L.col[i];
        const float w = (float)L.w[i];
        for (int k = 0; k < 3; k++) {
            const float v = (float)((c >> (8 * k)) & 255u) * 1.0f / 255.0f;
            s.mean[k] += v * w;
            s.meanw[k] += (s.pw[k] * v) * w;
        }
    }
    if (s.total_w) {
        const float inv = 1.0f / (float)s.total_w;
        for (int k = 0; k < 3; k++) {
            s.mean[k] *= inv;
            s.meanw[k] *= inv;
        }
    }
}

// compute_pca (crn_dxt1.cpp:192-256) over the weighted vectors
template <class Lds>
__device__ void compute_pca(St& s, const Lds& L, float* axis)
{
    double cov[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t i = 0; i < s.nu; i++) {
        float n[3], nw[3];
        norm_of(s, L, i, n, nw);
        const float r = nw[0], g = nw[1], b = nw[2];
        if (L.w[i] > 1) {
            const double w = L.w[i];
            cov[0] += r * r * w; cov[1] += r * g * w; cov[2] += r * b * w;
            cov[3] += g * g * w; cov[4] += g * b * w; cov[5] += b * b * w;
        } else {
            cov[0] += r * r; cov[1] += r * g; cov[2] += r * b;
            cov[3] += g * g; cov[4] += g * b; cov[5] += b * b;
        }
    }
    double vr = .9f, vg = 1.0f, vb = .7f;
    for (int it = 0; it < 8; it++) {
        double r = vr * cov[0] + vg * cov[1] + vb * cov[2];
        double g = vr * cov[1] + vg * cov[3] + vb * cov[4];
        double b = vr * cov[2] + vg * cov[4] + vb * cov[5];
        double m = fabs(r);
        if (fabs(g) > m) m = fabs(g);
        if (fabs(b) > m) m = fabs(b);
        if (m > 1e-10) {
            m = 1.0f / m;
            r *= m;
            g *= m;
            b *= m;
        }
        const double d = (vr - r) * (vr - r) + (vg - g) * (vg - g) + (vb - b) * (vb - b);
        vr = r;
        vg = g;
        vb = b;
        if (it > 2 && d < 1e-8) break;
    }
    double len = vr * vr + vg * vg + vb * vb;
    if (len < 1e-10) {
        axis[0] = .2837149f;
        axis[1] = 0.9540631f;
        axis[2] = 0.096277453f;
        return;
    }
    len = 1.0f / sqrt(len);
    axis[0] = (float)(vr * len);
    axis[1] = (float)(vg * len);
    axis[2] = (float)(vb * len);
}

// ray-box intersection (crn_intersect.h:44-120) for the unit cube
__device__ bool ray_unit_box(float* coord, const float* org, const float* dir)
{
    bool inside = true;
    int quad[3];
    float cand[3], maxt[3];
    for (int i = 0; i < 3; i++) {
        if (org[i] < 0.0f) { quad[i] = 1; cand[i] = 0.0f; inside = false; }
        else if (org[i] > 1.0f) { quad[i] = 0; cand[i] = 1.0f; inside = false; }
        else quad[i] = 2;
    }
    if (inside) return false;
    for (int i = 0; i < 3; i++) maxt[i] = (quad[i] != 2 && dir[i] != 0.0f) ? (cand[i] - org[i]) / dir[i] : -1.0f;
    int wp = 0;
    for (int i = 1; i < 3; i++)
        if (maxt[wp] < maxt[i]) wp = i;
    if (maxt[wp] < 0.0f) return false;
    float c[3];
    for (int i = 0; i < 3; i++) {
        if (i != wp) {
            c[i] = org[i] + maxt[wp] * dir[i];
            if (c[i] < 0.0f || c[i] > 1.0f) return false;
        } else {
            c[i] = cand[i];
        }
    }
    for (int i = 0; i < 3; i++) coord[i] = c[i];
    return true;
}

// refine_solution (crn_dxt1.cpp:525-698), levels 0 and 1.  The sums are of
// integers times integer weights (exact in double), so a lane-split sum is
// the reference's value.
template <class Lds>
__device__ bool refine(St& s, Lds& L, int level)
{
    compute_selectors(s, L);
    uint64_t A0 = 0, A1 = 0, A2 = 0, T1[3] = {0, 0, 0}, T2[3] = {0, 0, 0};
    for (uint32_t i = lane_id(); i < s.nu; i += 64) {
        const uint32_t c = L.col[i];
        const uint64_t w = L.w[i];
        const int step = L.sel[i] ^ 1;
        const uint32_t w1 = step == 0 ? 3 : step == 1 ? 0 : step == 2 ? 2 : 1;
        const uint32_t q0 = step >= 2 ? 2 : 0;
        const uint32_t q1 = step == 0 ? 0 : step == 1 ? 9 : step == 2 ? 1 : 4;
        const uint32_t q2 = step == 0 ? 9 : step == 1 ? 0 : step == 2 ? 4 : 1;
        A0 += q0 * w;
        A1 += q1 * w;
        A2 += q2 * w;
        for (int k = 0; k < 3; k++) {
            const uint64_t v = ((c >> (8 * k)) & 255u) * w;
            T1[k] += w1 * v;
            T2[k] += v;
        }
    }
    A0 = wave_sum(A0);
    A1 = wave_sum(A1);
    A2 = wave_sum(A2);
    double t1[3], t2[3];
    for (int k = 0; k < 3; k++) {
        t1[k] = (double)wave_sum(T1[k]);
        t2[k] = 3 * (double)wave_sum(T2[k]) - t1[k];
    }
    const double xx = (double)A2, yy = (double)A1, xy = (double)A0;
    const double t = xx * yy - xy * xy;
    if (!yy || !xx || fabs(t) < (double).0000125f) return false;
    const double frb = (double)(3.0f * 31.0f / 255.0f) / t;
    const double fg = frb * (double)(63.0f / 31.0f);
    int e0[3], e1[3];
    e0[0] = clampi((int)((t1[0] * yy - t2[0] * xy) * frb + (double)0.5f), 31);
    e0[1] = clampi((int)((t1[1] * yy - t2[1] * xy) * fg + (double)0.5f), 63);
    e0[2] = clampi((int)((t1[2] * yy - t2[2] * xy) * frb + (double)0.5f), 31);
    e1[0] = clampi((int)((t2[0] * xx - t1[0] * xy) * frb + (double)0.5f), 31);
    e1[1] = clampi((int)((t2[1] * xx - t1[1] * xy) * fg + (double)0.5f), 63);
    e1[2] = clampi((int)((t2[2] * xx - t1[2] * xy) * frb + (double)0.5f), 31);
    bool improved = false;
    if (level == 0) {
        improved |= eval_canon(s, L, (uint32_t)((e1[0] << 11) | (e1[1] << 5) | e1[2]),
                               (uint32_t)((e0[0] << 11) | (e0[1] << 5) | e0[2]));
    } else {
        for (int i = 0; i < 2; i++)
            for (int rr = -1; rr <= 1; rr++)
                for (int gr = -1; gr <= 1; gr++)
                    for (int br = -1; br <= 1; br++) {
                        int c0[3] = {e0[0], e0[1], e0[2]}, c1[3] = {e1[0], e1[1], e1[2]};
                        int* c = i ? c1 : c0;
                        c[0] = clampi(c[0] + rr, 31);
                        c[1] = clampi(c[1] + gr, 63);
                        c[2] = clampi(c[2] + br, 31);
                        improved |= eval_canon(s, L, pack565(c0[0], c0[1], c0[2], 0), pack565(c1[0], c1[1], c1[2], 0));
                    }
    }
    return improved;
}

// fast_random (crn_rand.cpp:310-405)
struct FastRandom {
    uint32_t jsr = 0xABCD917Au, jcong = 0x17F3DEADu;
    __device__ uint32_t u32()
    {
        jsr ^= jsr << 17;
        jsr ^= jsr >> 13;
        jsr ^= jsr << 5;
        jcong = 69069u * jcong + 1234567u;
        return jsr ^ jcong;
    }
    __device__ float frand(float l, float h)
    {
        if (l >= h) return l;
        const float v = (float)((double)l + (double)(h - l) * ((double)u32() * (1.0 / 4294967296.0)));
        return v < l ? l : (v > h ? h : v);
    }
};

// try_median4 (crn_dxt1.cpp:1181-1308): 4-means over the colours (sequential
// float sums), then the six pairs of means and refine_solution.
template <class Lds>
__device__ bool try_median4(St& s, Lds& L, const float* lo, const float* hi)
{
    float means[4][3];
    if (s.nu <= 4) {
        for (int i = 0; i < 4; i++) {
            float n[3], nw[3];
            norm_of(s, L, (uint32_t)i < s.nu - 1 ? (uint32_t)i : s.nu - 1, n, nw);
            for (int k = 0; k < 3; k++) means[i][k] = n[k];
        }
    } else {
        for (int k = 0; k < 3; k++) {
            means[0][k] = lo[k] - s.mean[k];
            means[3][k] = hi[k] - s.mean[k];
        }
        for (int k = 0; k < 3; k++) {
            means[1][k] = means[0][k] + (means[3][k] - means[0][k]) * (1.0f / 3.0f);
            means[2][k] = means[0][k] + (means[3][k] - means[0][k]) * (2.0f / 3.0f);
        }
        FastRandom rm;
        uint32_t rover = 0;
        float prev = 1.0e+37f;
        for (uint32_t it = 0; it < 8; it++) {
            float nm[4][3] = {}, nwt[4] = {0, 0, 0, 0};
            float total = 0;
            for (uint32_t i = 0; i < s.nu; i++) {
                float v[3], vw[3];
                norm_of(s, L, i, v, vw);
                float bd = 0;
                for (int k = 0; k < 3; k++) {
                    const float d = means[0][k] - v[k];
                    bd += d * d;
                }
                int bi = 0;
                for (int jj = 1; jj < 4; jj++) {
                    float dd = 0;
                    for (int k = 0; k < 3; k++) {
                        const float d = means[jj][k] - v[k];
                        dd += d * d;
                    }
                    if (dd < bd) {
                        bd = dd;
                        bi = jj;
                    }
                }
                total += bd;
                const float w = (float)L.w[i];
                for (int k = 0; k < 3; k++) nm[bi][k] += v[k] * w;
                nwt[bi] += w;
            }
            uint32_t hidx = 0;
            float hw = 0;
            bool empty = false;
            for (int jj = 0; jj < 4; jj++) {
                if (nwt[jj] > 0.0f) {
                    for (int k = 0; k < 3; k++) means[jj][k] = nm[jj][k] / nwt[jj];
                    if (nwt[jj] > hw) {
                        hw = nwt[jj];
                        hidx = (uint32_t)jj;
                    }
                } else {
                    empty = true;
                }
            }
            if (!empty) {
                if (fabsf(total - prev) < .00001f) break;
                prev = total;
            } else {
                prev = 1.0e+37f;
            }
            if (empty && it != 7) {
                const uint32_t ri = (hidx + rover) & 3u;
                rover++;
                for (int jj = 0; jj < 4; jj++)
                    if (nwt[jj] == 0.0f)
                        for (int k = 0; k < 3; k++) means[jj][k] = means[ri][k] + rm.frand(-.00196f, .00196f);
            }
        }
    }
    bool improved = false;
    for (int i = 0; i < 3; i++)
        for (int jj = i + 1; jj < 4; jj++) {
            float a[3], b[3];
            for (int k = 0; k < 3; k++) {
                a[k] = means[i][k] + s.mean[k];
                b[k] = means[jj][k] + s.mean[k];
            }
            uint32_t lo2 = pack565((int)floorf(.5f + a[0] * 31.0f), (int)floorf(.5f + a[1] * 63.0f),
                                   (int)floorf(.5f + a[2] * 31.0f), 0);
            uint32_t hi2 = pack565((int)floorf(.5f + b[0] * 31.0f), (int)floorf(.5f + b[1] * 63.0f),
                                   (int)floorf(.5f + b[2] * 31.0f), 0);
            improved |= eval_canon(s, L, lo2, hi2);
        }
    improved |= refine(s, L, 1);
    return improved;
}

// are_selectors_all_equal (crn_dxt1.h:300-316) on the best solution's
// selector vector as last computed: the hc evaluators leave it untouched, so
// it may belong to an earlier solution, and it is empty before the first
// compute_selectors.
template <class Lds>
__device__ bool sel_all_equal(const St& s, const Lds& L)
{
    if (!s.sel_nonempty) return false;
    uint32_t diff = 0;
    for (uint32_t i = lane_id(); i < s.nu; i += 64) diff |= L.sel[i] != L.sel[0];
    return wave_sum(diff) == 0;
}

// try_average_block_as_solid (crn_dxt1.cpp:93-153), uber branch
template <class Lds>
__device__ bool try_average_solid(St& s, Lds& L)
{
    uint64_t tr = 0, tg = 0, tb = 0, tw = 0;
    for (uint32_t i = lane_id(); i < s.nu; i += 64) {
        const uint32_t c = L.col[i];
        const uint64_t w = L.w[i];
        tw += w;
        tr += (c & 255u) * w;
        tg += ((c >> 8) & 255u) * w;
        tb += ((c >> 16) & 255u) * w;
    }
    tw = wave_sum(tw);
    tr = wave_sum(tr);
    tg = wave_sum(tg);
    tb = wave_sum(tb);
    const uint64_t half = tw >> 1;
    const uint32_t ar = (uint32_t)((tr + half) / tw), ag = (uint32_t)((tg + half) / tw), ab = (uint32_t)((tb + half) / tw);
    bool improved = evaluate(s, L, (g_c_om5[ar][0] << 11) | (g_c_om6[ag][0] << 5) | g_c_om5[ab][0],
                             (g_c_om5[ar][1] << 11) | (g_c_om6[ag][1] << 5) | g_c_om5[ab][1], 0);
    if (Lds::kUber && s.best.err)
        improved |= evaluate(s, L, (g_c_om5_3[ar][0] << 11) | (g_c_om6_3[ag][0] << 5) | g_c_om5_3[ab][0],
                             (g_c_om5_3[ar][1] << 11) | (g_c_om6_3[ag][1] << 5) | g_c_om5_3[ab][1], 0);
    for (uint32_t i = 0; i < s.nu; i++) {
        const uint32_t c = L.col[i];
        const uint32_t r = c & 255u, g = (c >> 8) & 255u, b = (c >> 16) & 255u;
        if (r == ar && g == ag && b == ab) continue;
        improved |= evaluate(s, L, (g_c_om5[r][0] << 11) | (g_c_om6[g][0] << 5) | g_c_om5[b][0],
                             (g_c_om5[r][1] << 11) | (g_c_om6[g][1] << 5) | g_c_om5[b][1], 0);
        if (Lds::kUber && s.best.err)
            improved |= evaluate(s, L, (g_c_om5_3[r][0] << 11) | (g_c_om6_3[g][0] << 5) | g_c_om5_3[b][0],
                                 (g_c_om5_3[r][1] << 11) | (g_c_om6_3[g][1] << 5) | g_c_om5_3[b][1], 0);
    }
    return improved;
}

// compute_endpoint_component_errors (crn_dxt1.cpp:369-418)
// Per-channel error of one selector class as a function of the class's value
// p: W p^2 - WP2 p + WPP (the reference's compute_endpoint_component_errors
// tables, crn_dxt1.cpp:369-418, evaluated where they are read).
struct QErr {
    uint64_t W[4], WP2[4], WPP[4];
    __device__ inline uint64_t at(int sl, uint64_t p) const { return W[sl] * p * p - WP2[sl] * p + WPP[sl]; }
};

__device__ inline uint32_t expand_comp(int comp, uint32_t c)
{
    return comp == 1 ? (((c << 2) | (c >> 4)) & 0xFF) : (((c << 3) | (c >> 2)) & 0xFF);
}

template <class Lds>
__device__ void comp_errors(St& s, Lds& L, int comp, QErr& q, uint64_t* brem)
{
    uint64_t W[4] = {0, 0, 0, 0}, WP2[4] = {0, 0, 0, 0}, WPP[4] = {0, 0, 0, 0};
    for (uint32_t i = lane_id(); i < s.nu; i += 64) {
        const uint64_t p = (L.col[i] >> (8 * comp)) & 255u;
        const uint64_t w = L.w[i];
        const uint32_t sl = L.sel[i];
        for (uint32_t k = 0; k < 4; k++)
            if (k == sl) {
                W[k] += w;
                WP2[k] += w * p * 2;
                WPP[k] += w * p * p;
            }
    }
    for (int k = 0; k < 4; k++) {
        q.W[k] = wave_sum(W[k]);
        q.WP2[k] = wave_sum(WP2[k]);
        q.WPP[k] = wave_sum(WPP[k]);
    }
    const uint32_t lane = lane_id();
    const uint32_t lim = comp == 1 ? 64 : 32;
    for (int sl = 0; sl < 4; sl++) {
        const uint32_t n = sl < 2 ? lim : 256;
        uint64_t m = ~0ull;
        for (uint32_t c = lane; c < n; c += 64) m = min(m, q.at(sl, sl < 2 ? expand_comp(comp, c) : c));
        for (int off = 32; off > 0; off >>= 1) {
            const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)m, off);
            const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(m >> 32), off);
            m = min(m, ((uint64_t)hi << 32) | lo);
        }
        brem[sl] = m;
    }
    for (int sl = 3; sl; sl--) brem[sl - 1] += brem[sl];
}

// optimize_endpoint_comps (crn_dxt1.cpp:420-486)
template <class Lds>
__device__ void optimize_comps(St& s, Lds& L)
{
    compute_selectors(s, L);
    if (s.best.alpha || !s.best.err) return;
    int slo[3], shi[3];
    unpack565(s.best.lo, 1, slo[0], slo[1], slo[2]);
    unpack565(s.best.hi, 1, shi[0], shi[1], shi[2]);
    uint64_t brem[4];
    QErr q;
    for (int comp = 0; comp < 3; comp++) {
        uint32_t p0 = (uint32_t)slo[comp], p1 = (uint32_t)shi[comp];
        int low[3], high[3];
        unpack565(s.best.lo, 0, low[0], low[1], low[2]);
        unpack565(s.best.hi, 0, high[0], high[1], high[2]);
        comp_errors(s, L, comp, q, brem);
        uint64_t best = q.at(0, p0) + q.at(1, p1) + q.at(2, (p0 * 2 + p1) / 3) + q.at(3, (p0 + p1 * 2) / 3);
        if (brem[0] >= best) continue;
        const uint32_t lim = comp == 1 ? 64 : 32;
        for (uint32_t c0 = 0; c0 < lim; c0++) {
            p0 = expand_comp(comp, c0);
            uint64_t e0 = q.at(0, p0);
            if (e0 + brem[1] >= best) continue;
            low[comp] = (int)c0;
            const uint32_t plow = pack565(low[0], low[1], low[2], 0);
            for (uint32_t c1 = 0; c1 < lim; c1++) {
                p1 = expand_comp(comp, c1);
                uint64_t e = e0 + q.at(1, p1);
                if (e + brem[2] >= best) continue;
                e += q.at(2, (p0 * 2 + p1) / 3);
                if (e + brem[3] >= best) continue;
                e += q.at(3, (p0 + p1 * 2) / 3);
                if (e >= best) continue;
                high[comp] = (int)c1;
                if (!evaluate(s, L, plow, pack565(high[0], high[1], high[2], 0), 0)) continue;
                if (!s.best.err) return;
                compute_selectors(s, L);
                comp_errors(s, L, comp, q, brem);
                best = q.at(0, p0) + q.at(1, p1) + q.at(2, (p0 * 2 + p1) / 3) + q.at(3, (p0 + p1 * 2) / 3);
                e0 = q.at(0, p0);
                if (e0 + brem[1] >= best) break;
            }
        }
    }
}

__device__ inline void lerp_color(int ar, int ag, int ab, int br, int bg, int bb, float f, int rounding, int& r,
                                  int& g, int& b)
{
    const float rr = rounding ? 1.0f : 0.0f;
    r = clampi((int)(rr + ((float)ar + ((float)br - (float)ar) * f)), 255);
    g = clampi((int)(rr + ((float)ag + ((float)bg - (float)ag) * f)), 255);
    b = clampi((int)(rr + ((float)ab + ((float)bb - (float)ab) * f)), 255);
}

// try_combinatorial_encoding (crn_dxt1.cpp:1886-1997): 2 to 4 colours only
template <class Lds>
__device__ void try_combinatorial(St& s, Lds& L)
{
    if (s.nu < 2 || s.nu > 4) return;
    int tr[52], tg[52], tb[52];
    uint32_t n = 0;
    for (uint32_t i = 0; i < s.nu; i++) {
        const uint32_t c = L.col[i];
        tr[n] = (int)(c & 255u);
        tg[n] = (int)((c >> 8) & 255u);
        tb[n] = (int)((c >> 16) & 255u);
        n++;
    }
    if (s.nu == 2) {
        const float fs[10] = {2.0f, 3.0f, .5f, 1.5f, -1.0f, 2.0f, -.5f, .5f, -2.0f, -1.0f};
        for (int k = 0; k < 2; k++)
            for (int q = 0; q < 2; q++) {
                const int r = q ^ 1;
                for (int f = 0; f < 10; f++) {
                    lerp_color(tr[q], tg[q], tb[q], tr[r], tg[r], tb[r], fs[f], k, tr[n], tg[n], tb[n]);
                    n++;
                }
            }
    } else if (s.nu == 3) {
        const float fs[4] = {1.5f, 2.0f / 3.0f, 1.0f / 3.0f, -.5f};
        for (int i = 0; i <= 2; i++)
            for (int jj = 0; jj <= 2; jj++) {
                if (i == jj) continue;
                for (int f = 0; f < 4; f++) {
                    lerp_color(tr[i], tg[i], tb[i], tr[jj], tg[jj], tb[jj], fs[f], 1, tr[n], tg[n], tb[n]);
                    n++;
                }
            }
    }
    uint32_t packed[64];
    uint32_t np = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t p = pack565(tr[i], tg[i], tb[i], 1);
        bool dup = false;
        for (uint32_t k = 0; k < np; k++)
            if (packed[k] == p) dup = true;
        if (!dup) packed[np++] = p;
    }
    for (uint32_t i = 0; s.best.err && i + 1 < np; i++)
        for (uint32_t jj = i + 1; s.best.err && jj < np; jj++) evaluate(s, L, packed[i], packed[jj], 0);
    const uint64_t err = s.best.err;
    if (err) s.best.err = 1;
    for (uint32_t i = 0; s.best.err && i + 1 < np; i++)
        for (uint32_t jj = i + 1; s.best.err && jj < np; jj++) evaluate(s, L, packed[i], packed[jj], 1);
    if (s.best.err) s.best.err = err;
}

__device__ uint32_t sweep(uint32_t range, const float* base, const float (*axis2)[3], uint32_t* out)
{
    int prev[2] = {-1, -1};
    uint32_t n = 0;
    const float init[3] = {base[0] + .5f, base[1] + .5f, base[2] + .5f};
    for (uint32_t i = 0; i < range; i++) {
        const int ls = i ? 0 : 1;
        const int x = c_uber_probe[i];
        for (int sg = ls; sg < 2; sg++) {
            const float pr = init[0] + axis2[sg][0] * (float)x;
            const float pg = init[1] + axis2[sg][1] * (float)x;
            const float pb = init[2] + axis2[sg][2] * (float)x;
            const int r = clampi((int)floorf(pr), 31), g = clampi((int)floorf(pg), 63), b = clampi((int)floorf(pb), 31);
            const int packed = b | (g << 5) | (r << 11);
            if (packed != prev[sg]) {
                out[n++] = (uint32_t)packed;
                prev[sg] = packed;
            }
        }
    }
    return n;
}

template <class Lds>
__device__ void lattice(St& s, Lds& L, int which)
{
    int lr, lg, lb;
    unpack565(which ? s.best.hi : s.best.lo, 0, lr, lg, lb);
    for (int i = 0; i < 26; i++) {
        const int r = lr + c_adj[i][0];
        if (r < 0 || r > 31) continue;
        const int g = lg + c_adj[i][1];
        if (g < 0 || g > 63) continue;
        const int b = lb + c_adj[i][2];
        if (b < 0 || b > 31) continue;
        const uint32_t c = pack565(r, g, b, 0);
        if (which) eval_canon(s, L, s.best.lo, c);
        else eval_canon(s, L, c, s.best.hi);
    }
    unpack565(which ? s.best.hi : s.best.lo, 0, lr, lg, lb);
    for (int a = 0; a < 3; a++) {
        const int limit = a == 1 ? 63 : 31;
        for (int sg = -2; sg <= 2; sg += 4) {
            int comp[3] = {lr, lg, lb};
            const int q = comp[a] + sg;
            if (q < 0 || q > limit) continue;
            comp[a] = q;
            const uint32_t c = pack565(comp[0], comp[1], comp[2], 0);
            if (which) eval_canon(s, L, s.best.lo, c);
            else eval_canon(s, L, c, s.best.hi);
        }
    }
}

// optimize_endpoints (crn_dxt1.cpp:703-1067), uber quality
template <class Lds>
__device__ void optimize_endpoints(St& s, Lds& L, float* lo, float* hi)
{
    const float olo[3] = {lo[0], lo[1], lo[2]}, ohi[3] = {hi[0], hi[1], hi[2]};
    const uint32_t range = 15, passes = 4;
    const float dist = .015625f;
    try_median4(s, L, olo, ohi);
    float ax[2][3];
    for (int k = 0; k < 3; k++) ax[1][k] = s.axis[k] * dist;
    ax[1][0] *= 31.0f;
    ax[1][1] *= 63.0f;
    ax[1][2] *= 31.0f;
    for (int k = 0; k < 3; k++) ax[0][k] = -ax[1][k];
    const float lim[3] = {31.0f, 63.0f, 31.0f};
    for (int k = 0; k < 3; k++) {
        const float a = lo[k] * lim[k];
        lo[k] = a < 0.0f ? 0.0f : (a > lim[k] ? lim[k] : a);
        const float b = hi[k] * lim[k];
        hi[k] = b < 0.0f ? 0.0f : (b > lim[k] ? lim[k] : b);
    }
    uint32_t plow[31], phigh[31];
    for (uint32_t pass = 0; pass < passes; pass++) {
        if (pass) {
            int r, g, b;
            unpack565(s.best.lo, 0, r, g, b);
            lo[0] = (float)r; lo[1] = (float)g; lo[2] = (float)b;
            unpack565(s.best.hi, 0, r, g, b);
            hi[0] = (float)r; hi[1] = (float)g; hi[2] = (float)b;
        }
        const uint64_t prev = s.best.err;
        if (!prev) break;
        const uint32_t nl = sweep(range, lo, ax, plow);
        const uint32_t nh = sweep(range, hi, ax, phigh);
        for (uint32_t i = 0; i < nl; i++)
            for (uint32_t jj = 0; jj < nh; jj++) eval_canon(s, L, plow[i], phigh[jj]);
        lattice(s, L, 0);
        lattice(s, L, 1);
        if (!s.best.err || (pass && s.best.err == prev)) break;
        refine(s, L, 1);
    }
    if (s.best.err && !s.have_alpha) {
        bool solid = false;
        // the uber evaluator keeps the best solution's selectors current
        if (Lds::kUber && s.sel_nonempty) compute_selectors(s, L);
        if (sel_all_equal(s, L)) solid = try_average_solid(s, L);
        if (!solid) optimize_comps(s, L);
    }
    if (s.best.err) try_combinatorial(s, L);
}

// handle_multicolor_block (crn_dxt1.cpp:1069-1178)
template <class Lds>
__device__ void handle_multicolor(St& s, Lds& L)
{
    uint32_t npass = 1;
    s.pw[0] = s.pw[1] = s.pw[2] = 1.0f;
    if (s.perceptual) {
        float red = 0, blue = 0, al = 0;
        for (uint32_t i = 0; i < s.nu; i++) {
            const uint32_t c = L.col[i];
            const int r = (int)(c & 255u), g = (int)((c >> 8) & 255u), b = (int)((c >> 16) & 255u);
            const int l = (r + g + b + 1) / 3;
            const float lf = (float)l;
            const float scale = (float)L.w[i] / (lf > 1.0f ? lf : 1.0f);
            red += scale * r;
            blue += scale * b;
            al += l;
        }
        red /= s.total_w;
        blue /= s.total_w;
        al /= s.total_w;
        al = al * 16.0f / 255.0f;
        al = 1.0f < al ? 1.0f : al;
        const float m = red > blue ? red : blue;
        float sat = m * 1.0f / 3.0f;
        sat = sat < 0.0f ? 0.0f : (sat > 1.0f ? 1.0f : sat);
        const float p = al * powf(sat, 2.75f);
        if (p >= 1.0f) {
            npass = 1;
        } else {
            npass = 2;
            const float lum[3] = {.212f, .72f, .072f};
            for (int k = 0; k < 3; k++) s.pw[k] = lum[k] + (s.pw[k] - lum[k]) * p;
        }
    }
    for (uint32_t pass = 0; pass < npass; pass++) {
        compute_vectors(s, L);
        compute_pca(s, L, s.axis);
        s.axis[0] /= s.pw[0];
        s.axis[1] /= s.pw[1];
        s.axis[2] /= s.pw[2];
        double n = s.axis[0] * s.axis[0];
        n += s.axis[1] * s.axis[1];
        n += s.axis[2] * s.axis[2];
        if (n != 0) {
            const float k = (float)(1.0f / sqrt(n));
            for (int i = 0; i < 3; i++) s.axis[i] *= k;
        }
        if (npass > 1) {
            if (fabsf(s.axis[0]) >= .795f) {
                s.pw[0] = .424f; s.pw[1] = .6f; s.pw[2] = .072f;
            } else if (fabsf(s.axis[2]) >= .795f) {
                s.pw[0] = .212f; s.pw[1] = .6f; s.pw[2] = .212f;
            } else {
                break;
            }
        }
    }
    float l = 1e+9f, h = -1e+9f;
    for (uint32_t i = 0; i < s.nu; i++) {
        float n[3], nw[3];
        norm_of(s, L, i, n, nw);
        float d = n[0] * s.axis[0];
        d += n[1] * s.axis[1];
        d += n[2] * s.axis[2];
        l = d < l ? d : l;
        h = h < d ? d : h;
    }
    float lo[3], hi[3];
    for (int k = 0; k < 3; k++) {
        lo[k] = s.mean[k] + l * s.axis[k];
        hi[k] = s.mean[k] + h * s.axis[k];
    }
    bool in = true;
    for (int k = 0; k < 3; k++)
        if (lo[k] < 0.0f || lo[k] > 1.0f) in = false;
    if (!in) {
        float c[3];
        if (ray_unit_box(c, lo, s.axis)) { lo[0] = c[0]; lo[1] = c[1]; lo[2] = c[2]; }
    }
    in = true;
    for (int k = 0; k < 3; k++)
        if (hi[k] < 0.0f || hi[k] > 1.0f) in = false;
    if (!in) {
        float c[3];
        const float nd[3] = {-s.axis[0], -s.axis[1], -s.axis[2]};
        if (ray_unit_box(c, hi, nd)) { hi[0] = c[0]; hi[1] = c[1]; hi[2] = c[2]; }
    }
    optimize_endpoints(s, L, lo, hi);
}

struct ClusterOut {
    uint64_t error;
    uint32_t lo_hi;    // results::m_low_color | m_high_color << 16
    uint32_t nunique;
    uint32_t flags;    // reordered | alternate_rounding << 2 | enforce << 3 | enforced << 4 | too_many_colours << 8
    uint32_t pad;
};
constexpr uint32_t kNotEncoded = 1u << 8;

// Open-addressing colour hash over L.kCapH slots (a power of two in LDS, any
// size in the global store): first slot and linear probe.
template <class Lds>
__device__ inline uint32_t hash_slot(const Lds& L, uint32_t key)
{
    return hash32(key) % L.kCapH;
}

template <class Lds>
__device__ inline uint32_t hash_next(const Lds& L, uint32_t h)
{
    return h + 1 == L.kCapH ? 0 : h + 1;
}

template <class Lds>
__device__ inline int hash_find(const Lds& L, uint32_t key)
{
    uint32_t h = hash_slot(L, key);
    while (L.hkey[h]) {
        if (L.hkey[h] == key) return (int)L.hval[h];
        h = hash_next(L, h);
    }
    return -1;
}

// Unique-colour index of texel i (key = its colour | 0xFF000000).
template <uint32_t MaxU>
__device__ inline uint32_t unique_index(const LdsT<MaxU>& L, uint32_t, uint32_t key)
{
    return (uint32_t)hash_find(L, key);
}

template <uint32_t MaxU>
__device__ inline uint32_t unique_index(const LdsS<MaxU>& L, uint32_t i, uint32_t)
{
    return L.tmap[i];
}

// Unique colours of a cluster of n <= MaxU texels in first-appearance order
// with their texel counts, and every texel's unique index, without a hash:
// bitonic sort of (colour, texel index) pairs, the first texel of each run
// of equal colours marked in a bitmap, a colour's rank = the number of marked
// texels before its first one.  One wave; the block is the wave.
template <uint32_t MaxU>
__device__ uint32_t unique_sorted(LdsS<MaxU>& L, const uint32_t* __restrict__ px, uint32_t p0, uint32_t n,
                              uint32_t thresh, uint32_t& opaque)
{
    uint32_t* key = reinterpret_cast<uint32_t*>(L.sel);
    uint16_t* idx = reinterpret_cast<uint16_t*>(key + MaxU);
    const uint32_t lane = lane_id();
    uint32_t P = 64;
    while (P < n) P <<= 1;   // n <= MaxU, a power of two
    uint32_t nop = 0;
    for (uint32_t i = lane; i < P; i += 64) {
        // transparent texels (DXT1A) sort after every colour, before the padding
        const uint32_t p = i < n ? px[p0 + i] : 0u;
        const bool tr = i < n && (p >> 24) < thresh;
        key[i] = i < n ? (tr ? 0xFFFFFFFEu : (p & 0x00FFFFFFu)) : 0xFFFFFFFFu;
        idx[i] = (uint16_t)i;
        nop += (i < n && !tr) ? 1u : 0u;
    }
    for (int off = 32; off; off >>= 1) nop += (uint32_t)__shfl_xor((int)nop, off);
    opaque = nop;
    const uint32_t nsrc = n;
    n = nop;
    for (uint32_t i = lane; i < MaxU / 32; i += 64) L.fbits[i] = 0;
    __syncthreads();
    for (uint32_t k = 2; k <= P; k <<= 1)
        for (uint32_t j = k >> 1; j > 0; j >>= 1) {
            for (uint32_t t = lane; t < P / 2; t += 64) {
                const uint32_t a = 2 * j * (t / j) + (t % j), b = a + j;
                const uint32_t ka = key[a], kb = key[b];
                const uint32_t ia = idx[a], ib = idx[b];
                const bool gt = ka > kb || (ka == kb && ia > ib);
                if (gt == ((a & k) == 0)) {
                    key[a] = kb;
                    key[b] = ka;
                    idx[a] = (uint16_t)ib;
                    idx[b] = (uint16_t)ia;
                }
            }
            __syncthreads();
        }
    // runs of equal colours (the padding sorts past the n real texels)
    for (uint32_t j = lane; j < n; j += 64)
        if (j == 0 || key[j] != key[j - 1]) atomicOr(&L.fbits[idx[j] >> 5], 1u << (idx[j] & 31u));
    __syncthreads();
    const uint32_t nw = (nsrc + 31) / 32;   // first-texel bits of all texels: at most 64 words, one per lane
    const uint32_t cnt = lane < nw ? (uint32_t)__popc(L.fbits[lane]) : 0u;
    uint32_t incl = cnt;
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t v = (uint32_t)__shfl_up((int)incl, off);
        if (lane >= off) incl += v;
    }
    if (lane < nw) L.fbase[lane] = incl - cnt;
    const uint32_t nu = (uint32_t)__shfl((int)incl, 63);
    __syncthreads();
    for (uint32_t j = lane; j < n; j += 64) {
        if (j != 0 && key[j] == key[j - 1]) continue;
        const uint32_t f = idx[j];   // the colour's first texel
        const uint32_t r = L.fbase[f >> 5] + (uint32_t)__popc(L.fbits[f >> 5] & ((1u << (f & 31u)) - 1u));
        uint32_t e = j;
        for (; e < n && key[e] == key[j]; e++) L.tmap[idx[e]] = (uint16_t)r;
        L.col[r] = key[j];
        L.w[r] = (uint16_t)(e - j);
    }
    __syncthreads();
    return nu;
}

// The colour set of a cluster with more distinct colours than the LDS store
// holds: colours, weights, selectors and the hash in global memory (a scratch
// region the cluster owns, sized by its texel count), the per-channel tables
// in LDS.  Members mirror LdsT's so every search step reads either store.
struct GStore {
    uint32_t kCapU, kCapH;
    uint32_t* col;
    uint32_t* w;
    uint8_t* sel;
    uint32_t* hkey;
    uint32_t* hval;
    uint32_t* hcnt;
    uint32_t* mark;
    unsigned long long (*pl)[64][3];
    uint64_t *rlo, *rhi, *glo, *ghi, *blo, *bhi;
};

__device__ inline uint32_t unique_index(const GStore& L, uint32_t, uint32_t key)
{
    return (uint32_t)hash_find(L, key);
}


struct GTables {
    unsigned long long pl[3][64][3];
    uint64_t rlo[32], rhi[32], glo[64], ghi[64], blo[32], bhi[32];
};

// Clusters by texel count into the four LDS sizes (any order within a class:
// each cluster's result depends only on its own texels).
__global__ void __launch_bounds__(256) dxt1_cluster_classes_kernel(const uint32_t* __restrict__ ofs,
                                                                  uint32_t nclusters, uint32_t* __restrict__ count,
                                                                  uint32_t* __restrict__ list)
{
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nclusters) return;
    const uint32_t n = ofs[c + 1] - ofs[c];
    const uint32_t k = n <= 512 ? 0 : n <= 1024 ? 1 : n <= 2048 ? 2 : 3;
    list[(size_t)k * nclusters + atomicAdd(&count[k], 1u)] = c;
}

// Everything after the unique colours: error planes (crn_dxt1.cpp:2134-2203),
// the 1-colour shortcut or handle_multicolor_block, selectors and
// return_solution, for either colour store.
template <class Lds>
__device__ void cluster_search(St& s, Lds& L, const uint32_t* __restrict__ px, uint32_t p0, uint32_t n,
                               int flags, uint32_t cl, ClusterOut* __restrict__ out,
                               uint8_t* __restrict__ selectors)
{
    const uint32_t lane = lane_id();
    ClusterOut o{};
    const uint32_t thresh = s.have_alpha ? ((uint32_t)flags >> 8) & 255u : 0u;
    if (!s.nu) {
        // no opaque texel (compute_internal, crn_dxt1.cpp:2207-2212): a 3-colour
        // block of zero endpoints, every selector transparent (error 0 here)
        if (selectors)
            for (uint32_t i = lane; i < n; i += 64) selectors[p0 + i] = 3;
        if (lane == 0) {
            o.flags = 2;
            out[cl] = o;
        }
        return;
    }
    for (uint32_t i = lane; i < s.nu; i += 64) {
        const uint32_t c = L.col[i];
        const unsigned long long w = L.w[i];
        const uint32_t R = c & 255u, G = (c >> 8) & 255u, B = (c >> 16) & 255u;
        const uint32_t r = (R >> 3) + (((R & 7) > (R >> 5)) ? 1 : 0);
        const uint32_t g = (G >> 2) + (((G & 3) > (G >> 6)) ? 1 : 0);
        const uint32_t b = (B >> 3) + (((B & 7) > (B >> 5)) ? 1 : 0);
        atomicAdd(&L.pl[0][r][0], w); atomicAdd(&L.pl[0][r][1], w * R); atomicAdd(&L.pl[0][r][2], w * R * R);
        atomicAdd(&L.pl[1][g][0], w); atomicAdd(&L.pl[1][g][1], w * G); atomicAdd(&L.pl[1][g][2], w * G * G);
        atomicAdd(&L.pl[2][b][0], w); atomicAdd(&L.pl[2][b][1], w * B); atomicAdd(&L.pl[2][b][2], w * B * B);
    }
    __syncthreads();
    if (lane == 0) {
        const uint64_t sc[3] = {s.perceptual ? 8u : 1u, s.perceptual ? 25u : 1u, 1u};
        for (int ch = 0; ch < 3; ch++) {
            const int lim = ch == 1 ? 64 : 32;
            for (int c = 0; c < lim; c++)
                for (int k = 0; k < 3; k++) {
                    L.pl[ch][c][k] *= sc[ch];
                    if (c) L.pl[ch][c][k] += L.pl[ch][c - 1][k];
                }
        }
        for (uint32_t c = 0; c < 32; c++) {
            const uint64_t C = (uint8_t)((c << 3) | (c >> 2));
            L.rlo[c] = L.pl[0][c][2] + C * C * L.pl[0][c][0] - 2 * C * L.pl[0][c][1];
            L.rhi[c] = L.pl[0][31][2] + C * C * L.pl[0][31][0] - 2 * C * L.pl[0][31][1] - L.rlo[c];
            L.blo[c] = L.pl[2][c][2] + C * C * L.pl[2][c][0] - 2 * C * L.pl[2][c][1];
            L.bhi[c] = L.pl[2][31][2] + C * C * L.pl[2][31][0] - 2 * C * L.pl[2][31][1] - L.blo[c];
        }
        for (uint32_t c = 0; c < 64; c++) {
            const uint64_t C = (uint8_t)((c << 2) | (c >> 4));
            L.glo[c] = L.pl[1][c][2] + C * C * L.pl[1][c][0] - 2 * C * L.pl[1][c][1];
            L.ghi[c] = L.pl[1][63][2] + C * C * L.pl[1][63][0] - 2 * C * L.pl[1][63][1] - L.glo[c];
        }
    }
    __syncthreads();
    s.best.lo = s.best.hi = 0;
    s.best.err = ~0ull;
    s.best.alt = s.best.enforce = s.best.enforced = s.best.alpha = 0;
    s.sel_valid = false;
    s.sel_nonempty = false;
    if (s.nu == 1 && !s.have_alpha) {
        const uint32_t c = L.col[0];
        const uint32_t r = c & 255u, g = (c >> 8) & 255u, b = (c >> 16) & 255u;
        evaluate(s, L, (g_c_om5[r][0] << 11) | (g_c_om6[g][0] << 5) | g_c_om5[b][0],
                 (g_c_om5[r][1] << 11) | (g_c_om6[g][1] << 5) | g_c_om5[b][1], 0);
        if (Lds::kUber && s.best.err)
            evaluate(s, L, (g_c_om5_3[r][0] << 11) | (g_c_om6_3[g][0] << 5) | g_c_om5_3[b][0],
                     (g_c_om5_3[r][1] << 11) | (g_c_om6_3[g][1] << 5) | g_c_om5_3[b][1], 0);
    } else {
        handle_multicolor(s, L);
    }
    compute_selectors(s, L);
    // return_solution (crn_dxt1.cpp:263-365): four-colour blocks keep low >
    // high (selectors through {1, 0, 3, 2}), 3-colour blocks low <= high
    // (through {1, 0, 2, 3})
    const bool inv = s.best.alpha ? s.best.lo > s.best.hi : s.best.lo < s.best.hi;
    const uint32_t lo = inv ? s.best.hi : s.best.lo, hi = inv ? s.best.lo : s.best.hi;
    if (selectors) {
        for (uint32_t i = lane; i < n; i += 64) {
            if ((px[p0 + i] >> 24) < thresh) {
                selectors[p0 + i] = 3;   // transparent (return_solution, crn_dxt1.cpp:300-343)
                continue;
            }
            const uint32_t key = (px[p0 + i] & 0x00FFFFFFu) | 0xFF000000u;
            uint32_t v = L.sel[unique_index(L, i, key)];
            if (inv) v = s.best.alpha ? (v < 2 ? v ^ 1u : v) : (v ^ 1u);
            selectors[p0 + i] = (uint8_t)v;
        }
    }
    if (lane == 0) {
        o.error = s.best.err;
        o.lo_hi = lo | (hi << 16);
        o.nunique = s.nu;
        o.flags = (uint32_t)inv | ((uint32_t)s.best.alpha << 1) | ((uint32_t)s.best.alt << 2) |
                  ((uint32_t)s.best.enforce << 3) | ((uint32_t)s.best.enforced << 4);
        out[cl] = o;
    }
}

template <uint32_t MaxU, bool Uber>
__global__ void __launch_bounds__(64) dxt1_cluster_kernel(const uint32_t* __restrict__ px,
                                                          const uint32_t* __restrict__ ofs,
                                                          const uint32_t* __restrict__ list,
                                                          const uint32_t* __restrict__ count, int perceptual,
                                                          ClusterOut* __restrict__ out,
                                                          uint8_t* __restrict__ selectors)
{
    typedef LdsU<std::conditional_t<(MaxU < kMaxU), LdsS<MaxU>, LdsT<MaxU>>, Uber> Lds;
    extern __shared__ __align__(16) uint8_t lds_raw[];
    Lds& L = *reinterpret_cast<Lds*>(lds_raw);
    if (blockIdx.x >= *count) return;
    const uint32_t cl = list[blockIdx.x];
    const uint32_t lane = lane_id();
    const uint32_t p0 = ofs[cl], p1 = ofs[cl + 1];
    const uint32_t n = p1 - p0;
    // DXT1A (qdxt1 with 3-colour blocks): texels with alpha below this are transparent
    const uint32_t thresh = Uber ? ((uint32_t)perceptual >> 8) & 255u : 0u;
    ClusterOut o{};
    if (!n) {
        if (lane == 0) out[cl] = o;
        return;
    }
    if constexpr (MaxU < kMaxU) {
        // this class holds at most MaxU texels (dxt1_cluster_classes_kernel)
        St s;
        s.perceptual = perceptual & 1;
        uint32_t opaque = n;
        s.nu = unique_sorted(L, px, p0, n, thresh, opaque);
        s.total_w = opaque;
        s.have_alpha = opaque != n;
        for (uint32_t i = lane; i < 3 * 64 * 3; i += 64) (&L.pl[0][0][0])[i] = 0;   // over the sort buffers
        __syncthreads();
        cluster_search(s, L, px, p0, n, perceptual, cl, out, selectors);
        return;
    }
    else {
    constexpr uint32_t kHashCap = LdsT<MaxU>::kCapH;
    for (uint32_t i = lane; i < kHashCap; i += 64) L.hkey[i] = 0;
    for (uint32_t i = lane; i < 3 * 64 * 3; i += 64) (&L.pl[0][0][0])[i] = 0;
    __builtin_amdgcn_wave_barrier();
    // unique colours in first-appearance order (one lane; the order fixes the float sums)
    St s;
    s.perceptual = perceptual & 1;
    s.nu = 0;
    uint32_t opaque = 0;
    bool overflow = false;
    if (lane == 0) {
        for (uint32_t i = 0; i < n; i++) {
            if ((px[p0 + i] >> 24) < thresh) continue;   // transparent (DXT1A)
            opaque++;
            const uint32_t key = (px[p0 + i] & 0x00FFFFFFu) | 0xFF000000u;
            uint32_t h = hash_slot(L, key);
            while (L.hkey[h] && L.hkey[h] != key) h = hash_next(L, h);
            if (L.hkey[h] == key) {
                L.w[L.hval[h]]++;
            } else if (s.nu < Lds::kCapU) {
                L.hkey[h] = key;
                L.hval[h] = (uint16_t)s.nu;
                L.col[s.nu] = key & 0x00FFFFFFu;
                L.w[s.nu] = 1;
                s.nu++;
            } else {
                overflow = true;
                break;
            }
        }
    }
    __builtin_amdgcn_wave_barrier();
    s.nu = (uint32_t)__shfl((int)s.nu, 0);
    s.total_w = (uint32_t)__shfl((int)opaque, 0);
    s.have_alpha = s.total_w != n;
    overflow = __shfl((int)overflow, 0) != 0;
    if (overflow) {
        // more distinct colours than this store holds: left to dxt1_cluster_big_kernel
        o.flags = kNotEncoded;
        if (lane == 0) out[cl] = o;
        return;
    }
    cluster_search(s, L, px, p0, n, perceptual, cl, out, selectors);
    }
}

// Clusters the LDS pass flagged (more than 4,096 distinct colours), one wave
// each, with the colour set in global scratch: the cluster's texels own
// scratch entries [p0, p1) of the texel-sized arrays and [2 p0, 2 p1) of the
// hash.  The first-appearance order is found in parallel: every texel inserts
// its colour with compare-and-swap and keeps the smallest texel index and
// the count per slot, the first texel of each colour is marked, and a prefix
// count over the marks gives each colour its rank.
template <bool Uber>
__global__ void __launch_bounds__(64) dxt1_cluster_big_kernel(const uint32_t* __restrict__ px,
                                                              const uint32_t* __restrict__ ofs, uint32_t nclusters,
                                                              int perceptual, ClusterOut* __restrict__ out,
                                                              uint8_t* __restrict__ selectors,
                                                              uint8_t* __restrict__ work)
{
    __shared__ GTables T;
    const uint32_t cl = blockIdx.x;
    if (cl >= nclusters || !(out[cl].flags & kNotEncoded)) return;
    const uint32_t lane = lane_id();
    const uint32_t p0 = ofs[cl], p1 = ofs[cl + 1], n = p1 - p0, total = ofs[nclusters];
    LdsU<GStore, Uber> L;
    L.kCapU = n;
    L.kCapH = 2 * n;
    // scratch layout, 37 bytes per texel of all clusters (`total`): u32
    // arrays col, w, mark of `total` entries, hkey, hval, hcnt of 2 * total,
    // then the u8 selectors of `total`
    uint32_t* base = reinterpret_cast<uint32_t*>(work);
    L.col = base + p0;
    L.w = base + (size_t)total + p0;
    L.mark = base + 2 * (size_t)total + p0;
    L.hkey = base + 3 * (size_t)total + 2 * (size_t)p0;
    L.hval = base + 5 * (size_t)total + 2 * (size_t)p0;
    L.hcnt = base + 7 * (size_t)total + 2 * (size_t)p0;
    L.sel = work + 36 * (size_t)total + p0;
    L.pl = T.pl;
    L.rlo = T.rlo;
    L.rhi = T.rhi;
    L.glo = T.glo;
    L.ghi = T.ghi;
    L.blo = T.blo;
    L.bhi = T.bhi;
    for (uint32_t i = lane; i < L.kCapH; i += 64) {
        L.hkey[i] = 0;
        L.hval[i] = 0xFFFFFFFFu;
        L.hcnt[i] = 0;
    }
    for (uint32_t i = lane; i < n; i += 64) L.mark[i] = 0;
    for (uint32_t i = lane; i < 3 * 64 * 3; i += 64) (&T.pl[0][0][0])[i] = 0;
    __threadfence();
    __syncthreads();
    const uint32_t thresh = Uber ? ((uint32_t)perceptual >> 8) & 255u : 0u;
    uint32_t opaque = 0;
    for (uint32_t i = lane; i < n; i += 64) {
        if ((px[p0 + i] >> 24) < thresh) continue;   // transparent (DXT1A)
        opaque++;
        const uint32_t key = (px[p0 + i] & 0x00FFFFFFu) | 0xFF000000u;
        uint32_t h = hash_slot(L, key);
        for (;;) {
            const uint32_t prev = atomicCAS(&L.hkey[h], 0u, key);
            if (prev == 0u || prev == key) {
                atomicMin(&L.hval[h], i);
                atomicAdd(&L.hcnt[h], 1u);
                break;
            }
            h = hash_next(L, h);
        }
    }
    __threadfence();
    __syncthreads();
    for (uint32_t h = lane; h < L.kCapH; h += 64)
        if (L.hkey[h]) L.mark[L.hval[h]] = 1;
    __threadfence();
    __syncthreads();
    // ranks in texel order: a running count over the marks, 64 texels per step
    uint32_t base_rank = 0;
    for (uint32_t i0 = 0; i0 < n; i0 += 64) {
        const uint32_t i = i0 + lane;
        const bool m = i < n && L.mark[i];
        const uint64_t bal = __ballot(m);
        const uint32_t r = base_rank + (uint32_t)__popcll(bal & ((1ull << lane) - 1ull));
        if (m) {
            const uint32_t key = (px[p0 + i] & 0x00FFFFFFu) | 0xFF000000u;
            uint32_t h = hash_slot(L, key);
            while (L.hkey[h] != key) h = hash_next(L, h);
            L.col[r] = key & 0x00FFFFFFu;
            L.w[r] = L.hcnt[h];
            L.hval[h] = r;
        }
        base_rank += (uint32_t)__popcll(bal);
    }
    __threadfence();
    __syncthreads();
    for (int off = 32; off; off >>= 1) opaque += (uint32_t)__shfl_xor((int)opaque, off);
    St s;
    s.perceptual = perceptual & 1;
    s.nu = base_rank;
    s.total_w = opaque;
    s.have_alpha = opaque != n;
    cluster_search(s, L, px, p0, n, perceptual, cl, out, selectors);
}

}  // namespace

namespace {
// Helper streams and events for crn2_dxt1_clusters_device, one set per device,
// kept until the process ends.  Callers hold g_side_mu from the fork to the
// join, so concurrent calls queue their work on the shared streams in turn.
std::mutex g_side_mu;

int side_streams(hipStream_t* side, hipEvent_t* fork, hipEvent_t* join)
{
    struct Set {
        bool ready = false;
        hipStream_t s[3];
        hipEvent_t fork, join[3];
    };
    static Set sets[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return CRN2_EHIP;
    Set& set = sets[dev];
    if (!set.ready) {
        for (int i = 0; i < 3; i++)
            if (hipStreamCreateWithFlags(&set.s[i], hipStreamNonBlocking) != hipSuccess) return CRN2_EHIP;
        if (hipEventCreateWithFlags(&set.fork, hipEventDisableTiming) != hipSuccess) return CRN2_EHIP;
        for (int i = 0; i < 3; i++)
            if (hipEventCreateWithFlags(&set.join[i], hipEventDisableTiming) != hipSuccess) return CRN2_EHIP;
        set.ready = true;
    }
    for (int i = 0; i < 3; i++) {
        side[i] = set.s[i];
        join[i] = set.join[i];
    }
    *fork = set.fork;
    return CRN2_OK;
}
}  // namespace

namespace {
// Scratch of crn2_dxt1_clusters_noalloc_device: four class counters, then four
// lists of at most kNoAllocMaxClusters cluster indices (1 MiB + 16 B on every
// GPU the library runs on, held until the process ends).  Host code never
// passes the array's name to a kernel: it takes the device address from
// hipGetSymbolAddress once per device (noalloc_scratch below).
constexpr uint32_t kNoAllocMaxClusters = 65536;
__device__ __align__(256) uint32_t scratch[4 + 4 * kNoAllocMaxClusters];

__global__ void clear_class_counts_kernel(uint32_t* __restrict__ count)
{
    if (threadIdx.x < 4) count[threadIdx.x] = 0;
}

constexpr int kMaxDevices = 64;
std::mutex g_init_mu;
bool g_init[kMaxDevices];
uint32_t* g_noalloc_scratch[kMaxDevices];

// Per-device set-up: the single-colour tables in device memory, the kernels'
// LDS limits and the device address of `scratch`.  Idempotent; runs on the
// device that is current.
template <bool Uber>
int set_lds()
{
    if (hipFuncSetAttribute(reinterpret_cast<const void*>(&dxt1_cluster_kernel<512, Uber>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)sizeof(LdsS<512>)) != hipSuccess ||
        hipFuncSetAttribute(reinterpret_cast<const void*>(&dxt1_cluster_kernel<1024, Uber>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)sizeof(LdsS<1024>)) != hipSuccess ||
        hipFuncSetAttribute(reinterpret_cast<const void*>(&dxt1_cluster_kernel<2048, Uber>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)sizeof(LdsS<2048>)) != hipSuccess ||
        hipFuncSetAttribute(reinterpret_cast<const void*>(&dxt1_cluster_kernel<kMaxU, Uber>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)sizeof(LdsT<kMaxU>)) != hipSuccess)
        return CRN2_EHIP;
    return CRN2_OK;
}

int clusters_init(int* dev_out)
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return CRN2_EHIP;
    *dev_out = dev;
    std::lock_guard<std::mutex> lk(g_init_mu);
    if (g_init[dev]) return CRN2_OK;
    // ryg single-colour tables (crn_ryg_dxt.cpp:58-99, 517-535), four- and
    // three-colour
    static uint8_t t5[256][2], t6[256][2], t5_3[256][2], t6_3[256][2];
    static bool tables = false;
    if (!tables) {
        for (int which = 0; which < 4; which++) {
            const int size = (which & 1) ? 64 : 32;
            const bool three = which >= 2;
            uint8_t (*t)[2] = which == 0 ? t5 : which == 1 ? t6 : which == 2 ? t5_3 : t6_3;
            for (int i = 0; i < 256; i++) {
                int best = 256;
                for (int mn = 0; mn < size; mn++)
                    for (int mx = 0; mx < size; mx++) {
                        const int me = size == 32 ? ((mn << 3) | (mn >> 2)) : ((mn << 2) | (mn >> 4));
                        const int xe = size == 32 ? ((mx << 3) | (mx >> 2)) : ((mx << 2) | (mx >> 4));
                        int err = three ? abs(((me + xe) >> 1) - i) : abs(((xe * 2 + me) / 3) - i);
                        err += (abs(xe - me) * 8) >> 8;
                        if (err < best) {
                            t[i][0] = (uint8_t)mx;
                            t[i][1] = (uint8_t)mn;
                            best = err;
                        }
                    }
            }
        }
        tables = true;
    }
    if (hipMemcpyToSymbol(HIP_SYMBOL(g_c_om5), t5, 512) != hipSuccess) return CRN2_EHIP;
    if (hipMemcpyToSymbol(HIP_SYMBOL(g_c_om6), t6, 512) != hipSuccess) return CRN2_EHIP;
    if (hipMemcpyToSymbol(HIP_SYMBOL(g_c_om5_3), t5_3, 512) != hipSuccess) return CRN2_EHIP;
    if (hipMemcpyToSymbol(HIP_SYMBOL(g_c_om6_3), t6_3, 512) != hipSuccess) return CRN2_EHIP;
    if (set_lds<false>() != CRN2_OK || set_lds<true>() != CRN2_OK) return CRN2_EHIP;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return CRN2_EHIP;
    g_noalloc_scratch[dev] = static_cast<uint32_t*>(p);
    g_init[dev] = true;
    return CRN2_OK;
}

template <bool Uber>
void launch_classes(const uint32_t* px, const uint32_t* offsets, uint32_t nclusters, const uint32_t* list,
                    const uint32_t* work, int perceptual, ClusterOut* o, uint8_t* selectors, const hipStream_t* cls)
{
    hipLaunchKernelGGL((dxt1_cluster_kernel<kMaxU, Uber>), dim3(nclusters), dim3(64), sizeof(LdsT<kMaxU>), cls[3], px,
                       offsets, list + 3 * (size_t)nclusters, work + 3, perceptual, o, selectors);
    hipLaunchKernelGGL((dxt1_cluster_kernel<2048, Uber>), dim3(nclusters), dim3(64), sizeof(LdsS<2048>), cls[2], px,
                       offsets, list + 2 * (size_t)nclusters, work + 2, perceptual, o, selectors);
    hipLaunchKernelGGL((dxt1_cluster_kernel<1024, Uber>), dim3(nclusters), dim3(64), sizeof(LdsS<1024>), cls[1], px,
                       offsets, list + (size_t)nclusters, work + 1, perceptual, o, selectors);
    hipLaunchKernelGGL((dxt1_cluster_kernel<512, Uber>), dim3(nclusters), dim3(64), sizeof(LdsS<512>), cls[0], px,
                       offsets, list, work, perceptual, o, selectors);
}

// The search proper.  work: 4 class counters + 4 lists of nclusters entries
// (device memory).  fork: run the four size classes side by side on helper
// streams; else every class runs on `st`, in turn (identical results).
int clusters_run(const void* rgba, const uint32_t* offsets, uint32_t nclusters, int flags, void* out,
                 uint8_t* selectors, hipStream_t st, uint32_t* work, bool fork)
{
    hipLaunchKernelGGL(clear_class_counts_kernel, dim3(1), dim3(64), 0, st, work);
    uint32_t* list = work + 4;
    const uint32_t* px = static_cast<const uint32_t*>(rgba);
    ClusterOut* o = static_cast<ClusterOut*>(out);
    hipLaunchKernelGGL(dxt1_cluster_classes_kernel, dim3((nclusters + 255) / 256), dim3(256), 0, st, offsets,
                       nclusters, work, list);
    hipStream_t cls[4] = {st, st, st, st};
    hipStream_t side[3];
    hipEvent_t fork_ev, join[3];
    std::unique_lock<std::mutex> lk(g_side_mu, std::defer_lock);
    int rc = CRN2_OK;
    if (fork) {
        // The four classes run side by side: the large clusters are few and
        // long, and alone would leave most CUs idle.  Three helper streams
        // (default priority, created once per device) fork from `st` and join
        // back into it.
        lk.lock();
        rc = side_streams(side, &fork_ev, join);
        if (rc == CRN2_OK && hipEventRecord(fork_ev, st) != hipSuccess) rc = CRN2_EHIP;
        for (int i = 0; rc == CRN2_OK && i < 3; i++)
            if (hipStreamWaitEvent(side[i], fork_ev, 0) != hipSuccess) rc = CRN2_EHIP;
        if (rc != CRN2_OK) return rc;
        cls[3] = side[0];
        cls[2] = side[1];
        cls[1] = side[2];
    }
    if (flags & 2)
        launch_classes<true>(px, offsets, nclusters, list, work, flags & 0xFF01, o, selectors, cls);
    else
        launch_classes<false>(px, offsets, nclusters, list, work, flags & 1, o, selectors, cls);
    if (hipGetLastError() != hipSuccess) rc = CRN2_EHIP;
    if (fork)
        for (int i = 0; i < 3; i++)
            if (hipEventRecord(join[i], side[i]) != hipSuccess || hipStreamWaitEvent(st, join[i], 0) != hipSuccess)
                rc = CRN2_EHIP;
    return rc;
}

// Second pass for clusters of more than 4,096 distinct colours (the LDS pass
// marks them kNotEncoded): reads the texel count once (the call waits for the
// stream there), and when some cluster may hold more colours than the largest
// LDS store, runs dxt1_cluster_big_kernel with stream-ordered global scratch.
int clusters_big_pass(const void* rgba, const uint32_t* offsets, uint32_t nclusters, int flags, void* out,
                      uint8_t* selectors, hipStream_t st)
{
    uint32_t total = 0;
    if (hipMemcpyAsync(&total, offsets + nclusters, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return CRN2_EHIP;
    if (total <= kMaxU) return CRN2_OK;      // no cluster can hold more colours than the LDS store
    uint8_t* big = nullptr;
    if (hipMallocAsync((void**)&big, 37 * (size_t)total, st) != hipSuccess) return CRN2_ENOMEM;
    if (flags & 2)
        hipLaunchKernelGGL(dxt1_cluster_big_kernel<true>, dim3(nclusters), dim3(64), 0, st,
                           static_cast<const uint32_t*>(rgba), offsets, nclusters, flags & 0xFF01,
                           static_cast<ClusterOut*>(out), selectors, big);
    else
        hipLaunchKernelGGL(dxt1_cluster_big_kernel<false>, dim3(nclusters), dim3(64), 0, st,
                           static_cast<const uint32_t*>(rgba), offsets, nclusters, flags & 1,
                           static_cast<ClusterOut*>(out), selectors, big);
    const int rc = hipGetLastError() == hipSuccess ? CRN2_OK : CRN2_EHIP;
    (void)hipFreeAsync(big, st);
    return rc;
}
}  // namespace

extern "C" int crn2_dxt1_clusters_device(const void* rgba, const uint32_t* offsets, uint32_t nclusters, int flags,
                                         void* out, uint8_t* selectors, void* stream)
{
    if (!rgba || !offsets || !out || (flags & ~0xFF03) || ((flags & 0xFF00) && !(flags & 2))) return CRN2_EINVAL;
    if (!nclusters) return CRN2_OK;
    int dev = 0;
    int rc = clusters_init(&dev);
    if (rc != CRN2_OK) return rc;
    hipStream_t st = static_cast<hipStream_t>(stream);
    uint32_t* work = nullptr;      // 4 class counts + 4 lists of nclusters
    if (hipMallocAsync((void**)&work, (4 + 4 * (size_t)nclusters) * sizeof(uint32_t), st) != hipSuccess)
        return CRN2_ENOMEM;
    rc = clusters_run(rgba, offsets, nclusters, flags, out, selectors, st, work, true);
    (void)hipFreeAsync(work, st);
    if (rc == CRN2_OK) rc = clusters_big_pass(rgba, offsets, nclusters, flags, out, selectors, st);
    return rc;
}

extern "C" int crn2_dxt1_clusters_noalloc_device(const void* rgba, const uint32_t* offsets, uint32_t nclusters,
                                                 int flags, void* out, uint8_t* selectors, void* stream)
{
    if (!rgba || !offsets || !out || (flags & ~0xFF03) || ((flags & 0xFF00) && !(flags & 2)) ||
        nclusters > kNoAllocMaxClusters)
        return CRN2_EINVAL;
    if (!nclusters) return CRN2_OK;
    int dev = 0;
    const int rc = clusters_init(&dev);
    if (rc != CRN2_OK) return rc;
    return clusters_run(rgba, offsets, nclusters, flags, out, selectors, static_cast<hipStream_t>(stream),
                        g_noalloc_scratch[dev], false);
}
