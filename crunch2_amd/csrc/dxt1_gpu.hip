// dxt1_gpu.hip -- DXT1 endpoint search (crnlib's dxt1_endpoint_optimizer,
// crnlib/crn_dxt1.cpp) for 16-texel blocks on gfx950, one lane per block.
//
// The search trajectory of crn_dxt1.cpp is reproduced step for step at every
// quality (superfast .. uber; probe tables and passes :716-746, the
// projection evaluator of the qualities below Better :1594-1757), with DXT1A
// transparency (:2116, :1380, :300-343) and the transparent-black trial of
// try_alpha_as_black_optimization (:2001-2079): unique colours in first-appearance order,
// perceptual weights (:1074-1108), PCA (:192-256), projection bounds and ray
// clamp (:1135-1175), k-means seeds with fast_random jitter (:1181-1308),
// probe sweeps along the axis, 26-neighbour and +-2 lattice moves, least-
// squares refinement (:525-698), solid-block tries (:93-153), per-channel
// exhaustive search (:369-486) and combinatorial recovery (:1886-1997).
// The float/double types and the order of every float expression follow the
// reference; the file is compiled with -ffp-contract=off.
//
// Two steps of the reference are result-invisible and are left out:
//   * the error-plane bound of evaluate_solution (:1316-1322) is a rigorous
//     lower bound on a candidate's error under the RGB metrics (every palette
//     entry lies between the endpoints on each channel), so a pruned
//     candidate would have been rejected by the full evaluation anyway.
//     Under grayscale sampling the error is a luma error, which the bound
//     does not bound: there it prunes candidates the evaluation would keep,
//     so it is evaluated (per unique colour, from the colours' planes);
//   * the tried-solutions set (:1324-1330) only skips re-evaluations, and a
//     re-evaluated candidate never replaces the best (strict <).
// State lives in registers: 16-entry arrays are only indexed from fully
// unrolled loops, selectors are 2-bit fields of one 32-bit word.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>

#include "autotune.h"
#include "crn2_internal.h"
#include "encode_set.h"

namespace {

__device__ __constant__ int16_t k_uber_probe[15] = {0, 1, 2, 3, 5, 7, 9, 10, 13, 15, 19, 27, 43, 59, 91};
__device__ __constant__ int16_t k_better_probe[10] = {0, 1, 2, 3, 5, 9, 15, 19, 27, 43};
__device__ __constant__ int16_t k_normal_probe[5] = {0, 1, 3, 5, 7};
__device__ __constant__ int16_t k_fast_probe[4] = {0, 1, 2, 3};
__device__ __constant__ int8_t k_adj[26][3] = {
    {-1, -1, -1}, {0, -1, -1}, {1, -1, -1}, {-1, 0, -1}, {0, 0, -1}, {1, 0, -1}, {-1, 1, -1}, {0, 1, -1},
    {1, 1, -1}, {-1, -1, 0}, {0, -1, 0}, {1, -1, 0}, {-1, 0, 0}, {1, 0, 0}, {-1, 1, 0}, {0, 1, 0},
    {1, 1, 0}, {-1, -1, 1}, {0, -1, 1}, {1, -1, 1}, {-1, 0, 1}, {0, 0, 1}, {1, 0, 1}, {-1, 1, 1},
    {0, 1, 1}, {1, 1, 1}};

// ryg single-colour tables, filled by crn2_dxt1_init (built on the host)
__device__ uint8_t g_om5[256][2], g_om6[256][2], g_om5_3[256][2], g_om6_3[256][2];

struct Sol {
    uint32_t lo, hi;
    uint64_t err;
    uint32_t sel;     // 2 bits per unique colour
    int nsel;         // 0: empty selector vector
    int alpha, alt, enforce, enforced;
};

struct Block {
    // parameters
    int quality, perceptual, use_alpha, grayscale, force_alpha, have_alpha, hc;
    // unique colours
    uint32_t nu, total_w;
    int r[16], g[16], b[16];
    uint32_t w[16];
    uint8_t pr[16], pg[16], pb[16]; // nearest 5/6/5-bit planes (grayscale bound)
    float nr[16], ng[16], nb[16];   // mean-subtracted normalised colours
    float mean[3], axis[3];
    Sol best, trial;
};

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

__device__ inline uint32_t dist(const Block& s, int dr, int dg, int db, int ar, int ag, int ab, int br, int bg, int bb)
{
    if (s.perceptual) return (uint32_t)(dr * dr * 8 + dg * dg * 25 + db * db);
    if (s.grayscale) {
        const int y0 = (ar * 19595 + ag * 38470 + ab * 7471 + 32768) >> 16;
        const int y1 = (br * 19595 + bg * 38470 + bb * 7471 + 32768) >> 16;
        return (uint32_t)((y0 - y1) * (y0 - y1));
    }
    return (uint32_t)(dr * dr + dg * dg + db * db);
}

__device__ inline uint32_t cd(const Block& s, int i, int pr, int pg, int pb)
{
    return dist(s, s.r[i] - pr, s.g[i] - pg, s.b[i] - pb, s.r[i], s.g[i], s.b[i], pr, pg, pb);
}

// evaluate_solution_uber (crn_dxt1.cpp:1370-1592)
__device__ bool eval_uber(Block& s, uint32_t lo, uint32_t hi, int alt)
{
    Sol& t = s.trial;
    t.lo = lo;
    t.hi = hi;
    t.nsel = (int)s.nu;
    t.err = s.best.err;
    t.alpha = 0;
    int first = 0, last = 1;
    if (s.have_alpha || s.force_alpha) first = 1;
    else if (!s.use_alpha) last = 0;
    int c0r, c0g, c0b, c1r, c1g, c1b;
    unpack565(lo, 1, c0r, c0g, c0b);
    unpack565(hi, 1, c1r, c1g, c1b);
    for (int type = first; type <= last; type++) {
        int c2r, c2g, c2b, c3r = 0, c3g = 0, c3b = 0;
        if (!type) {
            c2r = (c0r * 2 + c1r + alt) / 3; c2g = (c0g * 2 + c1g + alt) / 3; c2b = (c0b * 2 + c1b + alt) / 3;
            c3r = (c1r * 2 + c0r + alt) / 3; c3g = (c1g * 2 + c0g + alt) / 3; c3b = (c1b * 2 + c0b + alt) / 3;
        } else {
            c2r = (c0r + c1r + alt) >> 1; c2g = (c0g + c1g + alt) >> 1; c2b = (c0b + c1b + alt) >> 1;
        }
        uint64_t terr = 0;
        uint32_t tsel = 0;
        bool over = false;
#pragma unroll
        for (int k = 15; k >= 0; k--) {
            if (over || k >= (int)s.nu) continue;
            uint32_t be = cd(s, k, c0r, c0g, c0b), bi = 0;
            uint32_t e = cd(s, k, c1r, c1g, c1b);
            if (e < be) { be = e; bi = 1; }
            e = cd(s, k, c2r, c2g, c2b);
            if (e < be) { be = e; bi = 2; }
            if (!type) {
                e = cd(s, k, c3r, c3g, c3b);
                if (e < be) { be = e; bi = 3; }
            }
            terr += (uint64_t)be * s.w[k];
            if (terr >= t.err) { over = true; continue; }
            tsel |= bi << (2 * k);
        }
        if (!over && terr < t.err) {
            t.err = terr;
            t.alpha = type != 0;
            t.sel = tsel;
            t.alt = alt;
        }
    }
    t.enforce = !t.alpha && t.lo == t.hi;
    if (t.enforce) {
        int v;
        if ((t.lo & 31) != 31) { t.lo++; v = 1; }
        else { t.hi--; v = 0; }
        uint32_t all = 0;
#pragma unroll
        for (int k = 0; k < 16; k++) all |= (uint32_t)v << (2 * k);
        t.sel = all;
        t.enforced = v;
    }
    if (t.err < s.best.err) {
        s.best = t;
        s.best.nsel = (int)s.nu;
        return true;
    }
    return false;
}

// evaluate_solution_fast (crn_dxt1.cpp:1594-1757), the evaluator below
// cCRNDXTQualityBetter: each colour takes the palette entry whose projection
// onto the endpoint axis (green weighted 24 under the perceptual metric) is
// nearest, by midpoint thresholds, then one distance.  No candidate with the
// alternate rounding reaches it, so the trial's rounding flag stays 0.
__device__ bool eval_fast(Block& s, uint32_t lo, uint32_t hi)
{
    Sol& t = s.trial;
    t.lo = lo;
    t.hi = hi;
    t.nsel = (int)s.nu;
    t.err = s.best.err;
    t.alpha = 0;
    int first = 0, last = 1;
    if (s.have_alpha || s.force_alpha) first = 1;
    else if (!s.use_alpha) last = 0;
    int c0r, c0g, c0b, c1r, c1g, c1b;
    unpack565(lo, 1, c0r, c0g, c0b);
    unpack565(hi, 1, c1r, c1g, c1b);
    int vr = c1r - c0r, vg = c1g - c0g, vb = c1b - c0b;
    if (s.perceptual) {
        vr *= 8;
        vg *= 24;
    }
    const int s0 = c0r * vr + c0g * vg + c0b * vb, s1 = c1r * vr + c1g * vg + c1b * vb;
    const int dr = vr * 2, dg = vg * 2, db = vb * 2;
    for (int type = first; type <= last; type++) {
        int pr[4], pg[4], pb[4];
        pr[0] = c0r; pg[0] = c0g; pb[0] = c0b;
        pr[1] = c1r; pg[1] = c1g; pb[1] = c1b;
        int t0, t1, t2;   // thresholds of the projection (twice the stop midpoints)
        if (!type) {
            pr[2] = (c0r * 2 + c1r) / 3; pg[2] = (c0g * 2 + c1g) / 3; pb[2] = (c0b * 2 + c1b) / 3;
            pr[3] = (c1r * 2 + c0r) / 3; pg[3] = (c1g * 2 + c0g) / 3; pb[3] = (c1b * 2 + c0b) / 3;
            const int s2 = pr[2] * vr + pg[2] * vg + pb[2] * vb, s3 = pr[3] * vr + pg[3] * vg + pb[3] * vb;
            t0 = s2 + s0;      // below: entry 0, else 2
            t1 = s3 + s2;      // half point
            t2 = s1 + s3;      // below: entry 3, else 1
        } else {
            pr[2] = (c0r + c1r) >> 1; pg[2] = (c0g + c1g) >> 1; pb[2] = (c0b + c1b) >> 1;
            pr[3] = pg[3] = pb[3] = 0;
            const int s2 = pr[2] * vr + pg[2] * vg + pb[2] * vb;
            t0 = s0 + s2;
            t1 = s2 + s1;
            t2 = 0;
        }
        uint64_t terr = 0;
        uint32_t tsel = 0;
        bool over = false;
#pragma unroll
        for (int k = 15; k >= 0; k--) {
            if (over || k >= (int)s.nu) continue;
            const int dot = s.r[k] * dr + s.g[k] * dg + s.b[k] * db;
            uint32_t bi;
            if (!type) bi = dot >= t1 ? (dot < t2 ? 3u : 1u) : (dot < t0 ? 0u : 2u);
            else bi = dot < t0 ? 0u : (dot < t1 ? 2u : 1u);
            const int qr = bi == 0 ? pr[0] : bi == 1 ? pr[1] : bi == 2 ? pr[2] : pr[3];
            const int qg = bi == 0 ? pg[0] : bi == 1 ? pg[1] : bi == 2 ? pg[2] : pg[3];
            const int qb = bi == 0 ? pb[0] : bi == 1 ? pb[1] : bi == 2 ? pb[2] : pb[3];
            terr += (uint64_t)cd(s, k, qr, qg, qb) * s.w[k];
            if (terr >= t.err) { over = true; continue; }
            tsel |= bi << (2 * k);
        }
        if (!over && terr < t.err) {
            t.err = terr;
            t.alpha = type != 0;
            t.sel = tsel;
        }
    }
    t.alt = 0;
    t.enforce = !t.alpha && t.lo == t.hi;
    if (t.enforce) {
        int v;
        if ((t.lo & 31) != 31) { t.lo++; v = 1; }
        else { t.hi--; v = 0; }
        uint32_t all = 0;
#pragma unroll
        for (int k = 0; k < 16; k++) all |= (uint32_t)v << (2 * k);
        t.sel = all;
        t.enforced = v;
    }
    if (t.err < s.best.err) {
        s.best = t;
        s.best.nsel = (int)s.nu;
        return true;
    }
    return false;
}

// evaluate_solution_hc_* (crn_dxt1.cpp:1759-1835)
__device__ bool eval_hc(Block& s, uint32_t lo, uint32_t hi, int alt)
{
    int c0r, c0g, c0b, c1r, c1g, c1b;
    unpack565(lo, 1, c0r, c0g, c0b);
    unpack565(hi, 1, c1r, c1g, c1b);
    const int c2r = (c0r * 2 + c1r + alt) / 3, c2g = (c0g * 2 + c1g + alt) / 3, c2b = (c0b * 2 + c1b + alt) / 3;
    const int c3r = (c1r * 2 + c0r + alt) / 3, c3g = (c1g * 2 + c0g + alt) / 3, c3b = (c1b * 2 + c0b + alt) / 3;
    uint64_t err = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        if (k >= (int)s.nu) continue;
        const uint32_t e0 = cd(s, k, c0r, c0g, c0b), e1 = cd(s, k, c1r, c1g, c1b);
        const uint32_t e2 = cd(s, k, c2r, c2g, c2b), e3 = cd(s, k, c3r, c3g, c3b);
        const uint32_t e01 = min(e0, e1), e23 = min(e2, e3);
        err += (uint64_t)min(e01, e23) * s.w[k];
    }
    if (err >= s.best.err) return false;
    s.best.lo = lo;
    s.best.hi = hi;
    s.best.err = err;
    s.best.alpha = 0;
    s.best.alt = alt;
    s.best.enforce = lo == hi;
    if (s.best.enforce) {
        if ((s.best.lo & 31) != 31) { s.best.lo++; s.best.enforced = 1; }
        else { s.best.hi--; s.best.enforced = 0; }
    }
    return true;
}

__device__ void compute_selectors(Block& s)
{
    if (!s.hc) return;
    Sol& b = s.best;
    b.nsel = (int)s.nu;
    if (b.enforce) {
        uint32_t all = 0;
#pragma unroll
        for (int k = 0; k < 16; k++) all |= (uint32_t)b.enforced << (2 * k);
        b.sel = all;
        return;
    }
    int c0r, c0g, c0b, c1r, c1g, c1b;
    unpack565(b.lo, 1, c0r, c0g, c0b);
    unpack565(b.hi, 1, c1r, c1g, c1b);
    const int a = b.alt;
    const int c2r = (c0r * 2 + c1r + a) / 3, c2g = (c0g * 2 + c1g + a) / 3, c2b = (c0b * 2 + c1b + a) / 3;
    const int c3r = (c1r * 2 + c0r + a) / 3, c3g = (c1g * 2 + c0g + a) / 3, c3b = (c1b * 2 + c0b + a) / 3;
    uint32_t sel = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        if (k >= (int)s.nu) continue;
        const uint32_t e0 = cd(s, k, c0r, c0g, c0b), e1 = cd(s, k, c1r, c1g, c1b);
        const uint32_t e2 = cd(s, k, c2r, c2g, c2b), e3 = cd(s, k, c3r, c3g, c3b);
        const uint32_t e01 = min(e0, e1), e23 = min(e2, e3);
        const uint32_t v = e01 <= e23 ? (e01 == e0 ? 0u : 1u) : (e23 == e2 ? 2u : 3u);
        sel |= v << (2 * k);
    }
    b.sel = sel;
}

// the error planes' bound of evaluate_solution (crn_dxt1.cpp:1316-1322,
// planes :2134-2203): per channel, the colours whose plane is at or below the
// lower endpoint's measured against it, those above the higher endpoint's
// against that (the endpoints taken per channel in the order of the test
// c0 < c1); only consulted under grayscale sampling
__device__ uint64_t plane_bound(const Block& s, uint32_t lo, uint32_t hi)
{
    int ar, ag, ab, br, bg, bb;
    unpack565(lo, 0, ar, ag, ab);
    unpack565(hi, 0, br, bg, bb);
    const int Ar = (ar << 3) | (ar >> 2), Ag = (ag << 2) | (ag >> 4), Ab = (ab << 3) | (ab >> 2);
    const int Br = (br << 3) | (br >> 2), Bg = (bg << 2) | (bg >> 4), Bb = (bb << 3) | (bb >> 2);
    uint64_t e = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        if (k >= (int)s.nu) continue;
        const uint64_t w = s.w[k];
        int d;
        // red: a.r < b.r ? low(a) + high(b) : high(a) + low(b)
        if (ar < br) d = s.pr[k] <= ar ? s.r[k] - Ar : (s.pr[k] > br ? s.r[k] - Br : 0);
        else d = s.pr[k] > ar ? s.r[k] - Ar : (s.pr[k] <= br ? s.r[k] - Br : 0);
        e += w * (uint64_t)(d * d);
        if (ag < bg) d = s.pg[k] <= ag ? s.g[k] - Ag : (s.pg[k] > bg ? s.g[k] - Bg : 0);
        else d = s.pg[k] > ag ? s.g[k] - Ag : (s.pg[k] <= bg ? s.g[k] - Bg : 0);
        e += w * (uint64_t)(d * d);
        if (ab < bb) d = s.pb[k] <= ab ? s.b[k] - Ab : (s.pb[k] > bb ? s.b[k] - Bb : 0);
        else d = s.pb[k] > ab ? s.b[k] - Ab : (s.pb[k] <= bb ? s.b[k] - Bb : 0);
        e += w * (uint64_t)(d * d);
    }
    return e;
}

__device__ inline bool evaluate(Block& s, uint32_t lo, uint32_t hi, int alt)
{
    if (s.grayscale && plane_bound(s, lo, hi) >= s.best.err) return false;
    if (s.hc) return eval_hc(s, lo, hi, alt);
    return s.quality >= 3 ? eval_uber(s, lo, hi, alt) : eval_fast(s, lo, hi);
}

__device__ inline void canon(uint32_t& lo, uint32_t& hi)
{
    if (lo < hi) {
        const uint32_t t = lo; lo = hi; hi = t;
    }
}

__device__ inline bool eval_canon(Block& s, uint32_t lo, uint32_t hi)
{
    canon(lo, hi);
    return evaluate(s, lo, hi, 0);
}

__device__ inline int clampi(int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

// refine_solution (crn_dxt1.cpp:525-698), levels 0 and 1
__device__ bool refine(Block& s, int level)
{
    compute_selectors(s);
    double a0 = 0, a1 = 0, a2 = 0, t1r = 0, t1g = 0, t1b = 0, t2r = 0, t2g = 0, t2b = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        if (k >= (int)s.nu) continue;
        const double w = s.w[k];
        const double r = s.r[k] * w, g = s.g[k] * w, b = s.b[k] * w;
        const int step = (int)((s.best.sel >> (2 * k)) & 3u) ^ 1;
        const int w1 = step == 0 ? 3 : step == 1 ? 0 : step == 2 ? 2 : 1;
        const int q0 = step >= 2 ? 2 : 0;
        const int q1 = step == 0 ? 0 : step == 1 ? 9 : step == 2 ? 1 : 4;
        const int q2 = step == 0 ? 9 : step == 1 ? 0 : step == 2 ? 4 : 1;
        a0 += q0 * w;
        a1 += q1 * w;
        a2 += q2 * w;
        t1r += w1 * r; t1g += w1 * g; t1b += w1 * b;
        t2r += r; t2g += g; t2b += b;
    }
    t2r = 3 * t2r - t1r;
    t2g = 3 * t2g - t1g;
    t2b = 3 * t2b - t1b;
    const double xx = a2, yy = a1, xy = a0;
    const double t = xx * yy - xy * xy;
    if (!yy || !xx || fabs(t) < (double).0000125f) return false;
    const double frb = (double)(3.0f * 31.0f / 255.0f) / t;
    const double fg = frb * (double)(63.0f / 31.0f);
    int e0[3], e1[3];
    e0[0] = clampi((int)((t1r * yy - t2r * xy) * frb + (double)0.5f), 31);
    e0[1] = clampi((int)((t1g * yy - t2g * xy) * fg + (double)0.5f), 63);
    e0[2] = clampi((int)((t1b * yy - t2b * xy) * frb + (double)0.5f), 31);
    e1[0] = clampi((int)((t2r * xx - t1r * xy) * frb + (double)0.5f), 31);
    e1[1] = clampi((int)((t2g * xx - t1g * xy) * fg + (double)0.5f), 63);
    e1[2] = clampi((int)((t2b * xx - t1b * xy) * frb + (double)0.5f), 31);
    bool improved = false;
    if (level == 0) {
        uint32_t lo = (uint32_t)((e1[0] << 11) | (e1[1] << 5) | e1[2]);
        uint32_t hi = (uint32_t)((e0[0] << 11) | (e0[1] << 5) | e0[2]);
        improved |= eval_canon(s, lo, hi);
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
                        improved |= eval_canon(s, pack565(c0[0], c0[1], c0[2], 0), pack565(c1[0], c1[1], c1[2], 0));
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

__device__ bool try_median4(Block& s, const float* lo, const float* hi)
{
    float m[4][3];
    if (s.nu <= 4) {
        for (int i = 0; i < 4; i++) {
            const int k = min(i, (int)s.nu - 1);
            // k is at most 3: select through constant indices
            m[i][0] = k == 0 ? s.nr[0] : k == 1 ? s.nr[1] : k == 2 ? s.nr[2] : s.nr[3];
            m[i][1] = k == 0 ? s.ng[0] : k == 1 ? s.ng[1] : k == 2 ? s.ng[2] : s.ng[3];
            m[i][2] = k == 0 ? s.nb[0] : k == 1 ? s.nb[1] : k == 2 ? s.nb[2] : s.nb[3];
        }
    } else {
        for (int c = 0; c < 3; c++) {
            m[0][c] = lo[c] - s.mean[c];
            m[3][c] = hi[c] - s.mean[c];
        }
        for (int c = 0; c < 3; c++) {
            m[1][c] = m[0][c] + (m[3][c] - m[0][c]) * (1.0f / 3.0f);
            m[2][c] = m[0][c] + (m[3][c] - m[0][c]) * (2.0f / 3.0f);
        }
        FastRandom rm;
        uint32_t rover = 0;
        float prev = 1.0e+37f;
        for (uint32_t it = 0; it < 8; it++) {
            float nm[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            float nw[4] = {0, 0, 0, 0};
            float total = 0;
#pragma unroll
            for (int k = 0; k < 16; k++) {
                if (k >= (int)s.nu) continue;
                const float v0 = s.nr[k], v1 = s.ng[k], v2 = s.nb[k];
                float bd = 0;
                {
                    float d = m[0][0] - v0; bd += d * d;
                    d = m[0][1] - v1; bd += d * d;
                    d = m[0][2] - v2; bd += d * d;
                }
                int bi = 0;
                for (int j = 1; j < 4; j++) {
                    float dd = 0;
                    float d = m[j][0] - v0; dd += d * d;
                    d = m[j][1] - v1; dd += d * d;
                    d = m[j][2] - v2; dd += d * d;
                    if (dd < bd) { bd = dd; bi = j; }
                }
                total += bd;
                const float w = (float)s.w[k];
                for (int j = 0; j < 4; j++)
                    if (j == bi) {
                        nm[j][0] += v0 * w;
                        nm[j][1] += v1 * w;
                        nm[j][2] += v2 * w;
                        nw[j] += w;
                    }
            }
            uint32_t hidx = 0;
            float hw = 0;
            bool empty = false;
            for (int j = 0; j < 4; j++) {
                if (nw[j] > 0.0f) {
                    for (int c = 0; c < 3; c++) m[j][c] = nm[j][c] / nw[j];
                    if (nw[j] > hw) { hw = nw[j]; hidx = (uint32_t)j; }
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
                const uint32_t ri = (hidx + rover) & 3;
                rover++;
                for (int j = 0; j < 4; j++)
                    if (nw[j] == 0.0f) {
                        float src[3];
                        for (int c = 0; c < 3; c++)
                            src[c] = ri == 0 ? m[0][c] : ri == 1 ? m[1][c] : ri == 2 ? m[2][c] : m[3][c];
                        for (int c = 0; c < 3; c++) m[j][c] = src[c];
                        for (int c = 0; c < 3; c++) m[j][c] += rm.frand(-.00196f, .00196f);
                    }
            }
        }
    }
    bool improved = false;
    for (int i = 0; i < 3; i++)
        for (int j = i + 1; j < 4; j++) {
            float a[3], b[3];
            for (int c = 0; c < 3; c++) {
                a[c] = m[i][c] + s.mean[c];
                b[c] = m[j][c] + s.mean[c];
            }
            const uint32_t lo2 = pack565((int)floorf(.5f + a[0] * 31.0f), (int)floorf(.5f + a[1] * 63.0f),
                                         (int)floorf(.5f + a[2] * 31.0f), 0);
            const uint32_t hi2 = pack565((int)floorf(.5f + b[0] * 31.0f), (int)floorf(.5f + b[1] * 63.0f),
                                         (int)floorf(.5f + b[2] * 31.0f), 0);
            improved |= eval_canon(s, lo2, hi2);
        }
    improved |= refine(s, s.quality == 4 ? 1 : 0);
    return improved;
}

__device__ bool try_average_solid(Block& s)
{
    uint64_t tr = 0, tg = 0, tb = 0;
    uint32_t tw = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        if (k >= (int)s.nu) continue;
        tw += s.w[k];
        tr += (uint64_t)s.r[k] * s.w[k];
        tg += (uint64_t)s.g[k] * s.w[k];
        tb += (uint64_t)s.b[k] * s.w[k];
    }
    const uint32_t half = tw >> 1;
    const uint32_t ar = (uint32_t)((tr + half) / tw), ag = (uint32_t)((tg + half) / tw), ab = (uint32_t)((tb + half) / tw);
    bool improved = evaluate(s, (g_om5[ar][0] << 11) | (g_om6[ag][0] << 5) | g_om5[ab][0],
                             (g_om5[ar][1] << 11) | (g_om6[ag][1] << 5) | g_om5[ab][1], 0);
    if (s.use_alpha && s.best.err)
        improved |= evaluate(s, (g_om5_3[ar][0] << 11) | (g_om6_3[ag][0] << 5) | g_om5_3[ab][0],
                             (g_om5_3[ar][1] << 11) | (g_om6_3[ag][1] << 5) | g_om5_3[ab][1], 0);
    if (s.quality == 4) {
#pragma unroll
        for (int k = 0; k < 16; k++) {
            if (k >= (int)s.nu) continue;
            const uint32_t r = (uint32_t)s.r[k], g = (uint32_t)s.g[k], b = (uint32_t)s.b[k];
            if (r == ar && g == ag && b == ab) continue;
            improved |= evaluate(s, (g_om5[r][0] << 11) | (g_om6[g][0] << 5) | g_om5[b][0],
                                 (g_om5[r][1] << 11) | (g_om6[g][1] << 5) | g_om5[b][1], 0);
            if (s.use_alpha && s.best.err)
                improved |= evaluate(s, (g_om5_3[r][0] << 11) | (g_om6_3[g][0] << 5) | g_om5_3[b][0],
                                     (g_om5_3[r][1] << 11) | (g_om6_3[g][1] << 5) | g_om5_3[b][1], 0);
        }
    }
    return improved;
}

// compute_endpoint_component_errors + optimize_endpoint_comps (crn_dxt1.cpp:369-486).
// For 16 texels every table entry and every partial sum is below
// 16 * 255^2 * 4 < 2^32, so 32-bit arithmetic (wrapping in the intermediate
// W*p*p - WP2*p + WPP steps like the reference's 64-bit version) is exact.
// Per-channel error of one selector class as a function of the class's value
// p: W p^2 - WP2 p + WPP (exact in 32 bits: at most 16 texels x 255^2), the
// reference's compute_endpoint_component_errors tables (crn_dxt1.cpp:369-418)
// evaluated where they are read instead of stored.
struct QErr {
    uint32_t W[4], WP2[4], WPP[4];
    __device__ inline uint32_t at(int sl, uint32_t p) const { return W[sl] * p * p - WP2[sl] * p + WPP[sl]; }
};

__device__ inline uint32_t expand_comp(int comp, uint32_t c)
{
    return comp == 1 ? (((c << 2) | (c >> 4)) & 0xFF) : (((c << 3) | (c >> 2)) & 0xFF);
}

__device__ void comp_errors(const Block& s, int comp, QErr& q, uint32_t* brem)
{
    for (int k = 0; k < 4; k++) q.W[k] = q.WP2[k] = q.WPP[k] = 0;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        if (k >= (int)s.nu) continue;
        const uint32_t p = (uint32_t)(comp == 0 ? s.r[k] : comp == 1 ? s.g[k] : s.b[k]);
        const uint32_t w = s.w[k];
        const uint32_t sl = (s.best.sel >> (2 * k)) & 3u;
        for (uint32_t c = 0; c < 4; c++)
            if (c == sl) {
                q.W[c] += w;
                q.WP2[c] += w * p * 2;
                q.WPP[c] += w * p * p;
            }
    }
    const uint32_t lim = comp == 1 ? 64 : 32;
    for (int sl = 0; sl < 2; sl++) {
        uint32_t best = q.WPP[sl];
        for (uint32_t c = 1; c < lim; c++) {
            const uint32_t e = q.at(sl, expand_comp(comp, c));
            if (e < best) best = e;
        }
        brem[sl] = best;
    }
    for (int sl = 2; sl < 4; sl++) {
        uint32_t best = q.WPP[sl];
        for (uint32_t p = 1; p < 256; p++) {
            const uint32_t e = q.at(sl, p);
            if (e < best) best = e;
        }
        brem[sl] = best;
    }
    for (int sl = 3; sl; sl--) brem[sl - 1] += brem[sl];
}

__device__ void optimize_comps(Block& s)
{
    compute_selectors(s);
    if (s.best.alpha || !s.best.err) return;
    int slr, slg, slb, shr, shg, shb;
    unpack565(s.best.lo, 1, slr, slg, slb);
    unpack565(s.best.hi, 1, shr, shg, shb);
    uint32_t brem[4];
    QErr q;
    for (int comp = 0; comp < 3; comp++) {
        uint32_t p0 = (uint32_t)(comp == 0 ? slr : comp == 1 ? slg : slb);
        uint32_t p1 = (uint32_t)(comp == 0 ? shr : comp == 1 ? shg : shb);
        int low[3], high[3];
        unpack565(s.best.lo, 0, low[0], low[1], low[2]);
        unpack565(s.best.hi, 0, high[0], high[1], high[2]);
        comp_errors(s, comp, q, brem);
        uint32_t best = q.at(0, p0) + q.at(1, p1) + q.at(2, (p0 * 2 + p1) / 3) + q.at(3, (p0 + p1 * 2) / 3);
        if (brem[0] >= best) continue;
        const uint32_t lim = comp == 1 ? 64 : 32;
        for (uint32_t c0 = 0; c0 < lim; c0++) {
            p0 = expand_comp(comp, c0);
            uint32_t e0 = q.at(0, p0);
            if (e0 + brem[1] >= best) continue;
            if (comp == 0) low[0] = (int)c0; else if (comp == 1) low[1] = (int)c0; else low[2] = (int)c0;
            const uint32_t plow = pack565(low[0], low[1], low[2], 0);
            for (uint32_t c1 = 0; c1 < lim; c1++) {
                p1 = expand_comp(comp, c1);
                uint32_t e = e0 + q.at(1, p1);
                if (e + brem[2] >= best) continue;
                e += q.at(2, (p0 * 2 + p1) / 3);
                if (e + brem[3] >= best) continue;
                e += q.at(3, (p0 + p1 * 2) / 3);
                if (e >= best) continue;
                if (comp == 0) high[0] = (int)c1; else if (comp == 1) high[1] = (int)c1; else high[2] = (int)c1;
                if (!evaluate(s, plow, pack565(high[0], high[1], high[2], 0), 0)) continue;
                if (!s.best.err) return;
                compute_selectors(s);
                comp_errors(s, comp, q, brem);
                best = q.at(0, p0) + q.at(1, p1) + q.at(2, (p0 * 2 + p1) / 3) + q.at(3, (p0 + p1 * 2) / 3);
                e0 = q.at(0, p0);
                if (e0 + brem[1] >= best) break;
            }
        }
    }
}

__device__ inline void lerp_color(int ar, int ag, int ab, int br, int bg, int bb, float f, int rounding, int& r, int& g,
                                  int& b)
{
    const float rd = rounding ? 1.0f : 0.0f;
    r = clampi((int)(rd + ((float)ar + ((float)br - (float)ar) * f)), 255);
    g = clampi((int)(rd + ((float)ag + ((float)bg - (float)ag) * f)), 255);
    b = clampi((int)(rd + ((float)ab + ((float)bb - (float)ab) * f)), 255);
}

__device__ void try_combinatorial(Block& s)
{
    if (s.nu < 2 || s.nu > 4) return;
    uint32_t packed[64];
    uint32_t np = 0;
    auto add = [&](int r, int g, int b) {
        const uint32_t p = pack565(r, g, b, 1);
        for (uint32_t k = 0; k < np; k++)
            if (packed[k] == p) return;
        packed[np++] = p;
    };
    for (int k = 0; k < 4; k++)
        if (k < (int)s.nu) add(s.r[k], s.g[k], s.b[k]);
    if (s.nu == 2) {
        const float fs[10] = {2.0f, 3.0f, .5f, 1.5f, -1.0f, 2.0f, -.5f, .5f, -2.0f, -1.0f};
        for (int k = 0; k < 2; k++)
            for (int q = 0; q < 2; q++) {
                const int a = q, b2 = q ^ 1;
                const int ar = a ? s.r[1] : s.r[0], ag = a ? s.g[1] : s.g[0], ab = a ? s.b[1] : s.b[0];
                const int br = b2 ? s.r[1] : s.r[0], bg = b2 ? s.g[1] : s.g[0], bb = b2 ? s.b[1] : s.b[0];
                for (int f = 0; f < 10; f++) {
                    int r, g, b;
                    lerp_color(ar, ag, ab, br, bg, bb, fs[f], k, r, g, b);
                    add(r, g, b);
                }
            }
    } else if (s.nu == 3) {
        const float fs[4] = {1.5f, 2.0f / 3.0f, 1.0f / 3.0f, -.5f};
        for (int i = 0; i <= 2; i++)
            for (int j = 0; j <= 2; j++) {
                if (i == j) continue;
                const int ar = i == 0 ? s.r[0] : i == 1 ? s.r[1] : s.r[2];
                const int ag = i == 0 ? s.g[0] : i == 1 ? s.g[1] : s.g[2];
                const int ab = i == 0 ? s.b[0] : i == 1 ? s.b[1] : s.b[2];
                const int br = j == 0 ? s.r[0] : j == 1 ? s.r[1] : s.r[2];
                const int bg = j == 0 ? s.g[0] : j == 1 ? s.g[1] : s.g[2];
                const int bb = j == 0 ? s.b[0] : j == 1 ? s.b[1] : s.b[2];
                for (int f = 0; f < 4; f++) {
                    int r, g, b;
                    lerp_color(ar, ag, ab, br, bg, bb, fs[f], 1, r, g, b);
                    add(r, g, b);
                }
            }
    }
    for (uint32_t i = 0; s.best.err && i + 1 < np; i++)
        for (uint32_t j = i + 1; s.best.err && j < np; j++) evaluate(s, packed[i], packed[j], 0);
    const uint64_t err = s.best.err;
    if (err) s.best.err = 1;
    for (uint32_t i = 0; s.best.err && i + 1 < np; i++)
        for (uint32_t j = i + 1; s.best.err && j < np; j++) evaluate(s, packed[i], packed[j], 1);
    if (s.best.err) s.best.err = err;
}

__device__ void sweep(const int16_t* tab, uint32_t range, const float* base, const float (*ax)[3], uint32_t* out,
                      uint32_t& n)
{
    int prev[2] = {-1, -1};
    n = 0;
    const float i0 = base[0] + .5f, i1 = base[1] + .5f, i2 = base[2] + .5f;
    for (uint32_t i = 0; i < range; i++) {
        const int ls = i ? 0 : 1;
        const int x = tab[i];
        for (int sg = ls; sg < 2; sg++) {
            const float pr = i0 + ax[sg][0] * (float)x;
            const float pg = i1 + ax[sg][1] * (float)x;
            const float pb = i2 + ax[sg][2] * (float)x;
            const int r = clampi((int)floorf(pr), 31), g = clampi((int)floorf(pg), 63), b = clampi((int)floorf(pb), 31);
            const int packed = b | (g << 5) | (r << 11);
            if (packed != prev[sg]) {
                out[n++] = (uint32_t)packed;
                prev[sg] = packed;
            }
        }
    }
}

__device__ void lattice(Block& s, int which)
{
    int lr, lg, lb;
    unpack565(which ? s.best.hi : s.best.lo, 0, lr, lg, lb);
    for (int i = 0; i < 26; i++) {
        const int r = lr + k_adj[i][0];
        if (r < 0 || r > 31) continue;
        const int g = lg + k_adj[i][1];
        if (g < 0 || g > 63) continue;
        const int b = lb + k_adj[i][2];
        if (b < 0 || b > 31) continue;
        const uint32_t p = pack565(r, g, b, 0);
        if (which) eval_canon(s, s.best.lo, p);
        else eval_canon(s, p, s.best.hi);
    }
    if (s.quality == 4) {
        unpack565(which ? s.best.hi : s.best.lo, 0, lr, lg, lb);
        for (int a = 0; a < 3; a++) {
            const int limit = a == 1 ? 63 : 31;
            for (int sg = -2; sg <= 2; sg += 4) {
                int c[3] = {lr, lg, lb};
                const int q = c[a] + sg;
                if (q < 0 || q > limit) continue;
                c[a] = q;
                const uint32_t p = pack565(c[0], c[1], c[2], 0);
                if (which) eval_canon(s, s.best.lo, p);
                else eval_canon(s, p, s.best.hi);
            }
        }
    }
}

__device__ bool sel_all_equal(const Sol& b)
{
    if (!b.nsel) return false;
    const uint32_t f = b.sel & 3u;
    for (int k = 1; k < 16; k++)
        if (k < b.nsel && ((b.sel >> (2 * k)) & 3u) != f) return false;
    return true;
}

__device__ void optimize_endpoints(Block& s, float* lo, float* hi)
{
    float olo[3] = {lo[0], lo[1], lo[2]}, ohi[3] = {hi[0], hi[1], hi[2]};
    const int16_t* tab = k_uber_probe;
    uint32_t range = 15, passes = 4;
    float dpt = .015625f;
    if (s.quality == 3) {
        tab = k_better_probe;
        range = 10;
        passes = 2;
    } else if (s.quality == 2) {
        tab = k_normal_probe;
        range = 5;
        passes = 2;
        dpt = .027063293f;
    } else if (s.quality < 2) {
        tab = k_fast_probe;
        range = 4;
        passes = s.quality == 0 ? 1 : 2;
        dpt = .027063293f;
    }
    if (s.quality >= 3) try_median4(s, olo, ohi);
    float ax[2][3];
    for (int c = 0; c < 3; c++) ax[1][c] = s.axis[c] * dpt;
    ax[1][0] *= 31.0f;
    ax[1][1] *= 63.0f;
    ax[1][2] *= 31.0f;
    for (int c = 0; c < 3; c++) ax[0][c] = -ax[1][c];
    const float lim[3] = {31.0f, 63.0f, 31.0f};
    for (int c = 0; c < 3; c++) {
        const float a = lo[c] * lim[c];
        lo[c] = a < 0.0f ? 0.0f : (a > lim[c] ? lim[c] : a);
        const float b = hi[c] * lim[c];
        hi[c] = b < 0.0f ? 0.0f : (b > lim[c] ? lim[c] : b);
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
        uint32_t nl = 0, nh = 0;
        sweep(tab, range, lo, ax, plow, nl);
        sweep(tab, range, hi, ax, phigh, nh);
        for (uint32_t i = 0; i < nl; i++)
            for (uint32_t j = 0; j < nh; j++) eval_canon(s, plow[i], phigh[j]);
        if (s.quality >= 2) {
            lattice(s, 0);
            lattice(s, 1);
        }
        if (!s.best.err || (pass && s.best.err == prev)) break;
        if (s.quality >= 4) refine(s, 1);
    }
    if (s.quality >= 2) {
        if (s.best.err && !s.have_alpha) {
            bool solid = false;
            if (sel_all_equal(s.best)) solid = try_average_solid(s);
            if (!solid && s.quality == 4) optimize_comps(s);
        }
        if (s.quality == 4 && s.best.err) try_combinatorial(s);
    }
}

__device__ void compute_vectors(Block& s, const float* pw, float* meanw, float (*nw)[3])
{
    float m[3] = {0, 0, 0}, mw[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < 16; k++) {
        if (k >= (int)s.nu) continue;
        const float w = (float)s.w[k];
        const float n0 = s.r[k] * 1.0f / 255.0f, n1 = s.g[k] * 1.0f / 255.0f, n2 = s.b[k] * 1.0f / 255.0f;
        s.nr[k] = n0;
        s.ng[k] = n1;
        s.nb[k] = n2;
        nw[k][0] = pw[0] * n0;
        nw[k][1] = pw[1] * n1;
        nw[k][2] = pw[2] * n2;
        m[0] += n0 * w; m[1] += n1 * w; m[2] += n2 * w;
        mw[0] += nw[k][0] * w; mw[1] += nw[k][1] * w; mw[2] += nw[k][2] * w;
    }
    if (s.total_w) {
        const float inv = 1.0f / (float)s.total_w;
        for (int c = 0; c < 3; c++) {
            m[c] *= inv;
            mw[c] *= inv;
        }
    }
#pragma unroll
    for (int k = 0; k < 16; k++) {
        if (k >= (int)s.nu) continue;
        s.nr[k] -= m[0];
        s.ng[k] -= m[1];
        s.nb[k] -= m[2];
        nw[k][0] -= mw[0];
        nw[k][1] -= mw[1];
        nw[k][2] -= mw[2];
    }
    for (int c = 0; c < 3; c++) {
        s.mean[c] = m[c];
        meanw[c] = mw[c];
    }
}

__device__ void compute_pca(Block& s, const float (*nc)[3], float* axis)
{
    double cov[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k < 16; k++) {
        if (k >= (int)s.nu) continue;
        const float r = nc[k][0], g = nc[k][1], b = nc[k][2];
        if (s.w[k] > 1) {
            const double w = s.w[k];
            cov[0] += (double)(r * r) * w; cov[1] += (double)(r * g) * w; cov[2] += (double)(r * b) * w;
            cov[3] += (double)(g * g) * w; cov[4] += (double)(g * b) * w; cov[5] += (double)(b * b) * w;
        } else {
            cov[0] += (double)(r * r); cov[1] += (double)(r * g); cov[2] += (double)(r * b);
            cov[3] += (double)(g * g); cov[4] += (double)(g * b); cov[5] += (double)(b * b);
        }
    }
    double vr = (double).9f, vg = 1.0, vb = (double).7f;
    for (int it = 0; it < 8; it++) {
        double r = vr * cov[0] + vg * cov[1] + vb * cov[2];
        double g = vr * cov[1] + vg * cov[3] + vb * cov[4];
        double b = vr * cov[2] + vg * cov[4] + vb * cov[5];
        double m = fabs(r);
        if (fabs(g) > m) m = fabs(g);
        if (fabs(b) > m) m = fabs(b);
        if (m > 1e-10) {
            m = 1.0 / m;
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
    len = 1.0 / sqrt(len);
    axis[0] = (float)(vr * len);
    axis[1] = (float)(vg * len);
    axis[2] = (float)(vb * len);
}

__device__ bool ray_unit_box(float* coord, const float* org, const float* dir)
{
    int quad[3];
    float cand[3], maxt[3];
    bool inside = true;
    for (int i = 0; i < 3; i++) {
        if (org[i] < 0.0f) { quad[i] = 1; cand[i] = 0.0f; inside = false; }
        else if (org[i] > 1.0f) { quad[i] = 0; cand[i] = 1.0f; inside = false; }
        else { quad[i] = 2; cand[i] = 0.0f; }
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
    coord[0] = c[0];
    coord[1] = c[1];
    coord[2] = c[2];
    return true;
}

__device__ void handle_multicolor(Block& s)
{
    uint32_t npass = 1;
    float pw[3] = {1.0f, 1.0f, 1.0f};
    if (s.perceptual) {
        float red = 0, blue = 0, al = 0;
#pragma unroll
        for (int k = 0; k < 16; k++) {
            if (k >= (int)s.nu) continue;
            const int l = (s.r[k] + s.g[k] + s.b[k] + 1) / 3;
            const float lf = (float)l;
            const float scale = (float)s.w[k] / (lf > 1.0f ? lf : 1.0f);
            red += scale * (float)s.r[k];
            blue += scale * (float)s.b[k];
            al += (float)l;
        }
        red /= (float)s.total_w;
        blue /= (float)s.total_w;
        al /= (float)s.total_w;
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
            for (int c = 0; c < 3; c++) pw[c] = lum[c] + (pw[c] - lum[c]) * p;
        }
    }
    float nw[16][3];
    float meanw[3];
    for (uint32_t pass = 0; pass < npass; pass++) {
        compute_vectors(s, pw, meanw, nw);
        compute_pca(s, nw, s.axis);
        s.axis[0] /= pw[0];
        s.axis[1] /= pw[1];
        s.axis[2] /= pw[2];
        double n = (double)(s.axis[0] * s.axis[0]);
        n += (double)(s.axis[1] * s.axis[1]);
        n += (double)(s.axis[2] * s.axis[2]);
        if (n != 0) {
            const float k = (float)(1.0 / sqrt(n));
            s.axis[0] *= k;
            s.axis[1] *= k;
            s.axis[2] *= k;
        }
        if (npass > 1) {
            if (fabsf(s.axis[0]) >= .795f) { pw[0] = .424f; pw[1] = .6f; pw[2] = .072f; }
            else if (fabsf(s.axis[2]) >= .795f) { pw[0] = .212f; pw[1] = .6f; pw[2] = .212f; }
            else break;
        }
    }
    float l = 1e+9f, h = -1e+9f;
#pragma unroll
    for (int k = 0; k < 16; k++) {
        if (k >= (int)s.nu) continue;
        float d = s.nr[k] * s.axis[0];
        d += s.ng[k] * s.axis[1];
        d += s.nb[k] * s.axis[2];
        l = d < l ? d : l;
        h = h < d ? d : h;
    }
    float lo[3], hi[3];
    for (int c = 0; c < 3; c++) {
        lo[c] = s.mean[c] + l * s.axis[c];
        hi[c] = s.mean[c] + h * s.axis[c];
    }
    bool in = true;
    for (int c = 0; c < 3; c++)
        if (lo[c] < 0.0f || lo[c] > 1.0f) in = false;
    if (!in) {
        float c3[3];
        if (ray_unit_box(c3, lo, s.axis)) { lo[0] = c3[0]; lo[1] = c3[1]; lo[2] = c3[2]; }
    }
    in = true;
    for (int c = 0; c < 3; c++)
        if (hi[c] < 0.0f || hi[c] > 1.0f) in = false;
    if (!in) {
        float c3[3];
        const float nd[3] = {-s.axis[0], -s.axis[1], -s.axis[2]};
        if (ray_unit_box(c3, hi, nd)) { hi[0] = c3[0]; hi[1] = c3[1]; hi[2] = c3[2]; }
    }
    optimize_endpoints(s, lo, hi);
}

struct Dxt1Out {
    uint64_t error;
    uint32_t lo_hi;    // low | high << 16 (results::m_low_color / m_high_color)
    uint32_t selectors;// 2-bit DXT1 code of texel i at bit 2*i
    uint32_t flags;    // alpha_block | reordered << 1 | alternate_rounding << 2 | enforce << 3 | enforced << 4
    uint32_t pad;
};

// compute_internal + return_solution (crn_dxt1.cpp:2081-2232, 263-365).
// DXT1A (flags bit 4): texels with alpha below the threshold (flags bits
// 8..15) leave the colour set and take selector 3.
__device__ Dxt1Out encode_core(const uint32_t* px, int quality, int flags)
{
    Block s;
    s.quality = quality;
    const int perceptual_req = flags & 1;
    s.use_alpha = (flags >> 1) & 1;
    s.grayscale = (flags >> 2) & 1;
    s.force_alpha = (flags >> 3) & 1;
    s.have_alpha = (flags >> 4) & 1;
    const uint32_t thresh = ((uint32_t)flags >> 8) & 255u;
    s.hc = quality == 4 && !s.have_alpha && !s.force_alpha && !s.use_alpha && !s.grayscale;
    s.perceptual = perceptual_req && !s.grayscale;
    s.nu = 0;
    s.total_w = 0;
    uint32_t keys[16];
    uint32_t tmask = 0;  // transparent texels
    uint64_t map64 = 0;  // 4 bits per texel: unique index
#pragma unroll
    for (int i = 0; i < 16; i++) {
        if (s.have_alpha && (px[i] >> 24) < thresh) {
            tmask |= 1u << i;
            continue;
        }
        const uint32_t key = px[i] | 0xFF000000u;
        int k = -1;
#pragma unroll
        for (int u = 0; u < 16; u++)
            if (k < 0 && u < (int)s.nu && keys[u] == key) k = u;
        if (k < 0) {
            k = (int)s.nu;
#pragma unroll
            for (int u = 0; u < 16; u++)
                if (u == k) {
                    keys[u] = key;
                    const uint32_t R = key & 0xFF, G = (key >> 8) & 0xFF, B = (key >> 16) & 0xFF;
                    s.r[u] = (int)R;
                    s.g[u] = (int)G;
                    s.b[u] = (int)B;
                    s.pr[u] = (uint8_t)((R >> 3) + (((R & 7) > (R >> 5)) ? 1 : 0));
                    s.pg[u] = (uint8_t)((G >> 2) + (((G & 3) > (G >> 6)) ? 1 : 0));
                    s.pb[u] = (uint8_t)((B >> 3) + (((B & 7) > (B >> 5)) ? 1 : 0));
                    s.w[u] = 1;
                }
            s.nu++;
        } else {
#pragma unroll
            for (int u = 0; u < 16; u++)
                if (u == k) s.w[u]++;
        }
        s.total_w++;
        map64 |= (uint64_t)k << (4 * i);
    }
    Dxt1Out o;
    o.pad = 0;
    if (!s.nu) {
        // no opaque texel: transparent block (the reference leaves m_error unset; 0 here)
        o.error = 0;
        o.lo_hi = 0;
        o.selectors = 0xFFFFFFFFu;
        o.flags = 1;
        return o;
    }
    s.best.lo = s.best.hi = 0;
    s.best.nsel = 0;
    s.best.err = ~0ull;
    s.best.alpha = 0;
    s.best.sel = 0;
    s.best.alt = s.best.enforce = s.best.enforced = 0;
    s.trial = s.best;
    if (s.nu == 1 && !tmask) {
        const uint32_t r = (uint32_t)s.r[0], g = (uint32_t)s.g[0], b = (uint32_t)s.b[0];
        evaluate(s, (g_om5[r][0] << 11) | (g_om6[g][0] << 5) | g_om5[b][0], (g_om5[r][1] << 11) | (g_om6[g][1] << 5) | g_om5[b][1], 0);
        if (s.use_alpha && s.best.err)
            evaluate(s, (g_om5_3[r][0] << 11) | (g_om6_3[g][0] << 5) | g_om5_3[b][0],
                     (g_om5_3[r][1] << 11) | (g_om6_3[g][1] << 5) | g_om5_3[b][1], 0);
    } else {
        handle_multicolor(s);
    }
    compute_selectors(s);
    const bool inv = s.best.alpha ? (s.best.lo > s.best.hi) : (s.best.lo < s.best.hi);
    o.error = s.best.err;
    const uint32_t lo = inv ? s.best.hi : s.best.lo, hi = inv ? s.best.lo : s.best.hi;
    o.lo_hi = lo | (hi << 16);
    uint32_t sel = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        const uint32_t u = (uint32_t)((map64 >> (4 * i)) & 15u);
        uint32_t v = (s.best.sel >> (2 * u)) & 3u;
        if (inv) v = s.best.alpha ? (v < 2 ? v ^ 1u : v) : (v ^ 1u);
        if ((tmask >> i) & 1u) v = 3u;
        sel |= v << (2 * i);
    }
    o.selectors = sel;
    o.flags = (uint32_t)s.best.alpha | ((uint32_t)inv << 1) | ((uint32_t)s.best.alt << 2) |
              ((uint32_t)s.best.enforce << 3) | ((uint32_t)s.best.enforced << 4);
    return o;
}

__device__ inline uint32_t metric(int perceptual, int grayscale, uint32_t p, int r, int g, int b)
{
    const int pr = (int)(p & 255u), pg = (int)((p >> 8) & 255u), pb = (int)((p >> 16) & 255u);
    const int dr = pr - r, dg = pg - g, db = pb - b;
    if (perceptual) return (uint32_t)(dr * dr * 8 + dg * dg * 25 + db * db);
    if (grayscale) {
        const int y0 = (pr * 19595 + pg * 38470 + pb * 7471 + 32768) >> 16;
        const int y1 = (r * 19595 + g * 38470 + b * 7471 + 32768) >> 16;
        return (uint32_t)((y0 - y1) * (y0 - y1));
    }
    return (uint32_t)(dr * dr + dg * dg + db * db);
}

// dxt1_endpoint_optimizer::compute (crn_dxt1.cpp:2234-2246).  flags bit 6
// applies dxt_image's DXT1A rule first (crn_dxt_image.cpp:1439-1450: the
// block has alpha when a texel's alpha is below the threshold); bit 5 with
// alpha blocks allowed and no transparent texel runs
// try_alpha_as_black_optimization (:2001-2079): texels whose r, g and b are
// all <= 4 are made transparent, the block is searched again as DXT1A, and
// that 3-colour block replaces the first result if its error, counting those
// texels against black, is lower.  One copy of the search serves both runs.
__device__ Dxt1Out encode_block(const uint32_t* px, int quality, int flags)
{
    if ((flags >> 6) & 1) {
        const uint32_t thresh = ((uint32_t)flags >> 8) & 255u;
        bool any = false;
#pragma unroll
        for (int i = 0; i < 16; i++) any |= (px[i] >> 24) < thresh;
        flags = (flags & ~(64 | 16)) | (any ? 16 : 0);
    }
    uint32_t p[16];
    uint32_t dark = 0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
        p[i] = px[i];
        if ((p[i] & 255u) <= 4 && ((p[i] >> 8) & 255u) <= 4 && ((p[i] >> 16) & 255u) <= 4) dark |= 1u << i;
    }
    const bool black = ((flags >> 1) & 1) && ((flags >> 5) & 1) && !((flags >> 4) & 1) && dark && dark != 0xFFFFu;
    Dxt1Out first;
    for (int pass = 0;; pass++) {
        const Dxt1Out o = encode_core(p, quality, pass ? (flags | 16) : flags);
        if (!pass) {
            first = o;
            if (!black) return first;
#pragma unroll
            for (int i = 0; i < 16; i++)
                if ((dark >> i) & 1u) p[i] &= 0x00FFFFFFu;
            continue;
        }
        // get_block_colors3 of the trial's endpoints, entry 3 black
        int c0r, c0g, c0b, c1r, c1g, c1b;
        unpack565(o.lo_hi & 0xFFFFu, 1, c0r, c0g, c0b);
        unpack565(o.lo_hi >> 16, 1, c1r, c1g, c1b);
        const int c2r = (c0r + c1r) >> 1, c2g = (c0g + c1g) >> 1, c2b = (c0b + c1b) >> 1;
        const int gray = (flags >> 2) & 1, perc = (flags & 1) && !gray;
        uint64_t err = 0;
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const uint32_t v = (o.selectors >> (2 * i)) & 3u;
            const int r = v == 0 ? c0r : v == 1 ? c1r : v == 2 ? c2r : 0;
            const int g = v == 0 ? c0g : v == 1 ? c1g : v == 2 ? c2g : 0;
            const int b = v == 0 ? c0b : v == 1 ? c1b : v == 2 ? c2b : 0;
            err += metric(perc, gray, p[i], r, g, b);
        }
        if (err < first.error) {
            first.error = err;
            first.lo_hi = o.lo_hi;
            first.selectors = o.selectors;
            first.flags |= 1u;
        }
        return first;
    }
}

__global__ void __launch_bounds__(256, 2) dxt1_blocks_kernel(const uint32_t* __restrict__ px, uint32_t nblocks, int quality,
                                                         int flags, Dxt1Out* __restrict__ out)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblocks) return;
    uint32_t p[16];
    const uint4* src = reinterpret_cast<const uint4*>(px + 16 * (size_t)b);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const uint4 v = src[q];
        p[4 * q] = v.x;
        p[4 * q + 1] = v.y;
        p[4 * q + 2] = v.z;
        p[4 * q + 3] = v.w;
    }
    out[b] = encode_block(p, quality, flags);
}

// Whole image, as dxt_image::init_task and set_block_pixels do for cColorDXT1
// elements (crnlib/crn_dxt_image.cpp:283-349, 1436-1484): the block's 4x4
// texels with edge clamp, the search, then the 8-byte dxt1_block (low colour,
// high colour, 2-bit selectors row by row, texel (x, y) at bit 2 (x + 4 y) of
// the last four bytes) at blocks[i * stride]; errors (optional) = m_error.
__global__ void __launch_bounds__(256, 2) dxt1_image_kernel(const uint32_t* __restrict__ rgba, uint32_t width,
                                                           uint32_t height, uint32_t blocks_x, uint32_t nblocks,
                                                           int quality, int flags, uint64_t* __restrict__ out,
                                                           uint32_t stride, uint64_t* __restrict__ err)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblocks) return;
    const uint32_t bx = b % blocks_x, by = b / blocks_x;
    uint32_t p[16];
#pragma unroll
    for (int y = 0; y < 4; y++) {
        const uint32_t sy = min(by * 4 + y, height - 1);
#pragma unroll
        for (int x = 0; x < 4; x++) p[4 * y + x] = rgba[(size_t)sy * width + min(bx * 4 + x, width - 1)];
    }
    const Dxt1Out o = encode_block(p, quality, flags);
    out[(size_t)b * stride] = (uint64_t)o.lo_hi | ((uint64_t)o.selectors << 32);
    if (err) err[b] = o.error;
}

// Texture-set encode: colour elements of many level images in one flat
// order, each thread walking it with a grid stride (a persistent grid when
// the launch is smaller than the work).
__global__ void __launch_bounds__(256, 2) dxt1_set_kernel(const crn2::SetJob* __restrict__ jobs, uint32_t njobs,
                                                         uint64_t begin, uint64_t end, int quality)
{
    const uint64_t step = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = begin + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < end; i += step) {
        const crn2::SetJob j = jobs[crn2::set_job_of(jobs, njobs, i)];
        const uint32_t b = (uint32_t)(i - j.first);
        const uint32_t bx = b % j.blocks_x, by = b / j.blocks_x;
        uint32_t p[16];
#pragma unroll
        for (int y = 0; y < 4; y++) {
            const uint32_t sy = min(by * 4 + y, j.height - 1);
#pragma unroll
            for (int x = 0; x < 4; x++) p[4 * y + x] = j.rgba[(size_t)sy * j.width + min(bx * 4 + x, j.width - 1)];
        }
        const Dxt1Out o = encode_block(p, quality, (int)j.flags);
        j.out[(size_t)b * j.stride] = (uint64_t)o.lo_hi | ((uint64_t)o.selectors << 32);
    }
}

}  // namespace

extern "C" int crn2_dxt1_init(void)
{
    // ryg single-colour tables (crn_ryg_dxt.cpp:58-99, 517-535), built on the host
    static uint8_t t[4][256][2];
    static bool ready = false;
    if (!ready) {
        for (int which = 0; which < 4; which++) {
            const int size = (which == 1 || which == 3) ? 64 : 32;
            const bool three = which >= 2;
            for (int i = 0; i < 256; i++) {
                int best = 256;
                for (int mn = 0; mn < size; mn++)
                    for (int mx = 0; mx < size; mx++) {
                        const int me = size == 32 ? ((mn << 3) | (mn >> 2)) : ((mn << 2) | (mn >> 4));
                        const int xe = size == 32 ? ((mx << 3) | (mx >> 2)) : ((mx << 2) | (mx >> 4));
                        int err = three ? abs(((me + xe) >> 1) - i) : abs(((xe * 2 + me) / 3) - i);
                        err += (abs(xe - me) * 8) >> 8;
                        if (err < best) {
                            t[which][i][0] = (uint8_t)mx;
                            t[which][i][1] = (uint8_t)mn;
                            best = err;
                        }
                    }
            }
        }
        ready = true;
    }
    if (hipMemcpyToSymbol(HIP_SYMBOL(g_om5), t[0], 512) != hipSuccess) return CRN2_EHIP;
    if (hipMemcpyToSymbol(HIP_SYMBOL(g_om6), t[1], 512) != hipSuccess) return CRN2_EHIP;
    if (hipMemcpyToSymbol(HIP_SYMBOL(g_om5_3), t[2], 512) != hipSuccess) return CRN2_EHIP;
    if (hipMemcpyToSymbol(HIP_SYMBOL(g_om6_3), t[3], 512) != hipSuccess) return CRN2_EHIP;
    return CRN2_OK;
}

namespace {
// crn2_dxt1_init once per device (the tables are device memory of each GPU)
std::mutex g_tables_mu;
bool g_tables[64];

int dxt1_tables_ready()
{
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return CRN2_EHIP;
    std::lock_guard<std::mutex> lk(g_tables_mu);
    if (!g_tables[dev]) {
        const int rc = crn2_dxt1_init();
        if (rc) return rc;
        g_tables[dev] = true;
    }
    return CRN2_OK;
}
}  // namespace

extern "C" int crn2_dxt1_blocks_device(const void* rgba, uint32_t nblocks, int quality, int flags, void* out,
                                       void* stream)
{
    if (!rgba || !out || quality < 0 || quality > 4 || (flags & ~0xFF7F)) return CRN2_EINVAL;
    if (reinterpret_cast<uintptr_t>(rgba) & 15) return CRN2_EINVAL;
    if (!nblocks) return CRN2_OK;
    const int rc = dxt1_tables_ready();
    if (rc) return rc;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // launch shapes (threads per workgroup, one block per lane; identical results): 64, 128, 256
    auto launch = [&](int shape) -> hipError_t {
        const uint32_t tb = 64u << shape, nb = (nblocks + tb - 1) / tb;
        hipLaunchKernelGGL(dxt1_blocks_kernel, dim3(nb), dim3(tb), 0, s, static_cast<const uint32_t*>(rgba), nblocks,
                           quality, flags, static_cast<Dxt1Out*>(out));
        return hipGetLastError();
    };
    const int shape = crn2::tuner().pick(crn2::TUNE_DXT1, nblocks, 3, 0, s, launch);
    return launch(shape) == hipSuccess ? CRN2_OK : CRN2_EHIP;
}

extern "C" int crn2_encode_dxt1_image_device(const void* rgba, uint32_t width, uint32_t height, int quality, int flags,
                                             void* blocks, uint32_t block_stride_u64, void* errors, void* stream)
{
    if (!rgba || !blocks || !width || !height || !block_stride_u64 || quality < 0 || quality > 4 ||
        (flags & ~0xFF7F))
        return CRN2_EINVAL;
    if (reinterpret_cast<uintptr_t>(rgba) & 3) return CRN2_EINVAL;
    const uint64_t bx = (width + 3) / 4, by = (height + 3) / 4, n = bx * by;
    if (n > 0xFFFFFFFFull) return CRN2_EINVAL;
    const int rc = dxt1_tables_ready();
    if (rc) return rc;
    const uint32_t tb = 256;
    hipLaunchKernelGGL(dxt1_image_kernel, dim3((uint32_t)((n + tb - 1) / tb)), dim3(tb), 0,
                       static_cast<hipStream_t>(stream), static_cast<const uint32_t*>(rgba), width, height,
                       (uint32_t)bx, (uint32_t)n, quality, flags, static_cast<uint64_t*>(blocks), block_stride_u64,
                       static_cast<uint64_t*>(errors));
    return hipGetLastError() == hipSuccess ? CRN2_OK : CRN2_EHIP;
}

extern "C" int crn2_dxt1_set_launch(const crn2::SetJob* jobs, uint32_t njobs, uint64_t begin, uint64_t end,
                                    int quality, uint32_t grid, uint32_t tb, hipStream_t stream)
{
    if (!jobs || !njobs || !grid || (tb != 64 && tb != 128 && tb != 256) || quality < 0 || quality > 4)
        return CRN2_EINVAL;
    if (begin >= end) return CRN2_OK;
    const int rc = dxt1_tables_ready();
    if (rc) return rc;
    hipLaunchKernelGGL(dxt1_set_kernel, dim3(grid), dim3(tb), 0, stream, jobs, njobs, begin, end, quality);
    return hipGetLastError() == hipSuccess ? CRN2_OK : CRN2_EHIP;
}

extern "C" int crn2_dxt1_set_prepare(void) { return dxt1_tables_ready(); }
