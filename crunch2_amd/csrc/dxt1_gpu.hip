// dxt1_gpu.hip -- DXT1 endpoint search (crnlib's dxt1_endpoint_optimizer,
// crnlib/crn_dxt1.cpp) for 16-texel blocks on gfx950, one wavefront per block.
//
// The search trajectory of crn_dxt1.cpp is reproduced step for step at every
// quality (superfast .. uber; probe tables and passes :716-746, the
// projection evaluator of the qualities below Better :1594-1757), with DXT1A
// transparency (:2116, :1380, :300-343) and the transparent-black trial of
// try_alpha_as_black_optimization (:2001-2079): unique colours in
// first-appearance order, perceptual weights (:1074-1108), PCA (:192-256),
// projection bounds and ray clamp (:1135-1175), k-means seeds with
// fast_random jitter (:1181-1308), probe sweeps along the axis, 26-neighbour
// and +-2 lattice moves, least-squares refinement (:525-698), solid-block
// tries (:93-153), per-channel exhaustive search (:369-486) and combinatorial
// recovery (:1886-1997).  The float/double types and the order of every float
// expression follow the reference; the file is compiled with -ffp-contract=off.
//
// Work split.  Colour k of the block lives on lane k.  The search's decisions
// are the same on all lanes; the sums over colours that round (float means,
// covariance, k-means centres) run over the colours in order, the exact
// integer ones in any order.  Candidate endpoint pairs are evaluated one per
// lane, 64 at a time, and taken in the reference's order: evaluate_solution
// (:1312-1340) keeps a candidate when the error planes' bound of the pair and
// then its error are below the best error at its turn, so a lane's candidate
// is taken when both are below the best error left by the candidates before
// it (the bound is not a strict lower bound of the error -- a texel on the
// lower endpoint's plane may lie above it -- so it is applied exactly, not
// only as a prune).  Lists that do not depend on the search state (sweeps,
// refinement boxes, pairs) are first cut to the candidates whose bound is
// below the best error when they are generated and gathered 64 at a time;
// the lattice moves, which read the current best endpoints, are generated
// again after each candidate taken.  The tried-solutions set (:1324-1330)
// only skips re-evaluations of candidates that can no longer be taken and is
// left out.
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

constexpr uint32_t kNone = 0xFFFFFFFFu;
// LDS words per wave: error planes (r lo/hi 32+32, g 64+64, b 32+32), the
// pending candidate list (128 packed coordinates) and their bounds (128)
constexpr uint32_t kPlaneWords = 256;
constexpr uint32_t kLdsWords = kPlaneWords + 128 + 128;

__device__ inline int lane_id() { return (int)__lane_id(); }
__device__ inline int rdl(int v, int k) { return __builtin_amdgcn_readlane(v, k); }
__device__ inline uint32_t rdlu(uint32_t v, int k) { return (uint32_t)__builtin_amdgcn_readlane((int)v, k); }
__device__ inline float rdlf(float v, int k) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k)); }
__device__ inline double rdld(double v, int k)
{
    const uint64_t u = __double_as_longlong(v);
    const uint32_t lo = rdlu((uint32_t)u, k), hi = rdlu((uint32_t)(u >> 32), k);
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}
__device__ inline uint64_t ballot(bool p) { return __ballot(p); }
__device__ inline uint64_t lanes_below() { return (1ull << lane_id()) - 1ull; }
__device__ inline uint32_t wave_min(uint32_t v)
{
    for (int o = 32; o; o >>= 1) v = min(v, (uint32_t)__shfl_xor((int)v, o));
    return v;
}
__device__ inline float wave_minf(float v)
{
    for (int o = 32; o; o >>= 1) v = fminf(v, __shfl_xor(v, o));
    return v;
}
__device__ inline float wave_maxf(float v)
{
    for (int o = 32; o; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
    return v;
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

__device__ inline int clampi(int v, int hi) { return v < 0 ? 0 : (v > hi ? hi : v); }

__device__ inline void canon(uint32_t& lo, uint32_t& hi)
{
    if (lo < hi) {
        const uint32_t t = lo;
        lo = hi;
        hi = t;
    }
}

// color_distance (crn_color.h:728-745) and the grayscale variant of
// dxt1_endpoint_optimizer::color_distance (crn_dxt1.cpp:1342-1368):
// texel (ar, ag, ab) against palette entry (br, bg, bb)
__device__ inline uint32_t dist(int perceptual, int grayscale, int ar, int ag, int ab, int br, int bg, int bb)
{
    const int dr = ar - br, dg = ag - bg, db = ab - bb;
    if (perceptual) return (uint32_t)(dr * dr * 8 + dg * dg * 25 + db * db);
    if (grayscale) {
        const int y0 = (ar * 19595 + ag * 38470 + ab * 7471 + 32768) >> 16;
        const int y1 = (br * 19595 + bg * 38470 + bb * 7471 + 32768) >> 16;
        return (uint32_t)((y0 - y1) * (y0 - y1));
    }
    return (uint32_t)(dr * dr + dg * dg + db * db);
}

struct Best {
    uint32_t lo, hi, err;
    int alpha, alt, enforce, enforced;
    bool sel_valid;     // the selector vector holds values (m_selectors non-empty)
};

// One block, held by one wavefront.  Parameters, the best solution and the
// search's float state are the same on every lane; colour k of the block's
// unique colours (first-appearance order) lives on lane k, as does its
// selector in the best solution.  Candidate endpoints are evaluated one per
// lane; their acceptance follows the reference's sequential order.
struct Wb {
    int quality, perceptual, use_alpha, grayscale, force_alpha, have_alpha, hc;
    uint32_t nu, total_w;
    int cr, cg, cb;          // lane k: colour k (0 past nu)
    uint32_t cw;             // lane k: its weight (0 past nu)
    uint32_t cpk;            // lane k: r | g << 8 | b << 16 | weight << 24 (0 past nu)
    float nr, ng, nb;        // lane k: normalised colour minus the mean
    float mean[3], axis[3];
    Best best;
    uint32_t bsel;           // lane k: selector of colour k in the best solution
    uint32_t* planes;        // LDS
    uint32_t* pend;          // LDS: pending candidates (lo | hi << 16)
    uint32_t* pbound;        // LDS: their plane bounds
};

__device__ __forceinline__ int colour_r(const Wb& s, int k) { return rdl(s.cr, k); }
__device__ __forceinline__ int colour_g(const Wb& s, int k) { return rdl(s.cg, k); }
__device__ __forceinline__ int colour_b(const Wb& s, int k) { return rdl(s.cb, k); }
__device__ __forceinline__ uint32_t colour_w(const Wb& s, int k) { return rdlu(s.cw, k); }

// Error planes of compute_internal (crn_dxt1.cpp:2134-2203): for every 5/6-bit
// value c of a channel, lo[c] = the weighted error of the colours whose
// nearest plane is <= c against c's expanded value, hi[c] = of the others
// (x8 red, x25 green under the perceptual metric).  Every entry is below
// 16 * 255^2 * 25 < 2^32.
__device__ __forceinline__ void build_planes(Wb& s)
{
    for (int e = lane_id(); e < 128; e += 64) {
        const int ch = e < 32 ? 0 : (e < 96 ? 1 : 2);
        const uint32_t c = (uint32_t)(ch == 0 ? e : (ch == 1 ? e - 32 : e - 96));
        const uint32_t C = ch == 1 ? ((c << 2) | (c >> 4)) : ((c << 3) | (c >> 2));
        const uint32_t scale = s.perceptual ? (ch == 0 ? 8u : (ch == 1 ? 25u : 1u)) : 1u;
        uint32_t lo = 0, hi = 0;
        for (int k = 0; k < (int)s.nu; k++) {
            const uint32_t V = (uint32_t)(ch == 0 ? colour_r(s, k) : (ch == 1 ? colour_g(s, k) : colour_b(s, k)));
            const uint32_t p = ch == 1 ? (V >> 2) + (((V & 3) > (V >> 6)) ? 1u : 0u)
                                       : (V >> 3) + (((V & 7) > (V >> 5)) ? 1u : 0u);
            const int d = (int)V - (int)C;
            const uint32_t t = colour_w(s, k) * (uint32_t)(d * d) * scale;
            if (p <= c) lo += t;
            else hi += t;
        }
        const uint32_t base = ch == 0 ? 0u : (ch == 1 ? 64u : 192u), n = ch == 1 ? 64u : 32u;
        s.planes[base + c] = lo;
        s.planes[base + n + c] = hi;
    }
    __builtin_amdgcn_wave_barrier();
}

// the planes' bound of evaluate_solution (crn_dxt1.cpp:1316-1322)
__device__ __forceinline__ uint32_t plane_bound(const Wb& s, uint32_t lo, uint32_t hi)
{
    int ar, ag, ab, br, bg, bb;
    unpack565(lo, 0, ar, ag, ab);
    unpack565(hi, 0, br, bg, bb);
    const uint32_t* P = s.planes;
    const uint32_t re = ar < br ? P[ar] + P[32 + br] : P[32 + ar] + P[br];
    const uint32_t ge = ag < bg ? P[64 + ag] + P[128 + bg] : P[128 + ag] + P[64 + bg];
    const uint32_t be = ab < bb ? P[192 + ab] + P[224 + bb] : P[224 + ab] + P[192 + bb];
    return re + ge + be;
}

// lanes per candidate for a batch of m candidates: the colours of a
// candidate are split over its lanes when the batch leaves lanes free
__device__ inline uint32_t lanes_per(uint32_t m)
{
    return m <= 4 ? 16u : (m <= 8 ? 8u : (m <= 16 ? 4u : (m <= 32 ? 2u : 1u)));
}

// Error of candidate (lo, hi, alt) under the block's evaluator, and the block
// type it would take: evaluate_solution_hc_* (crn_dxt1.cpp:1759-1835), _uber
// (:1370-1592: 4-colour, then a 3-colour block only if strictly better) or
// _fast (:1594-1757, below Better).  The candidate is held by g aligned lanes
// (g a power of two <= 16), each summing the colours k = its index mod g; the
// integer sums are then added over the g lanes.  Each evaluator's early exits
// only decide whether the candidate is kept, which the caller decides from
// the full error.
__device__ __forceinline__ uint32_t cand_err(const Wb& s, uint32_t lo, uint32_t hi, int alt, int& type, uint32_t g)
{
    int c0r, c0g, c0b, c1r, c1g, c1b;
    unpack565(lo, 1, c0r, c0g, c0b);
    unpack565(hi, 1, c1r, c1g, c1b);
    const int c2r = (c0r * 2 + c1r + alt) / 3, c2g = (c0g * 2 + c1g + alt) / 3, c2b = (c0b * 2 + c1b + alt) / 3;
    const int c3r = (c1r * 2 + c0r + alt) / 3, c3g = (c1g * 2 + c0g + alt) / 3, c3b = (c1b * 2 + c0b + alt) / 3;
    const int h2r = (c0r + c1r + alt) >> 1, h2g = (c0g + c1g + alt) >> 1, h2b = (c0b + c1b + alt) >> 1;
    const int P = s.perceptual, G = s.grayscale;
    const uint32_t sub = (uint32_t)lane_id() & (g - 1), kend = ((s.nu + g - 1) / g) * g;
    type = 0;
    int first = 0, last = 1;
    if (s.hc) {
        first = last = 0;
    } else if (s.have_alpha || s.force_alpha) {
        first = 1;
    } else if (!s.use_alpha) {
        last = 0;
    }
    uint32_t e4 = 0, e3 = 0;
    // colour k: from a scalar register when every lane reads the same one
    // (g == 1), else from lane k
    auto colour = [&](uint32_t k) {
        return g == 1 ? rdlu(s.cpk, (int)k) : (uint32_t)__shfl((int)s.cpk, (int)k);
    };
    if (s.hc || s.quality >= 3) {
        auto body = [&](uint32_t pk) {
            const int r = (int)(pk & 255u), gg = (int)((pk >> 8) & 255u), b = (int)((pk >> 16) & 255u);
            const uint32_t w = pk >> 24;
            const uint32_t d0 = dist(P, G, r, gg, b, c0r, c0g, c0b), d1 = dist(P, G, r, gg, b, c1r, c1g, c1b);
            if (first == 0) {
                const uint32_t d2 = dist(P, G, r, gg, b, c2r, c2g, c2b), d3 = dist(P, G, r, gg, b, c3r, c3g, c3b);
                e4 += min(min(d0, d1), min(d2, d3)) * w;
            }
            if (last == 1) e3 += min(min(d0, d1), dist(P, G, r, gg, b, h2r, h2g, h2b)) * w;
        };
        if (g == 1) {
            for (uint32_t k = 0; k < s.nu; k++) body(rdlu(s.cpk, (int)k));
        } else {
            for (uint32_t k = sub; k < kend; k += g) body(colour(k));
        }
    } else {
        int vr = c1r - c0r, vg = c1g - c0g, vb = c1b - c0b;
        if (P) {
            vr *= 8;
            vg *= 24;
        }
        const int s0 = c0r * vr + c0g * vg + c0b * vb, s1 = c1r * vr + c1g * vg + c1b * vb;
        const int s2 = c2r * vr + c2g * vg + c2b * vb, s3 = c3r * vr + c3g * vg + c3b * vb;
        const int h2 = h2r * vr + h2g * vg + h2b * vb;
        const int dr = vr * 2, dg = vg * 2, db = vb * 2;
        for (uint32_t k = sub; k < kend; k += g) {
            const uint32_t pk = colour(k);
            const int r = (int)(pk & 255u), gg = (int)((pk >> 8) & 255u), b = (int)((pk >> 16) & 255u);
            const uint32_t w = pk >> 24;
            const int dot = r * dr + gg * dg + b * db;
            if (first == 0) {
                int pr, pg, pb;
                if (dot >= s3 + s2) {
                    const bool three = dot < s1 + s3;
                    pr = three ? c3r : c1r; pg = three ? c3g : c1g; pb = three ? c3b : c1b;
                } else {
                    const bool zero = dot < s2 + s0;
                    pr = zero ? c0r : c2r; pg = zero ? c0g : c2g; pb = zero ? c0b : c2b;
                }
                e4 += dist(P, G, r, gg, b, pr, pg, pb) * w;
            }
            if (last == 1) {
                int pr, pg, pb;
                if (dot < s0 + h2) { pr = c0r; pg = c0g; pb = c0b; }
                else if (dot < h2 + s1) { pr = h2r; pg = h2g; pb = h2b; }
                else { pr = c1r; pg = c1g; pb = c1b; }
                e3 += dist(P, G, r, gg, b, pr, pg, pb) * w;
            }
        }
    }
    for (uint32_t o = 1; o < g; o <<= 1) {
        e4 += (uint32_t)__shfl_xor((int)e4, (int)o);
        e3 += (uint32_t)__shfl_xor((int)e3, (int)o);
    }
    if (first == 1) {
        type = 1;
        return e3;
    }
    if (last == 1 && e3 < e4) {
        type = 1;
        return e3;
    }
    return e4;
}

// Lane k's selector of colour k under (lo, hi, alt, type) by the evaluator's
// rule: the first minimum of the palette (uber), the projection thresholds
// (fast).
__device__ __forceinline__ uint32_t colour_selector(const Wb& s, uint32_t lo, uint32_t hi, int alt, int type)
{
    int c0r, c0g, c0b, c1r, c1g, c1b;
    unpack565(lo, 1, c0r, c0g, c0b);
    unpack565(hi, 1, c1r, c1g, c1b);
    int pr[4], pg[4], pb[4];
    pr[0] = c0r; pg[0] = c0g; pb[0] = c0b;
    pr[1] = c1r; pg[1] = c1g; pb[1] = c1b;
    if (!type) {
        pr[2] = (c0r * 2 + c1r + alt) / 3; pg[2] = (c0g * 2 + c1g + alt) / 3; pb[2] = (c0b * 2 + c1b + alt) / 3;
        pr[3] = (c1r * 2 + c0r + alt) / 3; pg[3] = (c1g * 2 + c0g + alt) / 3; pb[3] = (c1b * 2 + c0b + alt) / 3;
    } else {
        pr[2] = (c0r + c1r + alt) >> 1; pg[2] = (c0g + c1g + alt) >> 1; pb[2] = (c0b + c1b + alt) >> 1;
        pr[3] = pg[3] = pb[3] = 0;
    }
    const int ncol = type ? 3 : 4;
    if (s.quality >= 3) {
        uint32_t be = dist(s.perceptual, s.grayscale, s.cr, s.cg, s.cb, pr[0], pg[0], pb[0]), bi = 0;
        for (int j = 1; j < 4; j++) {
            if (j >= ncol) break;
            const uint32_t e = dist(s.perceptual, s.grayscale, s.cr, s.cg, s.cb, pr[j], pg[j], pb[j]);
            if (e < be) {
                be = e;
                bi = (uint32_t)j;
            }
        }
        return bi;
    }
    int vr = c1r - c0r, vg = c1g - c0g, vb = c1b - c0b;
    if (s.perceptual) {
        vr *= 8;
        vg *= 24;
    }
    int st[4];
    for (int j = 0; j < 4; j++) st[j] = pr[j] * vr + pg[j] * vg + pb[j] * vb;
    const int dot = s.cr * vr * 2 + s.cg * vg * 2 + s.cb * vb * 2;
    if (!type) return dot >= st[3] + st[2] ? (dot < st[1] + st[3] ? 3u : 1u) : (dot < st[2] + st[0] ? 0u : 2u);
    return dot < st[0] + st[2] ? 0u : (dot < st[2] + st[1] ? 2u : 1u);
}

// The winner of an evaluate_solution call becomes the best solution.
__device__ __forceinline__ void accept(Wb& s, uint32_t lo, uint32_t hi, int alt, int type, uint32_t err)
{
    Best& b = s.best;
    b.lo = lo;
    b.hi = hi;
    b.err = err;
    if (s.hc) {
        // evaluate_solution_hc_*: selectors are computed later (compute_selectors)
        b.alpha = 0;
        b.alt = alt;
        b.enforce = lo == hi;
        if (b.enforce) {
            if ((b.lo & 31) != 31) { b.lo++; b.enforced = 1; }
            else { b.hi--; b.enforced = 0; }
        }
        return;
    }
    b.alpha = type;
    b.alt = s.quality >= 3 ? alt : 0;
    s.bsel = colour_selector(s, lo, hi, b.alt, type);
    b.sel_valid = true;
    b.enforce = !b.alpha && lo == hi;
    if (b.enforce) {
        int v;
        if ((b.lo & 31) != 31) { b.lo++; v = 1; }
        else { b.hi--; v = 0; }
        s.bsel = (uint32_t)v;
        b.enforced = v;
    }
}

// The reference's sequential evaluate_solution calls over the candidates on
// the valid lanes, in lane order: a candidate is taken when its plane bound
// and its error are both below the current best error.  stop_first returns
// after the first one taken (its lane, or -1).
__device__ __forceinline__ int take(Wb& s, bool valid, uint32_t lo, uint32_t hi, int alt, uint32_t bnd, uint32_t err, int type,
                    bool stop_first)
{
    int last = -1;
    uint64_t m = ballot(valid && bnd < s.best.err && err < s.best.err);
    while (m) {
        const int l = __ffsll((unsigned long long)m) - 1;
        accept(s, rdlu(lo, l), rdlu(hi, l), rdl(alt, l), rdl(type, l), rdlu(err, l));
        last = l;
        if (stop_first) break;
        m = ballot(valid && lane_id() > l && bnd < s.best.err && err < s.best.err);
    }
    return last;
}

// Evaluates the pending candidates (the first cnt entries) in order.
__device__ __forceinline__ bool flush(Wb& s, int alt, uint32_t cnt)
{
    const uint32_t g = lanes_per(cnt), c = (uint32_t)lane_id() / g;
    const bool v = c < cnt;
    const uint32_t pk = v ? s.pend[c] : 0u, bnd = v ? s.pbound[c] : kNone;
    const uint32_t lo = pk & 0xFFFFu, hi = pk >> 16;
    int type = 0;
    const uint32_t e = cand_err(s, lo, hi, alt, type, g);
    const bool rep = v && ((uint32_t)lane_id() & (g - 1)) == 0;
    return take(s, rep, lo, hi, alt, bnd, rep ? e : kNone, type, false) >= 0;
}

// n candidates that do not depend on the search state, candidate i =
// gen(i, lo, hi) (false: skipped), evaluated in order with rounding flag alt.
// gen runs on every lane (lanes past n with index 0), so that the shuffles in
// it read active lanes.
// Candidates whose bound already reaches the best error are dropped before
// the full evaluation (the bound is checked again, in order, when taken);
// the rest are gathered 64 at a time.
template <class Gen>
__device__ __forceinline__ bool eval_list(Wb& s, uint32_t n, int alt, Gen gen)
{
    bool any = false;
    uint32_t cnt = 0;
    for (uint32_t base = 0; base < n; base += 64) {
        const uint32_t i = base + (uint32_t)lane_id();
        uint32_t lo = 0, hi = 0;
        const bool made = gen(i < n ? i : 0u, lo, hi);
        const bool ok = i < n && made;
        const uint32_t bnd = ok ? plane_bound(s, lo, hi) : kNone;
        const bool alive = ok && bnd < s.best.err;
        const uint64_t am = ballot(alive);
        if (!am) continue;
        const uint32_t pos = cnt + (uint32_t)__popcll(am & lanes_below());
        if (alive) {
            s.pend[pos] = lo | (hi << 16);
            s.pbound[pos] = bnd;
        }
        cnt += (uint32_t)__popcll(am);
        __builtin_amdgcn_wave_barrier();
        if (cnt >= 64) {
            any |= flush(s, alt, 64);
            const uint32_t rest = cnt - 64;
            uint32_t a = 0, b = 0;
            if ((uint32_t)lane_id() < rest) {
                a = s.pend[64 + lane_id()];
                b = s.pbound[64 + lane_id()];
            }
            __builtin_amdgcn_wave_barrier();
            if ((uint32_t)lane_id() < rest) {
                s.pend[lane_id()] = a;
                s.pbound[lane_id()] = b;
            }
            __builtin_amdgcn_wave_barrier();
            cnt = rest;
        }
    }
    if (cnt) any |= flush(s, alt, cnt);
    return any;
}

// n <= 64 candidates that depend on the best solution (lattice moves):
// candidate i = gen(i, lo, hi) is generated from the state at its turn, so
// after a candidate is taken the ones after it are generated again.
template <class Gen>
__device__ __forceinline__ bool eval_chain(Wb& s, uint32_t n, Gen gen)
{
    bool any = false;
    uint32_t start = 0;
    while (start < n) {
        const uint32_t g = lanes_per(n - start), i = start + (uint32_t)lane_id() / g;
        uint32_t lo = 0, hi = 0;
        const bool made = gen(i < n ? i : 0u, lo, hi);
        const bool ok = i < n && made;
        const uint32_t bnd = ok ? plane_bound(s, lo, hi) : kNone;
        const bool rep = ok && ((uint32_t)lane_id() & (g - 1)) == 0;
        if (!ballot(rep && bnd < s.best.err)) break;
        int type = 0;
        const uint32_t e = cand_err(s, lo, hi, 0, type, g);
        const int l = take(s, rep, lo, hi, 0, bnd, rep ? e : kNone, type, true);
        if (l < 0) break;
        any = true;
        start += (uint32_t)l / g + 1;
    }
    return any;
}

// one candidate (the reference's single evaluate_solution call)
__device__ __forceinline__ bool eval_one(Wb& s, uint32_t lo, uint32_t hi, int alt)
{
    return eval_list(s, 1, alt, [&](uint32_t, uint32_t& a, uint32_t& b) {
        a = lo;
        b = hi;
        return true;
    });
}

// compute_selectors (crn_dxt1.cpp:1837-1869): the hc evaluator's selectors
// of the best solution
__device__ __forceinline__ void compute_selectors(Wb& s)
{
    if (!s.hc) return;
    Best& b = s.best;
    b.sel_valid = true;
    if (b.enforce) {
        s.bsel = (uint32_t)b.enforced;
        return;
    }
    int c0r, c0g, c0b, c1r, c1g, c1b;
    unpack565(b.lo, 1, c0r, c0g, c0b);
    unpack565(b.hi, 1, c1r, c1g, c1b);
    const int a = b.alt;
    const int c2r = (c0r * 2 + c1r + a) / 3, c2g = (c0g * 2 + c1g + a) / 3, c2b = (c0b * 2 + c1b + a) / 3;
    const int c3r = (c1r * 2 + c0r + a) / 3, c3g = (c1g * 2 + c0g + a) / 3, c3b = (c1b * 2 + c0b + a) / 3;
    const int P = s.perceptual, G = s.grayscale;
    const uint32_t e0 = dist(P, G, s.cr, s.cg, s.cb, c0r, c0g, c0b), e1 = dist(P, G, s.cr, s.cg, s.cb, c1r, c1g, c1b);
    const uint32_t e2 = dist(P, G, s.cr, s.cg, s.cb, c2r, c2g, c2b), e3 = dist(P, G, s.cr, s.cg, s.cb, c3r, c3g, c3b);
    const uint32_t e01 = min(e0, e1), e23 = min(e2, e3);
    s.bsel = e01 <= e23 ? (e01 == e0 ? 0u : 1u) : (e23 == e2 ? 2u : 3u);
}

__device__ __forceinline__ bool sel_all_equal(const Wb& s)
{
    if (!s.best.sel_valid) return false;
    const uint32_t f = rdlu(s.bsel, 0);
    return !ballot((uint32_t)lane_id() < s.nu && s.bsel != f);
}

// refine_solution (crn_dxt1.cpp:525-698), levels 0 and 1.  Every sum is of
// integers below 2^53, exact as the reference's doubles.
__device__ __forceinline__ bool refine(Wb& s, int level)
{
    compute_selectors(s);
    int64_t a0 = 0, a1 = 0, a2 = 0, t1r = 0, t1g = 0, t1b = 0, t2r = 0, t2g = 0, t2b = 0;
    for (int k = 0; k < (int)s.nu; k++) {
        const int64_t w = colour_w(s, k);
        const int64_t r = colour_r(s, k) * w, g = colour_g(s, k) * w, b = colour_b(s, k) * w;
        const int step = (int)rdlu(s.bsel, k) ^ 1;
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
    const double T1r = (double)t1r, T1g = (double)t1g, T1b = (double)t1b;
    const double T2r = 3 * (double)t2r - T1r, T2g = 3 * (double)t2g - T1g, T2b = 3 * (double)t2b - T1b;
    const double xx = (double)a2, yy = (double)a1, xy = (double)a0;
    const double t = xx * yy - xy * xy;
    if (!yy || !xx || fabs(t) < (double).0000125f) return false;
    const double frb = (double)(3.0f * 31.0f / 255.0f) / t;
    const double fg = frb * (double)(63.0f / 31.0f);
    int e0[3], e1[3];
    e0[0] = clampi((int)((T1r * yy - T2r * xy) * frb + (double)0.5f), 31);
    e0[1] = clampi((int)((T1g * yy - T2g * xy) * fg + (double)0.5f), 63);
    e0[2] = clampi((int)((T1b * yy - T2b * xy) * frb + (double)0.5f), 31);
    e1[0] = clampi((int)((T2r * xx - T1r * xy) * frb + (double)0.5f), 31);
    e1[1] = clampi((int)((T2g * xx - T1g * xy) * fg + (double)0.5f), 63);
    e1[2] = clampi((int)((T2b * xx - T1b * xy) * frb + (double)0.5f), 31);
    if (level == 0) {
        uint32_t lo = (uint32_t)((e1[0] << 11) | (e1[1] << 5) | e1[2]);
        uint32_t hi = (uint32_t)((e0[0] << 11) | (e0[1] << 5) | e0[2]);
        canon(lo, hi);
        return eval_one(s, lo, hi, 0);
    }
    // 2 x 27 candidates: one endpoint moved by -1..+1 per channel (i, r, g, b order)
    return eval_list(s, 54, 0, [&](uint32_t idx, uint32_t& lo, uint32_t& hi) {
        const int i = (int)idx / 27, rr = ((int)idx % 27) / 9 - 1, gr = ((int)idx % 9) / 3 - 1, br = (int)idx % 3 - 1;
        const int a0 = i ? 0 : rr, a1 = i ? 0 : gr, a2 = i ? 0 : br;
        const int b0 = i ? rr : 0, b1 = i ? gr : 0, b2 = i ? br : 0;
        lo = pack565(clampi(e0[0] + a0, 31), clampi(e0[1] + a1, 63), clampi(e0[2] + a2, 31), 0);
        hi = pack565(clampi(e1[0] + b0, 31), clampi(e1[1] + b1, 63), clampi(e1[2] + b2, 31), 0);
        canon(lo, hi);
        return true;
    });
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

// try_median4 (crn_dxt1.cpp:1181-1308).  Each colour's nearest centre is
// found on its lane; the float sums run over the colours in order.
__device__ __forceinline__ bool try_median4(Wb& s, const float* lo, const float* hi)
{
    float m[4][3];
    if (s.nu <= 4) {
        for (int i = 0; i < 4; i++) {
            const int k = min(i, (int)s.nu - 1);
            m[i][0] = rdlf(s.nr, k);
            m[i][1] = rdlf(s.ng, k);
            m[i][2] = rdlf(s.nb, k);
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
        const float wf = (float)s.cw;
        const float pr = s.nr * wf, pg = s.ng * wf, pb = s.nb * wf;
        for (uint32_t it = 0; it < 8; it++) {
            // this lane's colour: nearest centre (first minimum) and its distance
            float bd = 0;
            {
                float d = m[0][0] - s.nr; bd += d * d;
                d = m[0][1] - s.ng; bd += d * d;
                d = m[0][2] - s.nb; bd += d * d;
            }
            int bi = 0;
            for (int j = 1; j < 4; j++) {
                float dd = 0;
                float d = m[j][0] - s.nr; dd += d * d;
                d = m[j][1] - s.ng; dd += d * d;
                d = m[j][2] - s.nb; dd += d * d;
                if (dd < bd) { bd = dd; bi = j; }
            }
            float nm[4][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
            float nw[4] = {0, 0, 0, 0};
            float total = 0;
            for (int k = 0; k < (int)s.nu; k++) {
                total += rdlf(bd, k);
                const int j = rdl(bi, k);
                const float a = rdlf(pr, k), b = rdlf(pg, k), c = rdlf(pb, k), w = rdlf(wf, k);
                for (int q = 0; q < 4; q++)
                    if (q == j) {
                        nm[q][0] += a;
                        nm[q][1] += b;
                        nm[q][2] += c;
                        nw[q] += w;
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
    uint32_t cl[6], ch[6];
    int q = 0;
    for (int i = 0; i < 3; i++)
        for (int j = i + 1; j < 4; j++, q++) {
            float a[3], b[3];
            for (int c = 0; c < 3; c++) {
                a[c] = m[i][c] + s.mean[c];
                b[c] = m[j][c] + s.mean[c];
            }
            cl[q] = pack565((int)floorf(.5f + a[0] * 31.0f), (int)floorf(.5f + a[1] * 63.0f),
                            (int)floorf(.5f + a[2] * 31.0f), 0);
            ch[q] = pack565((int)floorf(.5f + b[0] * 31.0f), (int)floorf(.5f + b[1] * 63.0f),
                            (int)floorf(.5f + b[2] * 31.0f), 0);
        }
    bool improved = eval_list(s, 6, 0, [&](uint32_t i, uint32_t& lo2, uint32_t& hi2) {
        lo2 = i == 0 ? cl[0] : i == 1 ? cl[1] : i == 2 ? cl[2] : i == 3 ? cl[3] : i == 4 ? cl[4] : cl[5];
        hi2 = i == 0 ? ch[0] : i == 1 ? ch[1] : i == 2 ? ch[2] : i == 3 ? ch[3] : i == 4 ? ch[4] : ch[5];
        canon(lo2, hi2);
        return true;
    });
    improved |= refine(s, s.quality == 4 ? 1 : 0);
    return improved;
}

// try_average_block_as_solid (crn_dxt1.cpp:93-153).  A candidate after the
// error reaches 0 is never taken, so the reference's "&& error" conditions
// need no test.
__device__ __forceinline__ bool try_average_solid(Wb& s)
{
    uint32_t tr = 0, tg = 0, tb = 0, tw = 0;
    for (int k = 0; k < (int)s.nu; k++) {
        const uint32_t w = colour_w(s, k);
        tw += w;
        tr += (uint32_t)colour_r(s, k) * w;
        tg += (uint32_t)colour_g(s, k) * w;
        tb += (uint32_t)colour_b(s, k) * w;
    }
    const uint32_t half = tw >> 1;
    const uint32_t ar = (tr + half) / tw, ag = (tg + half) / tw, ab = (tb + half) / tw;
    const uint32_t per = s.use_alpha ? 2u : 1u;
    const uint32_t slots = s.quality == 4 ? 1u + s.nu : 1u;
    return eval_list(s, slots * per, 0, [&](uint32_t idx, uint32_t& lo, uint32_t& hi) {
        const uint32_t slot = idx / per, three = idx % per;
        const int from = slot ? (int)slot - 1 : 0;
        const uint32_t kr = (uint32_t)__shfl(s.cr, from), kg = (uint32_t)__shfl(s.cg, from),
                       kb = (uint32_t)__shfl(s.cb, from);
        uint32_t r = ar, g = ag, b = ab;
        if (slot) {
            if (kr == ar && kg == ag && kb == ab) return false;
            r = kr;
            g = kg;
            b = kb;
        }
        if (three) {
            lo = ((uint32_t)g_om5_3[r][0] << 11) | ((uint32_t)g_om6_3[g][0] << 5) | g_om5_3[b][0];
            hi = ((uint32_t)g_om5_3[r][1] << 11) | ((uint32_t)g_om6_3[g][1] << 5) | g_om5_3[b][1];
        } else {
            lo = ((uint32_t)g_om5[r][0] << 11) | ((uint32_t)g_om6[g][0] << 5) | g_om5[b][0];
            hi = ((uint32_t)g_om5[r][1] << 11) | ((uint32_t)g_om6[g][1] << 5) | g_om5[b][1];
        }
        return true;
    });
}

// compute_endpoint_component_errors + optimize_endpoint_comps
// (crn_dxt1.cpp:369-486).  Per selector class the channel error of value p
// is W p^2 - WP2 p + WPP, below 2^32 for 16 texels (32-bit arithmetic that
// wraps in the intermediate steps, like the reference's 64-bit one, is exact).
struct QErr {
    uint32_t W[4], WP2[4], WPP[4];
    __device__ inline uint32_t at(int sl, uint32_t p) const { return W[sl] * p * p - WP2[sl] * p + WPP[sl]; }
};

__device__ inline uint32_t expand_comp(int comp, uint32_t c)
{
    return comp == 1 ? (((c << 2) | (c >> 4)) & 0xFF) : (((c << 3) | (c >> 2)) & 0xFF);
}

__device__ __forceinline__ void comp_errors(const Wb& s, int comp, QErr& q, uint32_t* brem)
{
    for (int k = 0; k < 4; k++) q.W[k] = q.WP2[k] = q.WPP[k] = 0;
    for (int k = 0; k < (int)s.nu; k++) {
        const uint32_t p = (uint32_t)(comp == 0 ? colour_r(s, k) : comp == 1 ? colour_g(s, k) : colour_b(s, k));
        const uint32_t w = colour_w(s, k);
        const uint32_t sl = rdlu(s.bsel, k);
        for (uint32_t c = 0; c < 4; c++)
            if (c == sl) {
                q.W[c] += w;
                q.WP2[c] += w * p * 2;
                q.WPP[c] += w * p * p;
            }
    }
    const uint32_t lim = comp == 1 ? 64 : 32;
    const uint32_t l = (uint32_t)lane_id();
    for (int sl = 0; sl < 2; sl++) brem[sl] = wave_min(l < lim ? q.at(sl, expand_comp(comp, l)) : kNone);
    for (int sl = 2; sl < 4; sl++) {
        uint32_t v = q.at(sl, l);
        for (uint32_t p = l + 64; p < 256; p += 64) v = min(v, q.at(sl, p));
        brem[sl] = wave_min(v);
    }
    for (int sl = 3; sl; sl--) brem[sl - 1] += brem[sl];
}

// The candidates of one channel are the pairs (c0, c1) in order, each
// evaluated when the error-table bounds pass with the state at its turn
// (:437-470); after a candidate is taken the tables and bounds change, so the
// pairs after it are filtered again.  Passing pairs are gathered 64 at a time.
__device__ __forceinline__ void optimize_comps(Wb& s)
{
    compute_selectors(s);
    if (s.best.alpha || !s.best.err) return;
    int slr, slg, slb, shr, shg, shb;
    unpack565(s.best.lo, 1, slr, slg, slb);
    unpack565(s.best.hi, 1, shr, shg, shb);
    uint32_t brem[4];
    QErr q;
    for (int comp = 0; comp < 3; comp++) {
        const uint32_t p0 = (uint32_t)(comp == 0 ? slr : comp == 1 ? slg : slb);
        const uint32_t p1 = (uint32_t)(comp == 0 ? shr : comp == 1 ? shg : shb);
        int lr, lg, lb, hr, hg, hb;
        unpack565(s.best.lo, 0, lr, lg, lb);
        unpack565(s.best.hi, 0, hr, hg, hb);
        comp_errors(s, comp, q, brem);
        uint32_t best = q.at(0, p0) + q.at(1, p1) + q.at(2, (p0 * 2 + p1) / 3) + q.at(3, (p0 + p1 * 2) / 3);
        if (brem[0] >= best) continue;
        const uint32_t lim = comp == 1 ? 64 : 32, shift = comp == 1 ? 6 : 5, total = lim * lim;
        uint32_t pos = 0;      // first pair not yet decided
        uint32_t scan = 0;     // first pair not yet filtered
        uint32_t cnt = 0;      // pending pairs (indices) in s.pend
        while (pos < total) {
            // filter pairs from `scan` until 64 pass or the pairs run out
            while (cnt < 64 && scan < total) {
                const uint32_t idx = scan + (uint32_t)lane_id();
                const uint32_t c0 = idx >> shift, c1 = idx & (lim - 1);
                const uint32_t q0 = expand_comp(comp, c0), q1 = expand_comp(comp, c1);
                const uint32_t e0 = q.at(0, q0);
                uint32_t e = e0 + q.at(1, q1);
                bool pass = idx < total && !(e0 + brem[1] >= best) && !(e + brem[2] >= best);
                e += q.at(2, (q0 * 2 + q1) / 3);
                pass = pass && !(e + brem[3] >= best);
                e += q.at(3, (q0 + q1 * 2) / 3);
                pass = pass && !(e >= best);
                const uint64_t pm = ballot(pass);
                if (pass) s.pend[cnt + (uint32_t)__popcll(pm & lanes_below())] = idx;
                cnt += (uint32_t)__popcll(pm);
                scan += 64;
                __builtin_amdgcn_wave_barrier();
            }
            if (!cnt) break;
            const uint32_t take_n = min(cnt, 64u), g = lanes_per(take_n), cl = (uint32_t)lane_id() / g;
            const bool v = cl < take_n;
            const uint32_t idx = v ? s.pend[cl] : 0u;
            const uint32_t c0 = idx >> shift, c1 = idx & (lim - 1);
            const uint32_t plow = pack565(comp == 0 ? (int)c0 : lr, comp == 1 ? (int)c0 : lg, comp == 2 ? (int)c0 : lb, 0);
            const uint32_t phigh = pack565(comp == 0 ? (int)c1 : hr, comp == 1 ? (int)c1 : hg, comp == 2 ? (int)c1 : hb, 0);
            const uint32_t bnd = v ? plane_bound(s, plow, phigh) : kNone;
            int type = 0;
            const uint32_t ce = cand_err(s, plow, phigh, 0, type, g);
            const bool rep = v && ((uint32_t)lane_id() & (g - 1)) == 0;
            const int l = take(s, rep, plow, phigh, 0, bnd, rep ? ce : kNone, type, true);
            if (l < 0) {
                // none of these was taken: keep the rest of the pending pairs
                const uint32_t rest = cnt - take_n;
                const uint32_t moved = (uint32_t)lane_id() < rest ? s.pend[64 + lane_id()] : 0u;
                __builtin_amdgcn_wave_barrier();
                if ((uint32_t)lane_id() < rest) s.pend[lane_id()] = moved;
                __builtin_amdgcn_wave_barrier();
                cnt = rest;
                pos = rest ? rdlu(idx, (int)((take_n - 1) * g)) + 1 : scan;
                continue;
            }
            if (!s.best.err) return;
            compute_selectors(s);
            comp_errors(s, comp, q, brem);
            const uint32_t wi = rdlu(idx, l), w0 = expand_comp(comp, wi >> shift), w1 = expand_comp(comp, wi & (lim - 1));
            best = q.at(0, w0) + q.at(1, w1) + q.at(2, (w0 * 2 + w1) / 3) + q.at(3, (w0 + w1 * 2) / 3);
            // the pairs after the one taken, filtered again with the new tables
            pos = scan = wi + 1;
            cnt = 0;
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

// try_combinatorial_encoding (crn_dxt1.cpp:1886-1997): the packed colours on
// lanes 0..np-1, every pair i < j in order, then again with the alternate
// rounding where only an exact solution counts.
__device__ __forceinline__ void try_combinatorial(Wb& s)
{
    if (s.nu < 2 || s.nu > 4) return;
    uint32_t pk = 0;
    uint32_t np = 0;
    auto add = [&](int r, int g, int b) {
        const uint32_t p = pack565(r, g, b, 1);
        if (ballot((uint32_t)lane_id() < np && pk == p)) return;
        if ((uint32_t)lane_id() == np) pk = p;
        np++;
    };
    for (int k = 0; k < 4; k++)
        if (k < (int)s.nu) add(colour_r(s, k), colour_g(s, k), colour_b(s, k));
    if (s.nu == 2) {
        const float fs[10] = {2.0f, 3.0f, .5f, 1.5f, -1.0f, 2.0f, -.5f, .5f, -2.0f, -1.0f};
        for (int k = 0; k < 2; k++)
            for (int q = 0; q < 2; q++) {
                const int a = q, b2 = q ^ 1;
                const int ar = colour_r(s, a), ag = colour_g(s, a), ab = colour_b(s, a);
                const int br = colour_r(s, b2), bg = colour_g(s, b2), bb = colour_b(s, b2);
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
                const int ar = colour_r(s, i), ag = colour_g(s, i), ab = colour_b(s, i);
                const int br = colour_r(s, j), bg = colour_g(s, j), bb = colour_b(s, j);
                for (int f = 0; f < 4; f++) {
                    int r, g, b;
                    lerp_color(ar, ag, ab, br, bg, bb, fs[f], 1, r, g, b);
                    add(r, g, b);
                }
            }
    }
    const uint32_t npairs = np * (np - 1) / 2;
    auto pair = [&](uint32_t t, uint32_t& lo, uint32_t& hi) {
        uint32_t i = 0, row = np - 1;
        while (t >= row) {
            t -= row;
            i++;
            row--;
        }
        lo = (uint32_t)__shfl((int)pk, (int)i);
        hi = (uint32_t)__shfl((int)pk, (int)(i + 1 + t));
        return true;
    };
    eval_list(s, npairs, 0, pair);
    const uint32_t err = s.best.err;
    if (err) s.best.err = 1;
    eval_list(s, npairs, 1, pair);
    if (s.best.err) s.best.err = err;
}

// the probe sweep along the axis (crn_dxt1.cpp:840-892): the packed probes on
// lanes 0..n-1
__device__ __forceinline__ uint32_t sweep(const int16_t* tab, uint32_t range, const float* base, const float (*ax)[3], uint32_t& n)
{
    int prev[2] = {-1, -1};
    n = 0;
    uint32_t list = 0;
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
                if ((uint32_t)lane_id() == n) list = (uint32_t)packed;
                n++;
                prev[sg] = packed;
            }
        }
    }
    return list;
}

// the lattice moves of optimize_endpoints (crn_dxt1.cpp:905-1016) around the
// low (which = 0) or high endpoint
__device__ __forceinline__ void lattice(Wb& s, int which)
{
    int lr, lg, lb;
    unpack565(which ? s.best.hi : s.best.lo, 0, lr, lg, lb);
    eval_chain(s, 26, [&](uint32_t i, uint32_t& lo, uint32_t& hi) {
        const int r = lr + k_adj[i][0], g = lg + k_adj[i][1], b = lb + k_adj[i][2];
        if (r < 0 || r > 31 || g < 0 || g > 63 || b < 0 || b > 31) return false;
        const uint32_t p = pack565(r, g, b, 0);
        lo = which ? s.best.lo : p;
        hi = which ? p : s.best.hi;
        canon(lo, hi);
        return true;
    });
    if (s.quality == 4) {
        unpack565(which ? s.best.hi : s.best.lo, 0, lr, lg, lb);
        eval_chain(s, 6, [&](uint32_t i, uint32_t& lo, uint32_t& hi) {
            const int a = (int)i >> 1, sg = (i & 1) ? 2 : -2;
            const int r = lr + (a == 0 ? sg : 0), g = lg + (a == 1 ? sg : 0), b = lb + (a == 2 ? sg : 0);
            if (r < 0 || r > 31 || g < 0 || g > 63 || b < 0 || b > 31) return false;
            const uint32_t p = pack565(r, g, b, 0);
            lo = which ? s.best.lo : p;
            hi = which ? p : s.best.hi;
            canon(lo, hi);
            return true;
        });
    }
}

// optimize_endpoints (crn_dxt1.cpp:703-1067)
__device__ __forceinline__ void optimize_endpoints(Wb& s, float* lo, float* hi)
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
    for (uint32_t pass = 0; pass < passes; pass++) {
        if (pass) {
            int r, g, b;
            unpack565(s.best.lo, 0, r, g, b);
            lo[0] = (float)r; lo[1] = (float)g; lo[2] = (float)b;
            unpack565(s.best.hi, 0, r, g, b);
            hi[0] = (float)r; hi[1] = (float)g; hi[2] = (float)b;
        }
        const uint32_t prev = s.best.err;
        if (!prev) break;
        uint32_t nl = 0, nh = 0;
        const uint32_t plow = sweep(tab, range, lo, ax, nl);
        const uint32_t phigh = sweep(tab, range, hi, ax, nh);
        eval_list(s, nl * nh, 0, [&](uint32_t k, uint32_t& a, uint32_t& b) {
            a = (uint32_t)__shfl((int)plow, (int)(k / nh));
            b = (uint32_t)__shfl((int)phigh, (int)(k % nh));
            canon(a, b);
            return true;
        });
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
            if (sel_all_equal(s)) solid = try_average_solid(s);
            if (!solid && s.quality == 4) optimize_comps(s);
        }
        if (s.quality == 4 && s.best.err) try_combinatorial(s);
    }
}

// compute_vectors (crn_dxt1.cpp:155-189): normalised colours (this lane's),
// their weighted means summed in colour order; returns the weighted vector
// of this lane's colour minus the weighted mean
__device__ __forceinline__ void compute_vectors(Wb& s, const float* pw, float& nwr, float& nwg, float& nwb)
{
    const float w = (float)s.cw;
    const float n0 = (float)s.cr * 1.0f / 255.0f, n1 = (float)s.cg * 1.0f / 255.0f, n2 = (float)s.cb * 1.0f / 255.0f;
    nwr = pw[0] * n0;
    nwg = pw[1] * n1;
    nwb = pw[2] * n2;
    const float a0 = n0 * w, a1 = n1 * w, a2 = n2 * w, b0 = nwr * w, b1 = nwg * w, b2 = nwb * w;
    float m[3] = {0, 0, 0}, mw[3] = {0, 0, 0};
    for (int k = 0; k < (int)s.nu; k++) {
        m[0] += rdlf(a0, k); m[1] += rdlf(a1, k); m[2] += rdlf(a2, k);
        mw[0] += rdlf(b0, k); mw[1] += rdlf(b1, k); mw[2] += rdlf(b2, k);
    }
    if (s.total_w) {
        const float inv = 1.0f / (float)s.total_w;
        for (int c = 0; c < 3; c++) {
            m[c] *= inv;
            mw[c] *= inv;
        }
    }
    s.nr = n0 - m[0];
    s.ng = n1 - m[1];
    s.nb = n2 - m[2];
    nwr -= mw[0];
    nwg -= mw[1];
    nwb -= mw[2];
    for (int c = 0; c < 3; c++) s.mean[c] = m[c];
}

// compute_pca (crn_dxt1.cpp:192-256) of the lanes' vectors (r, g, b)
__device__ __forceinline__ void compute_pca(const Wb& s, float r, float g, float b, float* axis)
{
    double t[6];
    if (s.cw > 1) {
        const double w = s.cw;
        t[0] = (double)(r * r) * w; t[1] = (double)(r * g) * w; t[2] = (double)(r * b) * w;
        t[3] = (double)(g * g) * w; t[4] = (double)(g * b) * w; t[5] = (double)(b * b) * w;
    } else {
        t[0] = (double)(r * r); t[1] = (double)(r * g); t[2] = (double)(r * b);
        t[3] = (double)(g * g); t[4] = (double)(g * b); t[5] = (double)(b * b);
    }
    double cov[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < (int)s.nu; k++)
        for (int q = 0; q < 6; q++) cov[q] += rdld(t[q], k);
    double vr = (double).9f, vg = 1.0, vb = (double).7f;
    for (int it = 0; it < 8; it++) {
        double x = vr * cov[0] + vg * cov[1] + vb * cov[2];
        double y = vr * cov[1] + vg * cov[3] + vb * cov[4];
        double z = vr * cov[2] + vg * cov[4] + vb * cov[5];
        double m = fabs(x);
        if (fabs(y) > m) m = fabs(y);
        if (fabs(z) > m) m = fabs(z);
        if (m > 1e-10) {
            m = 1.0 / m;
            x *= m;
            y *= m;
            z *= m;
        }
        const double d = (vr - x) * (vr - x) + (vg - y) * (vg - y) + (vb - z) * (vb - z);
        vr = x;
        vg = y;
        vb = z;
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

// ray_aabb against the unit cube (crn_intersect.h:44-120)
__device__ __forceinline__ bool ray_unit_box(float* coord, const float* org, const float* dir)
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

// handle_multicolor_block (crn_dxt1.cpp:1069-1178)
__device__ __forceinline__ void handle_multicolor(Wb& s)
{
    uint32_t npass = 1;
    float pw[3] = {1.0f, 1.0f, 1.0f};
    if (s.perceptual) {
        const int l = (s.cr + s.cg + s.cb + 1) / 3;
        const float lf = (float)l;
        const float scale = (float)s.cw / (lf > 1.0f ? lf : 1.0f);
        const float tr = scale * (float)s.cr, tb = scale * (float)s.cb, tl = (float)l;
        float red = 0, blue = 0, al = 0;
        for (int k = 0; k < (int)s.nu; k++) {
            red += rdlf(tr, k);
            blue += rdlf(tb, k);
            al += rdlf(tl, k);
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
    for (uint32_t pass = 0; pass < npass; pass++) {
        float nwr, nwg, nwb;
        compute_vectors(s, pw, nwr, nwg, nwb);
        compute_pca(s, nwr, nwg, nwb, s.axis);
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
    float d = s.nr * s.axis[0];
    d += s.ng * s.axis[1];
    d += s.nb * s.axis[2];
    const bool mine = (uint32_t)lane_id() < s.nu;
    const float l = wave_minf(mine ? d : 1e+9f), h = wave_maxf(mine ? d : -1e+9f);
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

// the 2-bit values v of lanes 0..15 as one selector word (texel i at bit 2i)
__device__ inline uint32_t pack_selectors(uint32_t v)
{
    const uint32_t b0 = (uint32_t)ballot(((v & 1u) != 0) && lane_id() < 16) & 0xFFFFu;
    const uint32_t b1 = (uint32_t)ballot(((v & 2u) != 0) && lane_id() < 16) & 0xFFFFu;
    uint32_t w = 0;
    for (int i = 0; i < 16; i++) w |= (((b0 >> i) & 1u) << (2 * i)) | (((b1 >> i) & 1u) << (2 * i + 1));
    return w;
}

// compute_internal + return_solution (crn_dxt1.cpp:2081-2232, 263-365) for
// the block whose texel i is px on lane i (i < 16).  DXT1A (flags bit 4):
// texels with alpha below the threshold (flags bits 8..15) leave the colour
// set and take selector 3.
__device__ __forceinline__ Dxt1Out encode_core(uint32_t px, int quality, int flags, uint32_t* lds)
{
    Wb s;
    s.quality = quality;
    const int perceptual_req = flags & 1;
    s.use_alpha = (flags >> 1) & 1;
    s.grayscale = (flags >> 2) & 1;
    s.force_alpha = (flags >> 3) & 1;
    s.have_alpha = (flags >> 4) & 1;
    const uint32_t thresh = ((uint32_t)flags >> 8) & 255u;
    s.hc = quality == 4 && !s.have_alpha && !s.force_alpha && !s.use_alpha && !s.grayscale;
    s.perceptual = perceptual_req && !s.grayscale;
    s.planes = lds;
    s.pend = lds + kPlaneWords;
    s.pbound = lds + kPlaneWords + 128;
    const int me = lane_id();
    const bool texel = me < 16;
    const bool opaque = texel && !(s.have_alpha && (px >> 24) < thresh);
    const uint32_t key = px | 0xFF000000u;
    const uint32_t omask = (uint32_t)ballot(opaque) & 0xFFFFu;
    // first appearance of each colour among the opaque texels
    int firstj = -1;
    for (int j = 0; j < 16; j++) {
        const uint32_t kj = rdlu(key, j);
        if (firstj < 0 && j < me && ((omask >> j) & 1u) && kj == key) firstj = j;
    }
    const bool first = opaque && firstj < 0;
    const uint32_t fmask = (uint32_t)ballot(first) & 0xFFFFu;
    s.nu = (uint32_t)__popc(fmask);
    s.total_w = (uint32_t)__popc(omask);
    const uint32_t tidx = (uint32_t)__popc(fmask & ((1u << (first ? me : (firstj < 0 ? 0 : firstj))) - 1u));
    // colour k on lane k: the k-th first appearance, weight = its texels
    int src = 0;
    {
        int cnt = 0;
        for (int j = 0; j < 16; j++)
            if ((fmask >> j) & 1u) {
                if (cnt == me) src = j;
                cnt++;
            }
    }
    const uint32_t ckey = (uint32_t)__shfl((int)key, src);
    const bool colour = (uint32_t)me < s.nu;
    s.cr = colour ? (int)(ckey & 0xFF) : 0;
    s.cg = colour ? (int)((ckey >> 8) & 0xFF) : 0;
    s.cb = colour ? (int)((ckey >> 16) & 0xFF) : 0;
    uint32_t w = 0;
    for (int j = 0; j < 16; j++) {
        const uint32_t tj = rdlu(tidx, j);
        if (((omask >> j) & 1u) && tj == (uint32_t)me) w++;
    }
    s.cw = colour ? w : 0u;
    s.cpk = colour ? (ckey & 0x00FFFFFFu) | (w << 24) : 0u;
    s.nr = s.ng = s.nb = 0.0f;
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
    build_planes(s);
    s.best.lo = s.best.hi = 0;
    s.best.err = kNone;
    s.best.alpha = s.best.alt = s.best.enforce = s.best.enforced = 0;
    s.best.sel_valid = false;
    s.bsel = 0;
    s.mean[0] = s.mean[1] = s.mean[2] = 0.0f;
    s.axis[0] = s.axis[1] = s.axis[2] = 0.0f;
    if (s.nu == 1 && omask == 0xFFFFu) {
        const uint32_t r = (uint32_t)colour_r(s, 0), g = (uint32_t)colour_g(s, 0), b = (uint32_t)colour_b(s, 0);
        eval_list(s, s.use_alpha ? 2u : 1u, 0, [&](uint32_t i, uint32_t& lo, uint32_t& hi) {
            if (i) {
                lo = ((uint32_t)g_om5_3[r][0] << 11) | ((uint32_t)g_om6_3[g][0] << 5) | g_om5_3[b][0];
                hi = ((uint32_t)g_om5_3[r][1] << 11) | ((uint32_t)g_om6_3[g][1] << 5) | g_om5_3[b][1];
            } else {
                lo = ((uint32_t)g_om5[r][0] << 11) | ((uint32_t)g_om6[g][0] << 5) | g_om5[b][0];
                hi = ((uint32_t)g_om5[r][1] << 11) | ((uint32_t)g_om6[g][1] << 5) | g_om5[b][1];
            }
            return true;
        });
    } else {
        handle_multicolor(s);
    }
    compute_selectors(s);
    const bool inv = s.best.alpha ? (s.best.lo > s.best.hi) : (s.best.lo < s.best.hi);
    o.error = s.best.err;
    const uint32_t lo = inv ? s.best.hi : s.best.lo, hi = inv ? s.best.lo : s.best.hi;
    o.lo_hi = lo | (hi << 16);
    uint32_t v = (uint32_t)__shfl((int)s.bsel, (int)tidx);
    if (inv) v = s.best.alpha ? (v < 2 ? v ^ 1u : v) : (v ^ 1u);
    if (!opaque) v = 3u;
    o.selectors = pack_selectors(v);
    o.flags = (uint32_t)s.best.alpha | ((uint32_t)inv << 1) | ((uint32_t)s.best.alt << 2) |
              ((uint32_t)s.best.enforce << 3) | ((uint32_t)s.best.enforced << 4);
    return o;
}

__device__ inline uint32_t metric(int perceptual, int grayscale, uint32_t p, int r, int g, int b)
{
    return dist(perceptual, grayscale, (int)(p & 255u), (int)((p >> 8) & 255u), (int)((p >> 16) & 255u), r, g, b);
}

// dxt1_endpoint_optimizer::compute (crn_dxt1.cpp:2234-2246).  flags bit 6
// applies dxt_image's DXT1A rule first (crn_dxt_image.cpp:1439-1450: the
// block has alpha when a texel's alpha is below the threshold); bit 5 with
// alpha blocks allowed and no transparent texel runs
// try_alpha_as_black_optimization (:2001-2079): texels whose r, g and b are
// all <= 4 are made transparent, the block is searched again as DXT1A, and
// that 3-colour block replaces the first result if its error, counting those
// texels against black, is lower.
__device__ __forceinline__ Dxt1Out encode_block(uint32_t px, int quality, int flags, uint32_t* lds)
{
    const bool texel = lane_id() < 16;
    if ((flags >> 6) & 1) {
        const uint32_t thresh = ((uint32_t)flags >> 8) & 255u;
        const bool any = ballot(texel && (px >> 24) < thresh) != 0;
        flags = (flags & ~(64 | 16)) | (any ? 16 : 0);
    }
    const bool isdark = texel && (px & 255u) <= 4 && ((px >> 8) & 255u) <= 4 && ((px >> 16) & 255u) <= 4;
    const uint32_t dark = (uint32_t)ballot(isdark) & 0xFFFFu;
    const bool black = ((flags >> 1) & 1) && ((flags >> 5) & 1) && !((flags >> 4) & 1) && dark && dark != 0xFFFFu;
    Dxt1Out first = encode_core(px, quality, flags, lds);
    if (!black) return first;
    const uint32_t p2 = isdark ? (px & 0x00FFFFFFu) : px;
    const Dxt1Out o = encode_core(p2, quality, flags | 16, lds);
    // get_block_colors3 of the trial's endpoints, entry 3 black
    int c0r, c0g, c0b, c1r, c1g, c1b;
    unpack565(o.lo_hi & 0xFFFFu, 1, c0r, c0g, c0b);
    unpack565(o.lo_hi >> 16, 1, c1r, c1g, c1b);
    const int c2r = (c0r + c1r) >> 1, c2g = (c0g + c1g) >> 1, c2b = (c0b + c1b) >> 1;
    const int gray = (flags >> 2) & 1, perc = (flags & 1) && !gray;
    const uint32_t v = (o.selectors >> (2 * (lane_id() & 15))) & 3u;
    const int r = v == 0 ? c0r : v == 1 ? c1r : v == 2 ? c2r : 0;
    const int g = v == 0 ? c0g : v == 1 ? c1g : v == 2 ? c2g : 0;
    const int b = v == 0 ? c0b : v == 1 ? c1b : v == 2 ? c2b : 0;
    uint32_t e = texel ? metric(perc, gray, p2, r, g, b) : 0u;
    for (int off = 8; off; off >>= 1) e += (uint32_t)__shfl_xor((int)e, off, 16);
    const uint64_t err = rdlu(e, 0);
    if (err < first.error) {
        first.error = err;
        first.lo_hi = o.lo_hi;
        first.selectors = o.selectors;
        first.flags |= 1u;
    }
    return first;
}

constexpr uint32_t kWaves = 4;   // waves (blocks) per workgroup of the fixed-shape kernels

__global__ void __launch_bounds__(256) dxt1_blocks_kernel(const uint32_t* __restrict__ px, uint32_t nblocks, int quality,
                                                          int flags, Dxt1Out* __restrict__ out)
{
    __shared__ uint32_t lds[kWaves][kLdsWords];
    const uint32_t w = threadIdx.x >> 6, b = blockIdx.x * (blockDim.x >> 6) + w;
    if (b >= nblocks) return;
    const uint32_t p = lane_id() < 16 ? px[16 * (size_t)b + lane_id()] : 0u;
    const Dxt1Out o = encode_block(p, quality, flags, lds[w]);
    if (lane_id() == 0) out[b] = o;
}

// Whole image, as dxt_image::init_task and set_block_pixels do for cColorDXT1
// elements (crnlib/crn_dxt_image.cpp:283-349, 1436-1484): the block's 4x4
// texels with edge clamp, the search, then the 8-byte dxt1_block (low colour,
// high colour, 2-bit selectors row by row, texel (x, y) at bit 2 (x + 4 y) of
// the last four bytes) at blocks[i * stride]; errors (optional) = m_error.
__global__ void __launch_bounds__(256) dxt1_image_kernel(const uint32_t* __restrict__ rgba, uint32_t width,
                                                         uint32_t height, uint32_t blocks_x, uint32_t nblocks,
                                                         int quality, int flags, uint64_t* __restrict__ out,
                                                         uint32_t stride, uint64_t* __restrict__ err)
{
    __shared__ uint32_t lds[kWaves][kLdsWords];
    const uint32_t w = threadIdx.x >> 6, b = blockIdx.x * (blockDim.x >> 6) + w;
    if (b >= nblocks) return;
    const uint32_t bx = b % blocks_x, by = b / blocks_x;
    const uint32_t l = (uint32_t)lane_id() & 15u;
    const uint32_t p = rgba[(size_t)min(by * 4 + (l >> 2), height - 1) * width + min(bx * 4 + (l & 3), width - 1)];
    const Dxt1Out o = encode_block(p, quality, flags, lds[w]);
    if (lane_id() == 0) {
        out[(size_t)b * stride] = (uint64_t)o.lo_hi | ((uint64_t)o.selectors << 32);
        if (err) err[b] = o.error;
    }
}

// Texture-set encode: colour elements of many level images in one flat
// order, each wave walking it with a grid stride (a persistent grid when the
// launch is smaller than the work).
__global__ void __launch_bounds__(256) dxt1_set_kernel(const crn2::SetJob* __restrict__ jobs, uint32_t njobs,
                                                       uint64_t begin, uint64_t end, int quality)
{
    __shared__ uint32_t lds[kWaves][kLdsWords];
    const uint32_t w = threadIdx.x >> 6, wpg = blockDim.x >> 6;
    const uint64_t step = (uint64_t)gridDim.x * wpg;
    for (uint64_t i = begin + (uint64_t)blockIdx.x * wpg + w; i < end; i += step) {
        const crn2::SetJob j = jobs[crn2::set_job_of(jobs, njobs, i)];
        const uint32_t b = (uint32_t)(i - j.first);
        const uint32_t bx = b % j.blocks_x, by = b / j.blocks_x;
        const uint32_t l = (uint32_t)lane_id() & 15u;
        const uint32_t p =
            j.rgba[(size_t)min(by * 4 + (l >> 2), j.height - 1) * j.width + min(bx * 4 + (l & 3), j.width - 1)];
        const Dxt1Out o = encode_block(p, quality, (int)j.flags, lds[w]);
        if (lane_id() == 0) j.out[(size_t)b * j.stride] = (uint64_t)o.lo_hi | ((uint64_t)o.selectors << 32);
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
    // launch shapes (one block per wavefront; 1, 2 or 4 wavefronts per workgroup; identical results)
    auto launch = [&](int shape) -> hipError_t {
        const uint32_t tb = 64u << shape, per = tb / 64, nb = (nblocks + per - 1) / per;
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
    const uint32_t tb = 64 * kWaves;
    hipLaunchKernelGGL(dxt1_image_kernel, dim3((uint32_t)((n + kWaves - 1) / kWaves)), dim3(tb), 0,
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
