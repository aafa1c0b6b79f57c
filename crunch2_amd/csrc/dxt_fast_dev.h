// dxt_fast_dev.h -- device restatement of crnlib's dxt_fast colour block
// compressor (crnlib/crn_dxt_fast.cpp:25-764, refinement off) for N texels
// held in registers, shared by the dxt_hc tile kernels and the qdxt kernels.
//   optimize_block_colors .. :138-268 (box, covariance power iteration in
//                            double, projection extremes)
//   match_block_colors ..... :88-136 (selectors by projection thresholds)
//   refine_block ........... :270-337 (one least-squares refit)
//   determine_selectors .... :339-353
// Every function is an exact restatement with the reference's types; the
// translation units that include this are built with -ffp-contract=off.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace dxtf {

__device__ inline int ch(uint32_t p, int c) { return (int)((p >> (8 * c)) & 255u); }

__device__ inline int mul_8bit(int a, int b)
{
    const int t = a * b + 128;
    return (t + (t >> 8)) >> 8;
}

__device__ inline void unpack565(uint32_t v, int* c)
{
    const int r = (int)((v >> 11) & 31), g = (int)((v >> 5) & 63), b = (int)(v & 31);
    c[0] = (r << 3) | (r >> 2);
    c[1] = (g << 2) | (g >> 4);
    c[2] = (b << 3) | (b >> 2);
}

__device__ inline uint32_t pack_color(uint32_t p)
{
    return (uint32_t)((mul_8bit(ch(p, 0), 31) << 11) + (mul_8bit(ch(p, 1), 63) << 5) + mul_8bit(ch(p, 2), 31));
}

__device__ inline void eval_colors(int (&col)[4][3], uint32_t c0, uint32_t c1)
{
    unpack565(c0, col[0]);
    unpack565(c1, col[1]);
    for (int k = 0; k < 3; k++) {
        col[2][k] = (col[0][k] * 2 + col[1][k]) / 3;
        col[3][k] = (col[1][k] * 2 + col[0][k]) / 3;
    }
}

template <int N>
__device__ bool match_block_colors(const uint32_t (&px)[N], int (&col)[4][3], uint8_t (&sel)[N])
{
    const int dr = col[0][0] - col[1][0], dg = col[0][1] - col[1][1], db = col[0][2] - col[1][2];
    int stops[4];
    for (int i = 0; i < 4; i++) stops[i] = col[i][0] * dr + col[i][1] * dg + col[i][2] * db;
    const int c0p = (stops[1] + stops[3]) >> 1, half = (stops[3] + stops[2]) >> 1, c3p = (stops[2] + stops[0]) >> 1;
    bool status = false;
#pragma unroll
    for (int i = 0; i < N; i++) {
        const int dot = ch(px[i], 0) * dr + ch(px[i], 1) * dg + ch(px[i], 2) * db;
        uint8_t s;
        if (dot < half) s = dot < c0p ? 1 : 3;
        else s = dot < c3p ? 2 : 0;
        sel[i] = s;
        if (s != sel[0]) status = true;
    }
    return status;
}

template <int N>
__device__ bool optimize_block_colors(const uint32_t (&px)[N], uint32_t& max16, uint32_t& min16, uint32_t (&ave)[3])
{
    int mn[3], mx[3];
    for (int c = 0; c < 3; c++) {
        int mu = ch(px[0], c), a = mu, b = mu;
#pragma unroll
        for (int i = 1; i < N; i++) {
            const int v = ch(px[i], c);
            mu += v;
            a = a < v ? a : v;
            b = b > v ? b : v;
        }
        ave[c] = (uint32_t)((mu + N / 2) / N);
        mn[c] = a;
        mx[c] = b;
    }
    if (mn[0] == mx[0] && mn[1] == mx[1] && mn[2] == mx[2]) return false;
    double cov[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < N; i++) {
        const double r = ch(px[i], 0) - (int)ave[0];
        const double g = ch(px[i], 1) - (int)ave[1];
        const double b = ch(px[i], 2) - (int)ave[2];
        cov[0] += r * r;
        cov[1] += r * g;
        cov[2] += r * b;
        cov[3] += g * g;
        cov[4] += g * b;
        cov[5] += b * b;
    }
    double covf[6];
    for (int i = 0; i < 6; i++) covf[i] = cov[i] * (double)(1.0f / 255.0f);
    double vfr = mx[0] - mn[0], vfg = mx[1] - mn[1], vfb = mx[2] - mn[2];
    for (int it = 0; it < 4; it++) {
        const double r = vfr * covf[0] + vfg * covf[1] + vfb * covf[2];
        const double g = vfr * covf[1] + vfg * covf[3] + vfb * covf[4];
        const double b = vfr * covf[2] + vfg * covf[4] + vfb * covf[5];
        vfr = r;
        vfg = g;
        vfb = b;
    }
    double magn = fabs(vfr) > fabs(vfg) ? fabs(vfr) : fabs(vfg);
    magn = magn > fabs(vfb) ? magn : fabs(vfb);
    int vr, vg, vb;
    if (magn < 4.0f) {
        vr = 148;
        vg = 300;
        vb = 58;
    } else {
        magn = 512.0f / magn;
        vfr *= magn;
        vfg *= magn;
        vfb *= magn;
        vr = (int)vfr;
        vg = (int)vfg;
        vb = (int)vfb;
    }
    int mind = ch(px[0], 0) * vr + ch(px[0], 1) * vg + ch(px[0], 2) * vb, maxd = mind;
    uint32_t pmin = px[0], pmax = px[0];
#pragma unroll
    for (int i = 1; i < N; i++) {
        const int dot = ch(px[i], 0) * vr + ch(px[i], 1) * vg + ch(px[i], 2) * vb;
        if (dot < mind) {
            mind = dot;
            pmin = px[i];
        }
        if (dot > maxd) {
            maxd = dot;
            pmax = px[i];
        }
    }
    max16 = pack_color(pmax);
    min16 = pack_color(pmin);
    return true;
}

__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

template <int N>
__device__ bool refine_block(const uint32_t (&px)[N], uint32_t& max16, uint32_t& min16, const uint8_t (&sel)[N])
{
    const int w1t[4] = {3, 0, 2, 1}, p0[4] = {0, 0, 2, 2}, p1[4] = {0, 9, 1, 4}, p2[4] = {9, 0, 4, 1};
    double a0 = 0, a1 = 0, a2 = 0, t1r = 0, t1g = 0, t1b = 0, t2r = 0, t2g = 0, t2b = 0;
#pragma unroll
    for (int i = 0; i < N; i++) {
        const double r = ch(px[i], 0), g = ch(px[i], 1), b = ch(px[i], 2);
        const int st = sel[i], w1 = w1t[st];
        a0 += p0[st];
        a1 += p1[st];
        a2 += p2[st];
        t1r += w1 * r;
        t1g += w1 * g;
        t1b += w1 * b;
        t2r += r;
        t2g += g;
        t2b += b;
    }
    t2r = 3 * t2r - t1r;
    t2g = 3 * t2g - t1g;
    t2b = 3 * t2b - t1b;
    const double xx = a2, yy = a1, xy = a0;
    const double t = xx * yy - xy * xy;
    if (!yy || !xx || fabs(t) < .0000125f) return false;
    const double frb = (double)(3.0f * 31.0f / 255.0f) / t;
    const double fg = frb * (double)(63.0f / 31.0f);
    const uint32_t om = min16, ox = max16;
    uint32_t x = (uint32_t)clampi((int)((t1r * yy - t2r * xy) * frb + 0.5f), 0, 31) << 11;
    x |= (uint32_t)clampi((int)((t1g * yy - t2g * xy) * fg + 0.5f), 0, 63) << 5;
    x |= (uint32_t)clampi((int)((t1b * yy - t2b * xy) * frb + 0.5f), 0, 31);
    uint32_t m = (uint32_t)clampi((int)((t2r * xx - t1r * xy) * frb + 0.5f), 0, 31) << 11;
    m |= (uint32_t)clampi((int)((t2g * xx - t1g * xy) * fg + 0.5f), 0, 63) << 5;
    m |= (uint32_t)clampi((int)((t2b * xx - t1b * xy) * frb + 0.5f), 0, 31);
    max16 = x;
    min16 = m;
    return om != m || ox != x;
}

template <int N>
__device__ bool determine_selectors(const uint32_t (&px)[N], uint32_t min16, uint32_t max16, uint8_t (&sel)[N])
{
    if (max16 != min16) {
        int col[4][3];
        eval_colors(col, min16, max16);
        return match_block_colors<N>(px, col, sel);
    }
    for (int i = 0; i < N; i++) sel[i] = 0;
    return false;
}

// dxt_fast::compress_color_block (crn_dxt_fast.cpp:725-764, refine = false):
// low16 >= high16 on return, selectors flipped with the swap; solid blocks
// from the single-colour tables om5 / om6 (OMatch5 / OMatch6).
template <int N>
__device__ void compress_color(const uint32_t (&px)[N], uint32_t& lo, uint32_t& hi, uint8_t (&sel)[N],
                               const uint8_t (*om5)[2], const uint8_t (*om6)[2])
{
    uint32_t ave[3];
    bool solid = false;
    lo = hi = 0;
    if (!optimize_block_colors<N>(px, lo, hi, ave)) solid = true;
    else if (!determine_selectors<N>(px, lo, hi, sel)) solid = true;
    else if (refine_block<N>(px, lo, hi, sel)) determine_selectors<N>(px, lo, hi, sel);
    if (solid) {
        for (int i = 0; i < N; i++) sel[i] = 2;
        lo = (uint32_t)om5[ave[0]][0] << 11 | (uint32_t)om6[ave[1]][0] << 5 | om5[ave[2]][0];
        hi = (uint32_t)om5[ave[0]][1] << 11 | (uint32_t)om6[ave[1]][1] << 5 | om5[ave[2]][1];
    }
    if (lo < hi) {
        const uint32_t t = lo;
        lo = hi;
        hi = t;
        for (int i = 0; i < N; i++) sel[i] ^= 1;
    }
}

// The 4-colour single-colour tables (ryg's OMatch5 / OMatch6,
// crnlib/crn_ryg_dxt.cpp:58-99): for every 8-bit value the 5- or 6-bit
// endpoint pair whose 2/3 point is nearest, with a small penalty on the
// endpoints' spread.  Built on the host.
inline void build_omatch4(uint8_t (*t5)[2], uint8_t (*t6)[2])
{
    for (int which = 0; which < 2; which++) {
        const int size = which ? 64 : 32;
        uint8_t(*t)[2] = which ? t6 : t5;
        for (int i = 0; i < 256; i++) {
            int best = 256;
            for (int mn = 0; mn < size; mn++)
                for (int mx = 0; mx < size; mx++) {
                    const int me = size == 32 ? ((mn << 3) | (mn >> 2)) : ((mn << 2) | (mn >> 4));
                    const int xe = size == 32 ? ((mx << 3) | (mx >> 2)) : ((mx << 2) | (mx >> 4));
                    int err = ((xe * 2 + me) / 3) - i;
                    err = err < 0 ? -err : err;
                    const int spread = xe - me < 0 ? me - xe : xe - me;
                    err += (spread * 8) >> 8;
                    if (err < best) {
                        t[i][0] = (uint8_t)mx;
                        t[i][1] = (uint8_t)mn;
                        best = err;
                    }
                }
        }
    }
}

}  // namespace dxtf
