// synth.cpp -- seeded synthetic textures, box-filter mips and the simple CPU
// block quantizer that turns RGBA levels into a .crn through crn2_pack.
//
// Texture recipe (deterministic, integer-exact across hosts: only +,-,*,/ and
// sqrt, which IEEE 754 rounds correctly):
//   RGB   = clamp(sum of 4 triangle-wave ramps along seeded directions
//           (amplitude 255/4 each) + 3 octaves of smoothstep value noise with
//           lattice spacing 64/32/16 px and amplitude 40/20/10, 0, 255)
//   alpha = seeded radial gradient, 255 at a seeded centre falling to 0 at a
//           seeded radius
//   normal maps: unit normals of a 4-octave value-noise height field by
//           central differences, X and Y mapped to 0..255 in R and G.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <unordered_map>
#include <vector>

#include "crn2_internal.h"

namespace {

struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed) {}
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

static inline uint32_t hash2(uint32_t x, uint32_t y, uint32_t seed)
{
    uint32_t h = x * 0x8DA6B343u ^ y * 0xD8163841u ^ seed * 0xCB1AB31Fu;
    h ^= h >> 16;
    h *= 0x7FEB352Du;
    h ^= h >> 15;
    h *= 0x846CA68Bu;
    h ^= h >> 16;
    return h;
}

// value noise in [-1, 1], lattice spacing `cell` px
static inline double value_noise(double x, double y, double cell, uint32_t seed)
{
    double fx = x / cell, fy = y / cell;
    double ix = std::floor(fx), iy = std::floor(fy);
    double tx = fx - ix, ty = fy - iy;
    tx = tx * tx * (3.0 - 2.0 * tx);
    ty = ty * ty * (3.0 - 2.0 * ty);
    uint32_t x0 = (uint32_t)(int64_t)ix, y0 = (uint32_t)(int64_t)iy;
    auto v = [&](uint32_t a, uint32_t b) { return (double)(hash2(a, b, seed) >> 8) * (2.0 / 16777215.0) - 1.0; };
    double a = v(x0, y0), b = v(x0 + 1, y0), c = v(x0, y0 + 1), d = v(x0 + 1, y0 + 1);
    double top = a + (b - a) * tx, bot = c + (d - c) * tx;
    return top + (bot - top) * ty;
}

static inline double tri(double t)
{
    double f = t - std::floor(t);
    return f < 0.5 ? 2.0 * f : 2.0 - 2.0 * f;
}

static inline uint8_t clamp8(double v)
{
    if (v < 0.0) return 0;
    if (v > 255.0) return 255;
    return (uint8_t)(int)(v + 0.5);
}

}  // namespace

// kind: 0 = RGB (alpha 255), 1 = RGBA (radial alpha), 2 = normal map (XY in RG, B = Z, A = 255)
extern "C" int crn2_synth_rgba(uint32_t w, uint32_t h, uint64_t seed, int kind, uint8_t* out)
{
    if (!out || !w || !h || kind < 0 || kind > 2) return CRN2_EINVAL;
    Rng rng(seed);
    if (kind == 2) {
        const uint32_t s = (uint32_t)rng.next();
        std::vector<double> H((size_t)(w + 2) * (h + 2));
        for (uint32_t y = 0; y < h + 2; y++)
            for (uint32_t x = 0; x < w + 2; x++) {
                double hx = (double)x - 1.0, hy = (double)y - 1.0, v = 0.0, amp = 32.0, cell = 64.0;
                for (int o = 0; o < 4; o++, amp *= 0.5, cell *= 0.5) v += amp * value_noise(hx, hy, cell, s + o);
                H[(size_t)y * (w + 2) + x] = v;
            }
        for (uint32_t y = 0; y < h; y++)
            for (uint32_t x = 0; x < w; x++) {
                double dx = H[(size_t)(y + 1) * (w + 2) + x + 2] - H[(size_t)(y + 1) * (w + 2) + x];
                double dy = H[(size_t)(y + 2) * (w + 2) + x + 1] - H[(size_t)y * (w + 2) + x + 1];
                double nx = -dx * 0.5, ny = -dy * 0.5, nz = 1.0;
                double inv = 1.0 / std::sqrt(nx * nx + ny * ny + nz * nz);
                uint8_t* p = out + 4 * ((size_t)y * w + x);
                p[0] = clamp8(nx * inv * 127.5 + 127.5);
                p[1] = clamp8(ny * inv * 127.5 + 127.5);
                p[2] = clamp8(nz * inv * 127.5 + 127.5);
                p[3] = 255;
            }
        return CRN2_OK;
    }
    double dirx[3][4], diry[3][4], freq[3][4], phase[3][4];
    for (int c = 0; c < 3; c++)
        for (int g = 0; g < 4; g++) {
            double ax = rng.uniform() * 2.0 - 1.0, ay = rng.uniform() * 2.0 - 1.0;
            double n = std::sqrt(ax * ax + ay * ay) + 1e-9;
            dirx[c][g] = ax / n;
            diry[c][g] = ay / n;
            freq[c][g] = 0.5 + 1.5 * rng.uniform();
            phase[c][g] = rng.uniform();
        }
    uint32_t nseed[3];
    for (int c = 0; c < 3; c++) nseed[c] = (uint32_t)rng.next();
    const double cx = rng.uniform() * w, cy = rng.uniform() * h;
    const double rad = (0.35 + 0.5 * rng.uniform()) * (double)std::max(w, h);
    const double span = (double)std::max(w, h);
    for (uint32_t y = 0; y < h; y++)
        for (uint32_t x = 0; x < w; x++) {
            uint8_t* p = out + 4 * ((size_t)y * w + x);
            for (int c = 0; c < 3; c++) {
                double v = 0.0;
                for (int g = 0; g < 4; g++)
                    v += (255.0 / 4.0) * tri(freq[c][g] * ((double)x * dirx[c][g] + (double)y * diry[c][g]) / span + phase[c][g]);
                v += 40.0 * value_noise(x, y, 64.0, nseed[c]);
                v += 20.0 * value_noise(x, y, 32.0, nseed[c] + 1);
                v += 10.0 * value_noise(x, y, 16.0, nseed[c] + 2);
                p[c] = clamp8(v);
            }
            if (kind == 1) {
                double dx = (double)x - cx, dy = (double)y - cy;
                double d = std::sqrt(dx * dx + dy * dy) / rad;
                p[3] = clamp8(255.0 * (1.0 - std::min(d, 1.0)));
            } else {
                p[3] = 255;
            }
        }
    return CRN2_OK;
}

// 2x2 box filter with rounding; odd edges clamp.
extern "C" int crn2_downsample(const uint8_t* src, uint32_t w, uint32_t h, uint8_t* dst)
{
    if (!src || !dst || !w || !h) return CRN2_EINVAL;
    const uint32_t nw = std::max<uint32_t>(1, w >> 1), nh = std::max<uint32_t>(1, h >> 1);
    for (uint32_t y = 0; y < nh; y++)
        for (uint32_t x = 0; x < nw; x++) {
            uint32_t x0 = std::min(2 * x, w - 1), x1 = std::min(2 * x + 1, w - 1);
            uint32_t y0 = std::min(2 * y, h - 1), y1 = std::min(2 * y + 1, h - 1);
            for (int c = 0; c < 4; c++) {
                uint32_t s = src[4 * ((size_t)y0 * w + x0) + c] + src[4 * ((size_t)y0 * w + x1) + c] +
                             src[4 * ((size_t)y1 * w + x0) + c] + src[4 * ((size_t)y1 * w + x1) + c];
                dst[4 * ((size_t)y * nw + x) + c] = (uint8_t)((s + 2) >> 2);
            }
        }
    return CRN2_OK;
}

// ---------------------------------------------------------------------------
// Simple deterministic block quantizer (input production for the transcoder).
namespace {

struct Px { int r, g, b, a; };

static inline uint32_t pack565(int r, int g, int b)
{
    r = std::min(31, std::max(0, (r * 31 + 127) / 255));
    g = std::min(63, std::max(0, (g * 63 + 127) / 255));
    b = std::min(31, std::max(0, (b * 31 + 127) / 255));
    return (uint32_t)(b | (g << 5) | (r << 11));
}

static inline void unpack565(uint32_t c, int& r, int& g, int& b)
{
    int r5 = (c >> 11) & 31, g6 = (c >> 5) & 63, b5 = c & 31;
    r = (r5 << 3) | (r5 >> 2);
    g = (g6 << 2) | (g6 >> 4);
    b = (b5 << 3) | (b5 >> 2);
}

// gather the 4x4 texels of block (bx, by) with edge clamp
static void gather(const uint8_t* img, uint32_t w, uint32_t h, uint32_t bx, uint32_t by, Px* px)
{
    for (int y = 0; y < 4; y++)
        for (int x = 0; x < 4; x++) {
            uint32_t sx = std::min(bx * 4 + x, w - 1), sy = std::min(by * 4 + y, h - 1);
            const uint8_t* p = img + 4 * ((size_t)sy * w + sx);
            px[y * 4 + x] = {p[0], p[1], p[2], p[3]};
        }
}

// colour endpoints from the extremes of the projection on the luma-weighted
// principal direction (max-spread channel pair), 4-colour order c0 > c1
static uint32_t block_color_endpoints(const Px* px)
{
    int mn[3] = {255, 255, 255}, mx[3] = {0, 0, 0};
    double mean[3] = {0, 0, 0};
    for (int i = 0; i < 16; i++) {
        const int v[3] = {px[i].r, px[i].g, px[i].b};
        for (int c = 0; c < 3; c++) {
            mn[c] = std::min(mn[c], v[c]);
            mx[c] = std::max(mx[c], v[c]);
            mean[c] += v[c];
        }
    }
    for (int c = 0; c < 3; c++) mean[c] /= 16.0;
    double cov[6] = {0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 16; i++) {
        double d[3] = {px[i].r - mean[0], px[i].g - mean[1], px[i].b - mean[2]};
        cov[0] += d[0] * d[0]; cov[1] += d[0] * d[1]; cov[2] += d[0] * d[2];
        cov[3] += d[1] * d[1]; cov[4] += d[1] * d[2]; cov[5] += d[2] * d[2];
    }
    double ax[3] = {(double)(mx[0] - mn[0]), (double)(mx[1] - mn[1]), (double)(mx[2] - mn[2])};
    for (int it = 0; it < 4; it++) {
        double n[3] = {cov[0] * ax[0] + cov[1] * ax[1] + cov[2] * ax[2], cov[1] * ax[0] + cov[3] * ax[1] + cov[4] * ax[2],
                       cov[2] * ax[0] + cov[4] * ax[1] + cov[5] * ax[2]};
        double m = std::max(std::fabs(n[0]), std::max(std::fabs(n[1]), std::fabs(n[2])));
        if (m < 1e-9) break;
        for (int c = 0; c < 3; c++) ax[c] = n[c] / m;
    }
    double lo = 1e30, hi = -1e30;
    int ilo = 0, ihi = 0;
    for (int i = 0; i < 16; i++) {
        double t = (px[i].r - mean[0]) * ax[0] + (px[i].g - mean[1]) * ax[1] + (px[i].b - mean[2]) * ax[2];
        if (t < lo) { lo = t; ilo = i; }
        if (t > hi) { hi = t; ihi = i; }
    }
    uint32_t c0 = pack565(px[ihi].r, px[ihi].g, px[ihi].b), c1 = pack565(px[ilo].r, px[ilo].g, px[ilo].b);
    if (c0 < c1) std::swap(c0, c1);
    return c0 | (c1 << 16);
}

static uint32_t best_color_selectors(const Px* px, uint32_t ep)
{
    int r0, g0, b0, r1, g1, b1;
    unpack565(ep & 0xFFFF, r0, g0, b0);
    unpack565(ep >> 16, r1, g1, b1);
    int pal[4][3] = {{r0, g0, b0}, {(2 * r0 + r1) / 3, (2 * g0 + g1) / 3, (2 * b0 + b1) / 3},
                     {(r0 + 2 * r1) / 3, (g0 + 2 * g1) / 3, (b0 + 2 * b1) / 3}, {r1, g1, b1}};
    uint32_t s = 0;
    for (int i = 0; i < 16; i++) {
        int best = 0, bd = 1 << 30;
        for (int k = 0; k < 4; k++) {
            int dr = px[i].r - pal[k][0], dg = px[i].g - pal[k][1], db = px[i].b - pal[k][2];
            int d = dr * dr + dg * dg + db * db;
            if (d < bd) { bd = d; best = k; }
        }
        s |= (uint32_t)best << (2 * i);
    }
    return s;
}

static uint32_t block_alpha_endpoints(const Px* px, int comp)
{
    int mn = 255, mx = 0;
    for (int i = 0; i < 16; i++) {
        int v = comp == 0 ? px[i].r : comp == 1 ? px[i].g : comp == 2 ? px[i].b : px[i].a;
        mn = std::min(mn, v);
        mx = std::max(mx, v);
    }
    return (uint32_t)mx | ((uint32_t)mn << 8); // a0 >= a1: 8-value mode
}

static uint64_t best_alpha_selectors(const Px* px, int comp, uint32_t ep)
{
    int a0 = ep & 255, a1 = (ep >> 8) & 255;
    int vals[8];
    // linear order 0 = a0 .. 7 = a1 (8-value interpolation)
    for (int k = 0; k < 8; k++) vals[k] = (a0 * (7 - k) + a1 * k) / 7;
    uint64_t s = 0;
    for (int i = 0; i < 16; i++) {
        int v = comp == 0 ? px[i].r : comp == 1 ? px[i].g : comp == 2 ? px[i].b : px[i].a;
        int best = 0, bd = 1 << 30;
        for (int k = 0; k < 8; k++) {
            int d = (v - vals[k]) * (v - vals[k]);
            if (d < bd) { bd = d; best = k; }
        }
        s |= (uint64_t)best << (3 * i);
    }
    return s;
}

// Reduce a multiset of keys to at most `target` codewords: the most frequent
// keys become the codebook and every other key maps to its nearest codeword
// under `dist`.  Returns codebook and the key -> codeword map.
template <typename K, typename D>
static void vq_top(const std::vector<K>& keys, uint32_t target, D dist, std::vector<K>& book,
                   std::unordered_map<K, uint32_t>& map)
{
    std::unordered_map<K, uint32_t> freq;
    freq.reserve(keys.size());
    for (const K& k : keys) freq[k]++;
    std::vector<std::pair<K, uint32_t>> u(freq.begin(), freq.end());
    std::sort(u.begin(), u.end(), [](const std::pair<K, uint32_t>& a, const std::pair<K, uint32_t>& b) {
        return a.second != b.second ? a.second > b.second : a.first < b.first;
    });
    const size_t n = std::min<size_t>(u.size(), target);
    book.clear();
    map.clear();
    map.reserve(u.size());
    for (size_t i = 0; i < n; i++) {
        book.push_back(u[i].first);
        map[u[i].first] = (uint32_t)i;
    }
    std::vector<K> rest;
    for (size_t i = n; i < u.size(); i++) rest.push_back(u[i].first);
    std::vector<uint32_t> best(rest.size(), 0);
#pragma omp parallel for schedule(dynamic, 64)
    for (long i = 0; i < (long)rest.size(); i++) {
        uint64_t bd = ~0ull;
        uint32_t bi = 0;
        for (size_t j = 0; j < n; j++) {
            uint64_t d = dist(rest[i], book[j], bd);
            if (d < bd) { bd = d; bi = (uint32_t)j; }
        }
        best[i] = bi;
    }
    for (size_t i = 0; i < rest.size(); i++) map[rest[i]] = best[i];
}

static uint64_t ep_dist(uint32_t a, uint32_t b, uint64_t)
{
    uint64_t d = 0;
    for (int e = 0; e < 2; e++) {
        int ra, ga, ba, rb, gb, bb;
        unpack565((a >> (16 * e)) & 0xFFFF, ra, ga, ba);
        unpack565((b >> (16 * e)) & 0xFFFF, rb, gb, bb);
        d += (uint64_t)((ra - rb) * (ra - rb) + (ga - gb) * (ga - gb) + (ba - bb) * (ba - bb));
    }
    return d;
}

static uint64_t csel_dist(uint32_t a, uint32_t b, uint64_t bound)
{
    uint64_t d = 0;
    for (int i = 0; i < 16 && d < bound; i++) {
        int x = (a >> (2 * i)) & 3, y = (b >> (2 * i)) & 3;
        d += (uint64_t)((x - y) * (x - y));
    }
    return d;
}

static uint64_t aep_dist(uint32_t a, uint32_t b, uint64_t)
{
    int a0 = a & 255, a1 = (a >> 8) & 255, b0 = b & 255, b1 = (b >> 8) & 255;
    return (uint64_t)((a0 - b0) * (a0 - b0) + (a1 - b1) * (a1 - b1));
}

static uint64_t asel_dist(uint64_t a, uint64_t b, uint64_t bound)
{
    uint64_t d = 0;
    for (int i = 0; i < 16 && d < bound; i++) {
        int x = (int)((a >> (3 * i)) & 7), y = (int)((b >> (3 * i)) & 7);
        d += (uint64_t)((x - y) * (x - y));
    }
    return d;
}

}  // namespace

// Compresses RGBA8 levels (levels[f * nlevels + l], each max(1, w >> l) x
// max(1, h >> l)) to a .crn with palettes of at most the given sizes.
extern "C" int crn2_compress_rgba(const uint8_t* const* levels, uint32_t width, uint32_t height, uint32_t nlevels,
                                  uint32_t faces, uint32_t format, const uint32_t* palette_sizes /* ce, cs, ae, as */,
                                  uint8_t** out, uint32_t* out_size)
{
    if (!levels || !out || !out_size || !palette_sizes || !nlevels || nlevels > 16 || (faces != 1 && faces != 6))
        return CRN2_EINVAL;
    const bool has_color = format == CRN2_FMT_DXT1 || format == CRN2_FMT_DXT5;
    int ncomp_alpha = 0, acomp[2] = {3, 3};
    if (format == CRN2_FMT_DXT5 || format == CRN2_FMT_DXT5A) { ncomp_alpha = 1; acomp[0] = 3; }
    else if (format == CRN2_FMT_DXN_XY) { ncomp_alpha = 2; acomp[0] = 0; acomp[1] = 1; }
    else if (format == CRN2_FMT_DXN_YX) { ncomp_alpha = 2; acomp[0] = 1; acomp[1] = 0; }
    else if (format != CRN2_FMT_DXT1) return CRN2_EUNSUPPORTED;
    if (format == CRN2_FMT_DXT5A) acomp[0] = 3;

    // block list in pack order
    std::vector<uint32_t> ce_key, ae_key[2];
    std::vector<Px> pixels;
    for (uint32_t l = 0; l < nlevels; l++) {
        uint32_t w = std::max<uint32_t>(1, width >> l), h = std::max<uint32_t>(1, height >> l);
        uint32_t bw = (w + 3) >> 2, bh = (h + 3) >> 2;
        uint32_t W = (bw + 1) & ~1u, H = (bh + 1) & ~1u;
        for (uint32_t f = 0; f < faces; f++) {
            const uint8_t* img = levels[f * nlevels + l];
            for (uint32_t y = 0; y < H; y++)
                for (uint32_t x = 0; x < W; x++) {
                    Px px[16];
                    gather(img, w, h, std::min(x, bw - 1), std::min(y, bh - 1), px);
                    pixels.insert(pixels.end(), px, px + 16);
                    if (has_color) ce_key.push_back(block_color_endpoints(px));
                    for (int a = 0; a < ncomp_alpha; a++) ae_key[a].push_back(block_alpha_endpoints(px, acomp[a]));
                }
        }
    }
    const size_t nb = pixels.size() / 16;
    std::vector<uint32_t> ce_idx(nb, 0), cs_idx(nb, 0), a_idx[2], as_idx[2];
    std::vector<uint32_t> ce_book, cs_book, ae_book;
    std::vector<uint64_t> as_book;
    if (has_color) {
        std::unordered_map<uint32_t, uint32_t> m;
        vq_top(ce_key, palette_sizes[0], ep_dist, ce_book, m);
        std::vector<uint32_t> sel(nb);
        for (size_t b = 0; b < nb; b++) {
            ce_idx[b] = m[ce_key[b]];
            sel[b] = best_color_selectors(&pixels[16 * b], ce_book[ce_idx[b]]);
        }
        std::unordered_map<uint32_t, uint32_t> ms;
        vq_top(sel, palette_sizes[1], csel_dist, cs_book, ms);
        for (size_t b = 0; b < nb; b++) cs_idx[b] = ms[sel[b]];
    }
    if (ncomp_alpha) {
        std::vector<uint32_t> keys;
        for (int a = 0; a < ncomp_alpha; a++) keys.insert(keys.end(), ae_key[a].begin(), ae_key[a].end());
        std::unordered_map<uint32_t, uint32_t> m;
        vq_top(keys, palette_sizes[2], aep_dist, ae_book, m);
        std::vector<uint64_t> sel;
        for (int a = 0; a < ncomp_alpha; a++) {
            a_idx[a].resize(nb);
            for (size_t b = 0; b < nb; b++) {
                a_idx[a][b] = m[ae_key[a][b]];
                sel.push_back(best_alpha_selectors(&pixels[16 * b], acomp[a], ae_book[a_idx[a][b]]));
            }
        }
        std::unordered_map<uint64_t, uint32_t> ms;
        vq_top(sel, palette_sizes[3], asel_dist, as_book, ms);
        for (int a = 0; a < ncomp_alpha; a++) {
            as_idx[a].resize(nb);
            for (size_t b = 0; b < nb; b++) as_idx[a][b] = ms[sel[a * nb + b]];
        }
    }
    // order colour endpoints by luma of c0 so palette deltas stay small
    if (has_color) {
        std::vector<uint32_t> order(ce_book.size());
        std::iota(order.begin(), order.end(), 0u);
        auto luma = [&](uint32_t e) {
            int r, g, b;
            unpack565(e & 0xFFFF, r, g, b);
            return 2 * r + 5 * g + b;
        };
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return luma(ce_book[a]) < luma(ce_book[b]); });
        std::vector<uint32_t> inv(order.size()), nbk(order.size());
        for (size_t i = 0; i < order.size(); i++) { inv[order[i]] = (uint32_t)i; nbk[i] = ce_book[order[i]]; }
        ce_book.swap(nbk);
        for (auto& v : ce_idx) v = inv[v];
    }
    if (ncomp_alpha) {
        std::vector<uint32_t> order(ae_book.size());
        std::iota(order.begin(), order.end(), 0u);
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return (ae_book[a] & 255) < (ae_book[b] & 255); });
        std::vector<uint32_t> inv(order.size()), nbk(order.size());
        for (size_t i = 0; i < order.size(); i++) { inv[order[i]] = (uint32_t)i; nbk[i] = ae_book[order[i]]; }
        ae_book.swap(nbk);
        for (int a = 0; a < ncomp_alpha; a++)
            for (auto& v : a_idx[a]) v = inv[v];
    }
    crn2_pack_input in;
    memset(&in, 0, sizeof(in));
    in.width = width;
    in.height = height;
    in.levels = nlevels;
    in.faces = faces;
    in.format = format;
    if (has_color) {
        in.color_endpoints = ce_book.data();
        in.n_color_endpoints = (uint32_t)ce_book.size();
        in.color_selectors_linear = cs_book.data();
        in.n_color_selectors = (uint32_t)cs_book.size();
        in.color_endpoint_idx = ce_idx.data();
        in.color_selector_idx = cs_idx.data();
    }
    if (ncomp_alpha) {
        in.alpha_endpoints = ae_book.data();
        in.n_alpha_endpoints = (uint32_t)ae_book.size();
        in.alpha_selectors_linear = as_book.data();
        in.n_alpha_selectors = (uint32_t)as_book.size();
        in.alpha0_endpoint_idx = a_idx[0].data();
        in.alpha0_selector_idx = as_idx[0].data();
        if (ncomp_alpha == 2) {
            in.alpha1_endpoint_idx = a_idx[1].data();
            in.alpha1_selector_idx = as_idx[1].data();
        }
    }
    return crn2_pack(&in, out, out_size);
}
