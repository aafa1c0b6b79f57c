// mip_host.cpp -- host set-up of GPU mip generation: the resampling filters,
// the contributor lists, the gamma tables, level sizes and the per-level
// order of the two passes.  The arithmetic here is float and double exactly
// as crnlib evaluates it (built with -ffp-contract=off, glibc's libm), so the
// lists and tables are the reference's bit for bit:
//   filters ............ crnlib/crn_resample_filters.cpp:17-369 (box, tent,
//                        lanczos4, mitchell B = C = 1/3, kaiser 40 dB)
//   contributor lists .. Resampler::make_clist, crnlib/crn_resampler.cpp:117-416
//   edge handling ...... Resampler::reflect (clamp / wrap), :61-113
//   pass order ......... the Resampler constructor, :893-930
//   gamma tables ....... image_utils::resample_single_thread,
//                        crnlib/crn_image_utils.cpp:514-545
//   level sizes ........ mipmapped_texture::generate_mipmaps,
//                        crnlib/crn_mipmapped_texture.cpp:2145-2177
#include <math.h>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "mip_plan.h"

namespace crn2mip {
namespace {

double sinc(double x)
{
    x = x * M_PI;
    if (x < 0.01f && x > -0.01f) return 1.0f + x * x * (-1.0f / 6.0f + x * x * 1.0f / 120.0f);
    return std::sin(x) / x;
}

// values of magnitude below 1.25e-5 become exact zeros (dropped from lists)
float clean(double t)
{
    const float eps = .0000125f;
    return std::fabs(t) < eps ? 0.0f : static_cast<float>(t);
}

// modified Bessel function of the first kind, order 0, by its power series
double bessel_i0(double x)
{
    const double half = 0.5 * x;
    double sum = 1.0, term = 1.0, sq = 1.0;
    for (int k = 1; sq > sum * 1E-16; k++) {
        term = term * (half / k);
        sq = term * term;
        sum = sum + sq;
    }
    return sum;
}

float mitchell(float t, const float B, const float C)
{
    const float t2 = t * t;
    t = std::fabs(t);
    if (t < 1.0f) {
        const float v = ((12.0f - 9.0f * B - 6.0f * C) * (t * t2)) + ((-18.0f + 12.0f * B + 6.0f * C) * t2) +
                        (6.0f - 2.0f * B);
        return v / 6.0f;
    }
    if (t < 2.0f) {
        const float v = ((-1.0f * B - 6.0f * C) * (t * t2)) + ((6.0f * B + 30.0f * C) * t2) +
                        ((-12.0f * B - 48.0f * C) * t) + (8.0f * B + 24.0f * C);
        return v / 6.0f;
    }
    return 0.0f;
}

float kaiser(float t)
{
    t = std::fabs(t);
    if (!(t < 3)) return 0.0f;
    // window shape for 40 dB attenuation: Kaiser's empirical formula as crnlib
    // writes it, with the 0.58417 inside the power (beta = 4.124)
    const float att = 40.0f;
    const float beta = (float)(std::exp(std::log((double)0.58417 * (att - 20.96)) * 0.4) + 0.07886 * (att - 20.96));
    const double r = (double)t / 3.0;
    const double window = bessel_i0(beta * std::sqrt(1 - r * r)) / bessel_i0(beta);
    return clean(sinc(t) * window);
}

uint32_t reflect(int j, uint32_t src, bool wrap)
{
    const int n = (int)src;
    if (j >= 0 && j < n) return (uint32_t)j;
    if (wrap) {
        int m = j % n;
        return (uint32_t)(m < 0 ? m + n : m);
    }
    return j < 0 ? 0u : src - 1;
}

}  // namespace

float filter_support(uint32_t filter)
{
    static const float support[5] = {0.5f, 1.0f, 4.0f, 2.0f, 3.0f};
    return filter < 5 ? support[filter] : 0.0f;
}

float filter_value(uint32_t filter, float t)
{
    switch (filter) {
        case CRN2_MIP_BOX: return (t >= -0.5f && t < 0.5f) ? 1.0f : 0.0f;
        case CRN2_MIP_TENT: t = std::fabs(t); return t < 1.0f ? 1.0f - t : 0.0f;
        case CRN2_MIP_LANCZOS4: t = std::fabs(t); return t < 4.0f ? clean(sinc(t) * sinc(t / 4.0f)) : 0.0f;
        case CRN2_MIP_MITCHELL: return mitchell(t, 1.0f / 3.0f, 1.0f / 3.0f);
        case CRN2_MIP_KAISER: return kaiser(t);
    }
    return 0.0f;
}

bool make_list(uint32_t src, uint32_t dst, bool wrap, uint32_t filter, float filter_scale, List& L)
{
    L.offs.assign(1, 0);
    L.pix.clear();
    L.wt.clear();
    if (!src || !dst || filter > CRN2_MIP_KAISER || !(filter_scale > 0.0f)) return false;
    const float support = filter_support(filter);
    const float inv_fs = 1.0f / filter_scale;
    const float scale = dst / (float)src;
    const bool minify = scale < 1.0f;
    // half width of the filter in source samples
    const float half = minify ? (support / scale) * filter_scale : support * filter_scale;
    auto arg = [&](float center, int j) {
        return minify ? (center - (float)j) * scale * inv_fs : (center - (float)j) * inv_fs;
    };
    for (uint32_t i = 0; i < dst; i++) {
        // centre of destination sample i in source coordinates
        float center = ((float)i + 0.5f) / scale;
        center -= 0.5f;
        const int left = (int)std::floor(center - half);
        const int right = (int)std::ceil(center + half);
        float sum = 0;
        for (int j = left; j <= right; j++) sum += filter_value(filter, arg(center, j));
        const float norm = 1.0f / sum;
        sum = 0;
        size_t top = SIZE_MAX;
        float top_w = -1e+20f;
        for (int j = left; j <= right; j++) {
            const float wj = filter_value(filter, arg(center, j)) * norm;
            if (wj == 0.0f) continue;
            if (wj > top_w) {
                top_w = wj;
                top = L.pix.size();
            }
            L.pix.push_back(reflect(j, src, wrap));
            L.wt.push_back(wj);
            sum += wj;
        }
        // no contributor, or none above -1e20 (a zero filter sum gives NaNs)
        if (top == SIZE_MAX) return false;
        // the largest weight absorbs the rounding of the normalisation
        if (sum != 1.0f) L.wt[top] += 1.0f - sum;
        L.offs.push_back((uint32_t)L.pix.size());
    }
    return true;
}

void gamma_tables(float gamma, float* s2l, uint8_t* l2s)
{
    for (int i = 0; i < 256; ++i) s2l[i] = std::pow(i * 1.0f / 255.0f, gamma);
    const float step = 1.0f / 8192;
    const float inv_gamma = 1.0f / gamma;
    for (int i = 0; i < 8192; ++i) {
        const int k = (int)(255.0f * std::pow(i * step, inv_gamma) + .5f);
        l2s[i] = (uint8_t)std::min(255, std::max(0, k));
    }
}

uint32_t level_count(uint32_t w, uint32_t h, uint32_t min_size, uint32_t max_levels)
{
    uint32_t n = 1;
    for (; w > min_size || h > min_size; n++) {
        w >>= 1;
        h >>= 1;
    }
    return (max_levels && n > max_levels) ? max_levels : n;
}

bool plan_levels(uint32_t width, uint32_t height, const crn2_mip_params& p, std::vector<Level>& out)
{
    out.clear();
    if (!width || !height || width > 16384 || height > 16384 || p.filter > CRN2_MIP_KAISER ||
        !(p.filter_scale > 0.0f) || !(p.source_gamma > 0.0f))
        return false;
    const uint32_t n = level_count(width, height, p.min_mip_size, p.max_levels);
    const bool wrap = (p.flags & CRN2_MIP_WRAP) != 0;
    for (uint32_t l = 1; l < n; l++) {
        Level v;
        v.width = std::max<uint32_t>(1, width >> l);
        v.height = std::max<uint32_t>(1, height >> l);
        if (!make_list(width, v.width, wrap, p.filter, p.filter_scale, v.lx) ||
            !make_list(height, v.height, wrap, p.filter, p.filter_scale, v.ly))
            return false;
        v.xfirst = true;
        if (!(p.flags & CRN2_MIP_THREADED)) {
            // multiplies of either order; Y work weighs 4/3 (int arithmetic,
            // 32-bit wrap-around as in the reference)
            const uint32_t xo = v.lx.ops(), yo = v.ly.ops();
            const int xy = (int)((uint32_t)(int)(xo * height) + (uint32_t)((int)(4u * yo * v.width) / 3));
            const int yx = (int)((uint32_t)((int)(4u * yo * width) / 3) + (uint32_t)(int)(xo * v.height));
            v.xfirst = !(xy > yx || (xy == yx && width < v.width));
        }
        out.push_back(std::move(v));
    }
    return true;
}

}  // namespace crn2mip
