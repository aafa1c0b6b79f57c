/*
 * mip_oracle.c -- CPU restatement of crnlib's mipmap generation.
 *
 * TEST INFRASTRUCTURE ONLY (see crnd_oracle.c for the rule).
 *
 * Follows:
 *   - mipmapped_texture::generate_mipmaps (crnlib/crn_mipmapped_texture.cpp:
 *     2140-2220): the level count (halve both sides while either is above the
 *     minimum size, capped by max_mips when non-zero), level l of
 *     max(1, w >> l) x max(1, h >> l) resampled from level 0, 4 components
 *     when alpha is valid else 3 (alpha of the result stays 255), optional
 *     renorm_normal_map per generated level;
 *   - image_utils::resample_single_thread (crnlib/crn_image_utils.cpp:478-
 *     660): the gamma tables (powf of float arguments: the float overload of
 *     pow that C++ picks), colour channels through srgb_to_linear when sRGB,
 *     alpha always v * (1/255), output (int)(255 x + .5) or the 8,192-entry
 *     linear_to_srgb table, clamped;
 *   - image_utils::resample_multithreaded / threaded_resampler
 *     (crn_image_utils.cpp:668-865, crnlib/crn_threaded_resampler.cpp): the
 *     same tables, X pass over every source row then Y pass, clamp in the Y
 *     pass, a one-contributor Y list copies the X result;
 *   - Resampler (crnlib/crn_resampler.cpp): make_clist (:117-416) for
 *     minification and magnification, reflect() for clamp and wrap (:61-113),
 *     resample_x (:418-441: total = 0, += src * weight in list order),
 *     resample_y (:471-537: first contributor scale_y_mov, then scale_y_add;
 *     delayed X after Y; clamp), the order choice of the constructor
 *     (:893-930, int arithmetic as the reference's);
 *   - the filters box, tent, lanczos4, mitchell and kaiser with sinc, clean
 *     and bessel0 (crnlib/crn_resample_filters.cpp:17-369).
 * Float expressions are evaluated in the reference's order with
 * -ffp-contract=off; the libm calls are glibc's, as in the reference's Linux
 * build.
 *
 * Parity pins (tests/test_mip_oracle.py): the filters against numpy's sinc,
 * i0 and the closed forms of the tent and Mitchell kernels; box and tent
 * contributor lists against their closed forms ([1 1]/2, [1 1 1 1]/4 and
 * [1 3 3 1]/8 with clamped edges); whole box and tent levels against a numpy
 * float32 evaluation from those closed-form weights; constant images stay
 * constant under every filter; a circular shift by 2^l texels shifts level l
 * by one texel in wrap mode; the gamma tables against numpy; level counts;
 * renormalized texels are unit length or the flat normal.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define ORC_MIP_SRGB 1
#define ORC_MIP_WRAP 2
#define ORC_MIP_ALPHA 4
#define ORC_MIP_RENORMALIZE 8
#define ORC_MIP_THREADED 16

/* ------------------------------------------------------------- filters */

static double orc_sinc(double x)
{
    x = x * M_PI;
    if (x < 0.01f && x > -0.01f)
        return 1.0f + x * x * ((double)(-1.0f / 6.0f) + x * x * 1.0f / 120.0f);
    return sin(x) / x;
}

static float orc_clean(double t)
{
    const float eps = .0000125f;
    if (fabs(t) < eps) return 0.0f;
    return (float)t;
}

static double orc_bessel0(double x)
{
    double xh = 0.5 * x, sum = 1.0, pw = 1.0, ds = 1.0;
    int k = 0;
    while (ds > sum * 1E-16) {
        ++k;
        pw = pw * (xh / k);
        ds = pw * pw;
        sum = sum + ds;
    }
    return sum;
}

static float orc_mitchell(float t, const float B, const float C)
{
    float tt = t * t;
    if (t < 0.0f) t = -t;
    if (t < 1.0f) {
        t = (((12.0f - 9.0f * B - 6.0f * C) * (t * tt)) + ((-18.0f + 12.0f * B + 6.0f * C) * tt) + (6.0f - 2.0f * B));
        return t / 6.0f;
    }
    if (t < 2.0f) {
        t = (((-1.0f * B - 6.0f * C) * (t * tt)) + ((6.0f * B + 30.0f * C) * tt) + ((-12.0f * B - 48.0f * C) * t) +
             (8.0f * B + 24.0f * C));
        return t / 6.0f;
    }
    return 0.0f;
}

/* filter 0 box, 1 tent, 2 lanczos4, 3 mitchell, 4 kaiser (crn_mip_filter) */
float orc_mip_filter(int filter, float t)
{
    switch (filter) {
        case 0:
            return (t >= -0.5f && t < 0.5f) ? 1.0f : 0.0f;
        case 1:
            if (t < 0.0f) t = -t;
            return t < 1.0f ? 1.0f - t : 0.0f;
        case 2:
            if (t < 0.0f) t = -t;
            return t < 4.0f ? orc_clean(orc_sinc(t) * orc_sinc(t / 4.0f)) : 0.0f;
        case 3:
            return orc_mitchell(t, 1.0f / 3.0f, 1.0f / 3.0f);
        case 4:
            if (t < 0.0f) t = -t;
            if (t < 3) {
                const float att = 40.0f;
                const float alpha = (float)(exp(log((double)0.58417 * (att - 20.96)) * 0.4) + 0.07886 * (att - 20.96));
                const double ratio = (double)t / 3.0;
                const double kw = orc_bessel0((double)alpha * sqrt(1 - ratio * ratio)) / orc_bessel0(alpha);
                return (float)orc_clean(orc_sinc(t) * kw);
            }
            return 0.0f;
    }
    return 0.0f;
}

float orc_mip_support(int filter)
{
    static const float s[5] = {0.5f, 1.0f, 4.0f, 2.0f, 3.0f};
    return (filter >= 0 && filter < 5) ? s[filter] : 0.0f;
}

/* ----------------------------------------------------- contributor lists */

typedef struct {
    int n;
    int* pix;
    float* w;
} orc_clist;

static int orc_reflect(int j, int src, int wrap)
{
    if (j < 0) {
        if (wrap) { int m = j % src; return m < 0 ? m + src : m; }
        return 0;
    }
    if (j >= src) {
        if (wrap) return j % src;
        return src - 1;
    }
    return j;
}

/* The contributors of every destination sample; NULL on failure.  The pool
 * of all lists is lists[0].pix / lists[0].w. */
static orc_clist* orc_make_clist(int src, int dst, int wrap, int filter, float filter_scale)
{
    const float support = orc_mip_support(filter);
    const float oo_scale = 1.0f / filter_scale;
    const float xscale = dst / (float)src;
    const int down = xscale < 1.0f;
    const float half_width = down ? (support / xscale) * filter_scale : support * filter_scale;
    orc_clist* L = (orc_clist*)calloc((size_t)dst, sizeof(orc_clist));
    float* centers = (float*)malloc(sizeof(float) * (size_t)dst);
    int* lefts = (int*)malloc(sizeof(int) * (size_t)dst);
    int* rights = (int*)malloc(sizeof(int) * (size_t)dst);
    long total = 0;
    for (int i = 0; i < dst; i++) {
        float center = ((float)i + 0.5f) / xscale;
        center -= 0.5f;
        center += 0.0f;
        centers[i] = center;
        lefts[i] = (int)floorf(center - half_width);
        rights[i] = (int)ceilf(center + half_width);
        total += rights[i] - lefts[i] + 1;
    }
    int* pool_p = (int*)malloc(sizeof(int) * (size_t)(total > 0 ? total : 1));
    float* pool_w = (float*)malloc(sizeof(float) * (size_t)(total > 0 ? total : 1));
    long next = 0;
    int ok = total > 0;
    for (int i = 0; i < dst && ok; i++) {
        const float center = centers[i];
        L[i].pix = pool_p + next;
        L[i].w = pool_w + next;
        L[i].n = 0;
        next += rights[i] - lefts[i] + 1;
        float tw = 0;
        for (int j = lefts[i]; j <= rights[i]; j++) {
            const float arg = down ? (center - (float)j) * xscale * oo_scale : (center - (float)j) * oo_scale;
            tw += orc_mip_filter(filter, arg);
        }
        const float norm = 1.0f / tw;
        tw = 0;
        int max_k = -1;
        float max_w = -1e+20f;
        for (int j = lefts[i]; j <= rights[i]; j++) {
            const float arg = down ? (center - (float)j) * xscale * oo_scale : (center - (float)j) * oo_scale;
            const float w = orc_mip_filter(filter, arg) * norm;
            if (w == 0.0f) continue;
            const int k = L[i].n++;
            L[i].pix[k] = orc_reflect(j, src, wrap);
            L[i].w[k] = w;
            tw += w;
            if (w > max_w) { max_w = w; max_k = k; }
        }
        if (max_k == -1 || L[i].n == 0) { ok = 0; break; }
        if (tw != 1.0f) L[i].w[max_k] += 1.0f - tw;
    }
    free(centers); free(lefts); free(rights);
    if (!ok) { free(pool_p); free(pool_w); free(L); return NULL; }
    if (dst > 0) { L[0].pix = pool_p; L[0].w = pool_w; }
    return L;
}

static void orc_free_clist(orc_clist* L)
{
    if (!L) return;
    free(L[0].pix);
    free(L[0].w);
    free(L);
}

/* Flattened contributors for the tests: offs[dst + 1], pix[], w[]; returns
 * the contributor count (or -1), writing at most cap entries. */
int orc_mip_contributors(int src, int dst, int wrap, int filter, float filter_scale, uint32_t* offs, uint32_t* pix,
                         float* w, int cap)
{
    orc_clist* L = orc_make_clist(src, dst, wrap, filter, filter_scale);
    if (!L) return -1;
    int n = 0;
    for (int i = 0; i < dst; i++) {
        if (offs) offs[i] = (uint32_t)n;
        for (int k = 0; k < L[i].n; k++, n++)
            if (n < cap) { pix[n] = (uint32_t)L[i].pix[k]; w[n] = L[i].w[k]; }
    }
    if (offs) offs[dst] = (uint32_t)n;
    orc_free_clist(L);
    return n;
}

/* The Resampler constructor's order: 1 = X after Y (delay_x_resample). */
static int orc_delay_x(const orc_clist* lx, const orc_clist* ly, int src_x, int src_y, int dst_x, int dst_y)
{
    uint32_t x_ops = 0, y_ops = 0;
    for (int i = 0; i < dst_x; i++) x_ops += (uint32_t)lx[i].n;
    for (int i = 0; i < dst_y; i++) y_ops += (uint32_t)ly[i].n;
    /* int arithmetic of the reference; products and sums wrap as 32-bit two's
     * complement (they overflow only for sides near 16,384) */
    const int xy_ops = (int)((uint32_t)(int)(x_ops * (uint32_t)src_y) +
                             (uint32_t)((int)(4u * y_ops * (uint32_t)dst_x) / 3));
    const int yx_ops = (int)((uint32_t)((int)(4u * y_ops * (uint32_t)src_x) / 3) +
                             (uint32_t)(int)(x_ops * (uint32_t)dst_y));
    return (xy_ops > yx_ops) || (xy_ops == yx_ops && src_x < dst_x);
}

/* --------------------------------------------------------------- tables */

void orc_mip_tables(float gamma, float* s2l, uint8_t* l2s)
{
    for (int i = 0; i < 256; ++i) s2l[i] = powf(i * 1.0f / 255.0f, gamma);
    const float inv_size = 1.0f / 8192;
    const float inv_gamma = 1.0f / gamma;
    for (int i = 0; i < 8192; ++i) {
        int k = (int)(255.0f * powf(i * inv_size, inv_gamma) + .5f);
        if (k < 0) k = 0; else if (k > 255) k = 255;
        l2s[i] = (uint8_t)k;
    }
}

static float orc_clamp01(float x)
{
    if (x < 0.0f) x = 0.0f;
    else if (x > 1.0f) x = 1.0f;
    return x;
}

static uint8_t orc_to_u8(float v, int linear, const uint8_t* l2s)
{
    if (linear) {
        int c = (int)(255.0f * v + .5f);
        if (c < 0) c = 0; else if (c > 255) c = 255;
        return (uint8_t)c;
    }
    int j = (int)(8192 * v + .5f);
    if (j < 0) j = 0; else if (j >= 8192) j = 8191;
    return l2s[j];
}

/* --------------------------------------------------------- resampling */

static void orc_resample_x(float* dst, const float* src, const orc_clist* L, int dst_x)
{
    for (int i = 0; i < dst_x; i++) {
        float total = 0;
        for (int k = 0; k < L[i].n; k++) total += src[L[i].pix[k]] * L[i].w[k];
        dst[i] = total;
    }
}

/* One component plane (src_x * src_y floats) -> dst_x * dst_y floats, as the
 * single-thread Resampler produces them (clamped to [0, 1]). */
static void orc_resample_plane(const float* src, int sx, int sy, float* dst, int dx, int dy, const orc_clist* lx,
                               const orc_clist* ly, int delay_x)
{
    if (!delay_x) {
        /* X-Y order: every source row through resample_x, then the Y lists */
        float* rows = (float*)malloc(sizeof(float) * (size_t)dx * (size_t)sy);
        for (int y = 0; y < sy; y++) orc_resample_x(rows + (size_t)y * dx, src + (size_t)y * sx, lx, dx);
        for (int y = 0; y < dy; y++) {
            float* out = dst + (size_t)y * dx;
            for (int k = 0; k < ly[y].n; k++) {
                const float* line = rows + (size_t)ly[y].pix[k] * dx;
                const float w = ly[y].w[k];
                if (!k) for (int x = 0; x < dx; x++) out[x] = line[x] * w;
                else for (int x = 0; x < dx; x++) out[x] += line[x] * w;
            }
            for (int x = 0; x < dx; x++) out[x] = orc_clamp01(out[x]);
        }
        free(rows);
    } else {
        /* Y-X order: raw source rows combined by the Y lists, then resample_x */
        float* tmp = (float*)malloc(sizeof(float) * (size_t)sx);
        for (int y = 0; y < dy; y++) {
            for (int k = 0; k < ly[y].n; k++) {
                const float* line = src + (size_t)ly[y].pix[k] * sx;
                const float w = ly[y].w[k];
                if (!k) for (int x = 0; x < sx; x++) tmp[x] = line[x] * w;
                else for (int x = 0; x < sx; x++) tmp[x] += line[x] * w;
            }
            float* out = dst + (size_t)y * dx;
            orc_resample_x(out, tmp, lx, dx);
            for (int x = 0; x < dx; x++) out[x] = orc_clamp01(out[x]);
        }
        free(tmp);
    }
}

/* threaded_resampler: X pass over all rows, Y pass with the one-contributor
 * copy, clamp. */
static void orc_resample_plane_threaded(const float* src, int sx, int sy, float* dst, int dx, int dy,
                                        const orc_clist* lx, const orc_clist* ly)
{
    float* rows = (float*)malloc(sizeof(float) * (size_t)dx * (size_t)sy);
    float* tmp = (float*)malloc(sizeof(float) * (size_t)dx);
    for (int y = 0; y < sy; y++) orc_resample_x(rows + (size_t)y * dx, src + (size_t)y * sx, lx, dx);
    for (int y = 0; y < dy; y++) {
        const float* p;
        if (ly[y].n == 1) p = rows + (size_t)ly[y].pix[0] * dx;
        else {
            for (int k = 0; k < ly[y].n; k++) {
                const float* line = rows + (size_t)ly[y].pix[k] * dx;
                const float w = ly[y].w[k];
                if (!k) for (int x = 0; x < dx; x++) tmp[x] = line[x] * w;
                else for (int x = 0; x < dx; x++) tmp[x] += line[x] * w;
            }
            p = tmp;
        }
        for (int x = 0; x < dx; x++) dst[(size_t)y * dx + x] = orc_clamp01(p[x]);
    }
    free(rows);
    free(tmp);
}

/* image_utils::renorm_normal_map for one texel (crn_image_utils.cpp:369-420) */
static void orc_renorm(uint8_t* c)
{
    if (c[0] == 128 && c[1] == 128 && c[2] == 128) return;
    float v[3];
    for (int i = 0; i < 3; i++) {
        v[i] = (float)c[i];
        v[i] *= 1.0f / 255.0f;
        v[i] *= 2.0f;
        v[i] -= 1.0f;
        if (v[i] < -1.0f) v[i] = -1.0f; else if (v[i] > 1.0f) v[i] = 1.0f;
    }
    float sum = v[0] * v[0];
    sum += v[1] * v[1];
    sum += v[2] * v[2];
    const float length = sqrtf(sum);
    if (length < .077f) {
        c[0] = c[1] = c[2] = 128;
    } else if (fabsf(length - 1.0f) > .077f) {
        if (length != 0.0f)
            for (int i = 0; i < 3; i++) v[i] /= length;
        for (int i = 0; i < 3; i++) {
            float f = floorf((v[i] + 1.0f) * .5f * 255.0f + .5f);
            if (f < 0.0f) f = 0.0f; else if (f > 255.0f) f = 255.0f;
            c[i] = (uint8_t)f;
        }
        if (c[0] == 128 && c[1] == 128) c[2] = c[2] < 128 ? 0 : 255;
    }
}

uint32_t orc_mip_level_count(uint32_t w, uint32_t h, uint32_t min_size, uint32_t max_levels)
{
    uint32_t n = 1;
    while (w > min_size || h > min_size) {
        w >>= 1U;
        h >>= 1U;
        n++;
    }
    if (max_levels > 0 && n > max_levels) n = max_levels;
    return n;
}

/* Resamples RGBA8 level 0 (w x h) to dst_w x dst_h as resample() does for one
 * generated level; *order receives 1 when X ran after Y. Returns 0 or -1. */
int orc_mip_resample(const uint8_t* rgba, int w, int h, int dst_w, int dst_h, int filter, float filter_scale,
                     float gamma, int flags, uint8_t* out, int* order)
{
    const int wrap = (flags & ORC_MIP_WRAP) != 0, srgb = (flags & ORC_MIP_SRGB) != 0;
    const int ncomp = (flags & ORC_MIP_ALPHA) ? 4 : 3;
    if (order) *order = 0;
    if (w == dst_w && h == dst_h) {          /* dst = src */
        memcpy(out, rgba, (size_t)w * h * 4);
        return 0;
    }
    orc_clist* lx = orc_make_clist(w, dst_w, wrap, filter, filter_scale);
    orc_clist* ly = orc_make_clist(h, dst_h, wrap, filter, filter_scale);
    if (!lx || !ly) { orc_free_clist(lx); orc_free_clist(ly); return -1; }
    const int delay = (flags & ORC_MIP_THREADED) ? 0 : orc_delay_x(lx, ly, w, h, dst_w, dst_h);
    if (order) *order = delay;
    float s2l[256];
    uint8_t* l2s = (uint8_t*)malloc(8192);
    orc_mip_tables(gamma, s2l, l2s);
    float* plane = (float*)malloc(sizeof(float) * (size_t)w * h);
    float* res = (float*)malloc(sizeof(float) * (size_t)dst_w * dst_h);
    for (size_t i = 0; i < (size_t)dst_w * dst_h; i++) {
        out[4 * i + 0] = out[4 * i + 1] = out[4 * i + 2] = 0;
        out[4 * i + 3] = 255;
    }
    for (int c = 0; c < ncomp; c++) {
        const int linear = !srgb || c == 3;
        for (size_t i = 0; i < (size_t)w * h; i++) {
            const uint8_t v = rgba[4 * i + c];
            plane[i] = linear ? v * (1.0f / 255.0f) : s2l[v];
        }
        if (flags & ORC_MIP_THREADED) orc_resample_plane_threaded(plane, w, h, res, dst_w, dst_h, lx, ly);
        else orc_resample_plane(plane, w, h, res, dst_w, dst_h, lx, ly, delay);
        for (size_t i = 0; i < (size_t)dst_w * dst_h; i++) out[4 * i + c] = orc_to_u8(res[i], linear, l2s);
    }
    free(plane); free(res); free(l2s);
    orc_free_clist(lx);
    orc_free_clist(ly);
    return 0;
}

/* The whole chain: levels 0..n-1 concatenated (level 0 copied), n returned
 * in *levels; orders[l] = 1 where X ran after Y. */
int orc_mip_generate(const uint8_t* rgba, uint32_t w, uint32_t h, int filter, float filter_scale, float gamma,
                     int flags, uint32_t max_levels, uint32_t min_size, uint8_t* out, uint32_t* levels, int* orders)
{
    const uint32_t n = orc_mip_level_count(w, h, min_size, max_levels);
    *levels = n;
    memcpy(out, rgba, (size_t)w * h * 4);
    if (orders) orders[0] = 0;
    size_t ofs = (size_t)w * h * 4;
    for (uint32_t l = 1; l < n; l++) {
        const uint32_t mw = (w >> l) > 1 ? (w >> l) : 1, mh = (h >> l) > 1 ? (h >> l) : 1;
        int ord = 0;
        if (orc_mip_resample(rgba, (int)w, (int)h, (int)mw, (int)mh, filter, filter_scale, gamma, flags, out + ofs,
                             &ord))
            return -1;
        if (orders) orders[l] = ord;
        if (flags & ORC_MIP_RENORMALIZE)
            for (size_t i = 0; i < (size_t)mw * mh; i++) orc_renorm(out + ofs + 4 * i);
        ofs += (size_t)mw * mh * 4;
    }
    return 0;
}
