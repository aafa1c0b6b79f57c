/*
 * dxt1_oracle.c -- CPU restatement of crnlib's dxt1_endpoint_optimizer.
 *
 * TEST INFRASTRUCTURE ONLY (see crnd_oracle.c for the rule).
 *
 * Follows crnlib/crn_dxt1.cpp (reference) with the float/double types and
 * operation order of each expression kept, compiled with -ffp-contract=off
 * so that no multiply-add is fused (the reference's x86-64 GCC build has no
 * FMA):
 *   compute_internal ............. :2081-2232 (unique colours, error planes,
 *                                   1-colour shortcut via the ryg tables)
 *   handle_multicolor_block ...... :1069-1178 (perceptual weights, vectors,
 *                                   PCA :192-256, projection bounds, ray clamp
 *                                   crn_intersect.h:44-120)
 *   optimize_endpoints ........... :703-1067 (probe sweeps, lattice, refine)
 *   try_median4 .................. :1181-1308 (k-means, fast_random jitter
 *                                   crn_rand.cpp:310-405)
 *   refine_solution .............. :525-698
 *   evaluate_solution ............ :1312-1340 (plane bound, dedupe)
 *   evaluate_solution_uber ....... :1370-1592
 *   evaluate_solution_hc_* ....... :1759-1835, compute_selectors_hc :1837-1869
 *   try_average_block_as_solid ... :93-153 (ryg tables crn_ryg_dxt.cpp:58-99, 517-535)
 *   optimize_endpoint_comps ...... :369-486
 *   try_combinatorial_encoding ... :1871-1997
 *   return_solution .............. :263-365
 *   evaluate_solution_fast ....... :1594-1757 (qualities below Better:
 *                                   selectors by projection onto the endpoint
 *                                   axis, green weight 24)
 *   optimize_endpoints per quality  :716-746 (probe tables, passes, spacing),
 *                                   :905/:1030 (lattice and solid steps from
 *                                   Normal up)
 *   DXT1A transparency ........... :2116 (pixels below the alpha threshold
 *                                   leave the colour set), :1380/:1604 (3-colour
 *                                   blocks only), :300-343 (selector 3 for them)
 *   try_alpha_as_black_optimization :2001-2079, called from compute :2234-2246
 * Not restated: endpoint caching (parity is defined with
 * cCRNCompFlagDisableEndpointCaching); the solution sort of optimize_endpoints
 * (:798-816) only orders the early exits of the hc evaluators, whose result
 * it cannot change, so it is left out.  A block with no opaque texel leaves
 * the reference's m_error unassigned (:2207-2212); here it is 0.
 *
 * Parity pins (tests/test_dxt1_oracle.py): single-colour blocks reach the
 * optimum of the exhaustive single-colour table search; reported error equals
 * the weighted error recomputed from the emitted endpoints and selectors with
 * the DXT1 palettes; canonical endpoint order and selector inversion; error
 * never below the exhaustive per-channel lower bound.  The search trajectory
 * itself (probe order, k-means, PCA) is fixed only by the reference source:
 * parity unpinned beyond these checks.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { uint8_t r, g, b, a; } c4;
typedef struct { c4 c; uint32_t w; } ucolor;
typedef struct { float v[3]; } v3;
typedef struct { uint16_t lo, hi; } coords;

typedef struct {
    coords c;
    uint8_t* sel;     /* n unique */
    int nsel;         /* 0 = empty selector vector */
    uint64_t err;
    int alpha_block, alt_round, enforce, enforced;
} solution;

typedef struct {
    int quality, perceptual_req, use_alpha_blocks, grayscale, force_alpha_blocks;
    int have_alpha, alpha_thresh;
    /* state */
    int evaluate_hc, perceptual;
    ucolor* uc;
    uint32_t nu, cap;
    uint32_t total_w;
    uint64_t rlo[32], rhi[32], glo[64], ghi[64], blo[32], bhi[32];
    v3 *norm, *normw;
    v3 mean, meanw, axis;
    solution best, trial;
    uint8_t* trial_sel;
    /* hash sets: unique colours, tried solutions */
    uint32_t* ckeys; uint32_t* cvals; uint32_t cmask;
    uint32_t* skeys; uint8_t* sused; uint32_t smask, scount;
    uint64_t ndist;   /* colour distances evaluated (work count) */
} e1;

/* ---------------------------------------------------------------- tables */
static uint8_t om5[256][2], om6[256][2], om5_3[256][2], om6_3[256][2];
static int tables_ready = 0;

static void opt_table(uint8_t (*t)[2], int size, int three)
{
    for (int i = 0; i < 256; i++) {
        int best = 256;
        for (int mn = 0; mn < size; mn++)
            for (int mx = 0; mx < size; mx++) {
                int me = size == 32 ? ((mn << 3) | (mn >> 2)) : ((mn << 2) | (mn >> 4));
                int xe = size == 32 ? ((mx << 3) | (mx >> 2)) : ((mx << 2) | (mx >> 4));
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

static void init_tables(void)
{
    if (tables_ready) return;
    opt_table(om5, 32, 0);
    opt_table(om6, 64, 0);
    opt_table(om5_3, 32, 1);
    opt_table(om6_3, 64, 1);
    tables_ready = 1;
}

/* exposed for tests: which = 0 OMatch5, 1 OMatch6, 2 OMatch5_3, 3 OMatch6_3 */
void orc_dxt1_omatch(int which, uint8_t* out512)
{
    init_tables();
    uint8_t (*t)[2] = which == 0 ? om5 : which == 1 ? om6 : which == 2 ? om5_3 : om6_3;
    memcpy(out512, t, 512);
}

/* ---------------------------------------------------------------- helpers */
static inline c4 unpack565(uint16_t p, int scaled)
{
    uint32_t b = p & 31u, g = (p >> 5) & 63u, r = (p >> 11) & 31u;
    if (scaled) {
        b = (b << 3) | (b >> 2);
        g = (g << 2) | (g >> 4);
        r = (r << 3) | (r >> 2);
    }
    c4 c = {(uint8_t)r, (uint8_t)g, (uint8_t)b, 255};
    return c;
}

static inline uint16_t pack565(int r, int g, int b, int scaled)
{
    /* the values pass through a colour_quad_u8 (clamped to 0..255) first */
    uint32_t R = (uint32_t)(r < 0 ? 0 : r > 255 ? 255 : r);
    uint32_t G = (uint32_t)(g < 0 ? 0 : g > 255 ? 255 : g);
    uint32_t B = (uint32_t)(b < 0 ? 0 : b > 255 ? 255 : b);
    if (scaled) {
        R = (R * 31u + 127u) / 255u;
        G = (G * 63u + 127u) / 255u;
        B = (B * 31u + 127u) / 255u;
    }
    if (R > 31) R = 31;
    if (G > 63) G = 63;
    if (B > 31) B = 31;
    return (uint16_t)(B | (G << 5) | (R << 11));
}

static inline coords canon(coords c)
{
    if (c.lo < c.hi) {
        uint16_t t = c.lo; c.lo = c.hi; c.hi = t;
    }
    return c;
}

static inline uint32_t cdist(const e1* s, int perceptual, c4 a, c4 b)
{
    int dr = (int)a.r - (int)b.r, dg = (int)a.g - (int)b.g, db = (int)a.b - (int)b.b;
    if (perceptual) return (uint32_t)(dr * dr * 8 + dg * dg * 25 + db * db);
    if (s->grayscale) {
        int y0 = ((int)a.r * 19595 + (int)a.g * 38470 + (int)a.b * 7471 + 32768) >> 16;
        int y1 = ((int)b.r * 19595 + (int)b.g * 38470 + (int)b.b * 7471 + 32768) >> 16;
        return (uint32_t)((y0 - y1) * (y0 - y1));
    }
    return (uint32_t)(dr * dr + dg * dg + db * db);
}

static void sol_copy(solution* d, const solution* s, uint32_t n)
{
    uint8_t* keep = d->sel;
    *d = *s;
    d->sel = keep;
    memcpy(d->sel, s->sel, n);
}

static void sol_clear(solution* s)
{
    s->c.lo = s->c.hi = 0;
    s->nsel = 0;
    s->err = ~0ull;
    s->alpha_block = 0;
}

static int sel_all_equal(const solution* s)
{
    if (!s->nsel) return 0;
    for (int i = 1; i < s->nsel; i++)
        if (s->sel[i] != s->sel[0]) return 0;
    return 1;
}

/* ------------------------------------------------------------ hash sets */
static uint32_t hash32(uint32_t x)
{
    x ^= x >> 16; x *= 0x7FEB352Du; x ^= x >> 15; x *= 0x846CA68Bu; x ^= x >> 16;
    return x;
}

/* tried-solution set: returns 1 if newly inserted */
static int tried_insert(e1* s, uint32_t key)
{
    /* dedupe only skips re-evaluations, which cannot change the best
       solution; a full table simply stops deduplicating */
    if (s->scount * 4 >= (s->smask + 1) * 3) return 1;
    uint32_t h = hash32(key) & s->smask;
    while (s->sused[h]) {
        if (s->skeys[h] == key) return 0;
        h = (h + 1) & s->smask;
    }
    s->sused[h] = 1;
    s->skeys[h] = key;
    s->scount++;
    return 1;
}

/* ------------------------------------------------------------ evaluators */
static int eval_uber(e1* s, coords c, int alt)
{
    solution* t = &s->trial;
    t->c = c;
    t->nsel = (int)s->nu;
    t->err = s->best.err;
    t->alpha_block = 0;
    int first = 0, last = 1;
    if (s->have_alpha || s->force_alpha_blocks) first = 1;
    else if (!s->use_alpha_blocks) last = 0;
    c4 col[4];
    col[0] = unpack565(c.lo, 1);
    col[1] = unpack565(c.hi, 1);
    for (int type = first; type <= last; type++) {
        uint64_t terr = 0;
        int ncol;
        if (!type) {
            col[2].r = (uint8_t)((col[0].r * 2 + col[1].r + alt) / 3);
            col[2].g = (uint8_t)((col[0].g * 2 + col[1].g + alt) / 3);
            col[2].b = (uint8_t)((col[0].b * 2 + col[1].b + alt) / 3);
            col[3].r = (uint8_t)((col[1].r * 2 + col[0].r + alt) / 3);
            col[3].g = (uint8_t)((col[1].g * 2 + col[0].g + alt) / 3);
            col[3].b = (uint8_t)((col[1].b * 2 + col[0].b + alt) / 3);
            ncol = 4;
        } else {
            col[2].r = (uint8_t)((col[0].r + col[1].r + alt) >> 1);
            col[2].g = (uint8_t)((col[0].g + col[1].g + alt) >> 1);
            col[2].b = (uint8_t)((col[0].b + col[1].b + alt) >> 1);
            ncol = 3;
        }
        for (int i = (int)s->nu - 1; i >= 0; i--) {
            const c4 px = s->uc[i].c;
            s->ndist += (uint64_t)ncol;
            uint32_t be = cdist(s, s->perceptual, px, col[0]), bi = 0;
            for (int k = 1; k < ncol; k++) {
                uint32_t e = cdist(s, s->perceptual, px, col[k]);
                if (e < be) { be = e; bi = (uint32_t)k; }
            }
            terr += (uint64_t)be * s->uc[i].w;
            if (terr >= t->err) break;
            s->trial_sel[i] = (uint8_t)bi;
        }
        if (terr < t->err) {
            t->err = terr;
            t->alpha_block = type != 0;
            memcpy(t->sel, s->trial_sel, s->nu);
            t->alt_round = alt;
        }
    }
    t->enforce = !t->alpha_block && t->c.lo == t->c.hi;
    if (t->enforce) {
        int sv;
        if ((t->c.lo & 31) != 31) { t->c.lo++; sv = 1; }
        else { t->c.hi--; sv = 0; }
        memset(t->sel, sv, s->nu);
        t->enforced = sv;
    }
    if (t->err < s->best.err) {
        sol_copy(&s->best, t, s->nu);
        s->best.nsel = (int)s->nu;
        return 1;
    }
    return 0;
}

static int eval_hc(e1* s, coords c, int alt)
{
    c4 c0 = unpack565(c.lo, 1), c1 = unpack565(c.hi, 1), c2, c3;
    c2.r = (uint8_t)((c0.r * 2 + c1.r + alt) / 3); c2.g = (uint8_t)((c0.g * 2 + c1.g + alt) / 3);
    c2.b = (uint8_t)((c0.b * 2 + c1.b + alt) / 3);
    c3.r = (uint8_t)((c1.r * 2 + c0.r + alt) / 3); c3.g = (uint8_t)((c1.g * 2 + c0.g + alt) / 3);
    c3.b = (uint8_t)((c1.b * 2 + c0.b + alt) / 3);
    uint64_t err = 0;
    for (uint32_t i = 0; i < s->nu; i++) {
        const c4 px = s->uc[i].c;
        s->ndist += 4;
        uint32_t e0 = cdist(s, s->perceptual, px, c0), e1v = cdist(s, s->perceptual, px, c1);
        uint32_t e2 = cdist(s, s->perceptual, px, c2), e3 = cdist(s, s->perceptual, px, c3);
        uint32_t e01 = e0 < e1v ? e0 : e1v, e23 = e2 < e3 ? e2 : e3;
        err += (uint64_t)(e01 < e23 ? e01 : e23) * s->uc[i].w;
        if (err >= s->best.err) return 0;
    }
    if (err >= s->best.err) return 0;
    s->best.c = c;
    s->best.err = err;
    s->best.alpha_block = 0;
    s->best.alt_round = alt;
    s->best.enforce = c.lo == c.hi;
    if (s->best.enforce) {
        if ((s->best.c.lo & 31) != 31) { s->best.c.lo++; s->best.enforced = 1; }
        else { s->best.c.hi--; s->best.enforced = 0; }
    }
    return 1;
}

/* evaluate_solution_fast (crn_dxt1.cpp:1594-1757): each colour takes the
   palette entry its projection onto the endpoint axis falls nearest to
   (green weighted 24 under the perceptual metric), then one distance. */
static int eval_fast(e1* s, coords c, int alt)
{
    solution* t = &s->trial;
    t->c = c;
    t->nsel = (int)s->nu;
    t->err = s->best.err;
    t->alpha_block = 0;
    int first = 0, last = 1;
    if (s->have_alpha || s->force_alpha_blocks) first = 1;
    else if (!s->use_alpha_blocks) last = 0;
    c4 col[4];
    col[0] = unpack565(c.lo, 1);
    col[1] = unpack565(c.hi, 1);
    int vr = col[1].r - col[0].r, vg = col[1].g - col[0].g, vb = col[1].b - col[0].b;
    if (s->perceptual) {
        vr *= 8;
        vg *= 24;
    }
    int stops[4];
    stops[0] = col[0].r * vr + col[0].g * vg + col[0].b * vb;
    stops[1] = col[1].r * vr + col[1].g * vg + col[1].b * vb;
    const int dr = vr * 2, dg = vg * 2, db = vb * 2;
    for (int type = first; type <= last; type++) {
        uint64_t terr = 0;
        if (!type) {
            col[2].r = (uint8_t)((col[0].r * 2 + col[1].r + alt) / 3);
            col[2].g = (uint8_t)((col[0].g * 2 + col[1].g + alt) / 3);
            col[2].b = (uint8_t)((col[0].b * 2 + col[1].b + alt) / 3);
            col[3].r = (uint8_t)((col[1].r * 2 + col[0].r + alt) / 3);
            col[3].g = (uint8_t)((col[1].g * 2 + col[0].g + alt) / 3);
            col[3].b = (uint8_t)((col[1].b * 2 + col[0].b + alt) / 3);
            stops[2] = col[2].r * vr + col[2].g * vg + col[2].b * vb;
            stops[3] = col[3].r * vr + col[3].g * vg + col[3].b * vb;
            const int p31 = stops[1] + stops[3], p32 = stops[3] + stops[2], p20 = stops[2] + stops[0];
            for (int i = (int)s->nu - 1; i >= 0; i--) {
                const c4 px = s->uc[i].c;
                const int dot = px.r * dr + px.g * dg + px.b * db;
                uint8_t bi;
                if (dot >= p32) bi = dot < p31 ? 3 : 1;
                else bi = dot < p20 ? 0 : 2;
                s->ndist++;
                terr += (uint64_t)cdist(s, s->perceptual, px, col[bi]) * s->uc[i].w;
                if (terr >= t->err) break;
                s->trial_sel[i] = bi;
            }
        } else {
            col[2].r = (uint8_t)((col[0].r + col[1].r + alt) >> 1);
            col[2].g = (uint8_t)((col[0].g + col[1].g + alt) >> 1);
            col[2].b = (uint8_t)((col[0].b + col[1].b + alt) >> 1);
            stops[2] = col[2].r * vr + col[2].g * vg + col[2].b * vb;
            const int p02 = stops[0] + stops[2], p21 = stops[2] + stops[1];
            for (int i = (int)s->nu - 1; i >= 0; i--) {
                const c4 px = s->uc[i].c;
                const int dot = px.r * dr + px.g * dg + px.b * db;
                const uint8_t bi = dot < p02 ? 0 : dot < p21 ? 2 : 1;
                s->ndist++;
                terr += (uint64_t)cdist(s, s->perceptual, px, col[bi]) * s->uc[i].w;
                if (terr >= t->err) break;
                s->trial_sel[i] = bi;
            }
        }
        if (terr < t->err) {
            t->err = terr;
            t->alpha_block = type != 0;
            memcpy(t->sel, s->trial_sel, s->nu);
        }
    }
    /* the fast evaluator leaves the trial's rounding flag as it was: no
       alternate-rounding candidate is ever evaluated below Better */
    t->alt_round = 0;
    t->enforce = !t->alpha_block && t->c.lo == t->c.hi;
    if (t->enforce) {
        int sv;
        if ((t->c.lo & 31) != 31) { t->c.lo++; sv = 1; }
        else { t->c.hi--; sv = 0; }
        memset(t->sel, sv, s->nu);
        t->enforced = sv;
    }
    if (t->err < s->best.err) {
        sol_copy(&s->best, t, s->nu);
        s->best.nsel = (int)s->nu;
        return 1;
    }
    return 0;
}

static void compute_selectors(e1* s)
{
    if (!s->evaluate_hc) return;
    solution* b = &s->best;
    b->nsel = (int)s->nu;
    if (b->enforce) {
        memset(b->sel, b->enforced, s->nu);
        return;
    }
    c4 c0 = unpack565(b->c.lo, 1), c1 = unpack565(b->c.hi, 1), c2, c3;
    int a = b->alt_round;
    c2.r = (uint8_t)((c0.r * 2 + c1.r + a) / 3); c2.g = (uint8_t)((c0.g * 2 + c1.g + a) / 3);
    c2.b = (uint8_t)((c0.b * 2 + c1.b + a) / 3);
    c3.r = (uint8_t)((c1.r * 2 + c0.r + a) / 3); c3.g = (uint8_t)((c1.g * 2 + c0.g + a) / 3);
    c3.b = (uint8_t)((c1.b * 2 + c0.b + a) / 3);
    for (uint32_t i = 0; i < s->nu; i++) {
        const c4 px = s->uc[i].c;
        uint32_t e0 = cdist(s, s->perceptual, px, c0), e1v = cdist(s, s->perceptual, px, c1);
        uint32_t e2 = cdist(s, s->perceptual, px, c2), e3 = cdist(s, s->perceptual, px, c3);
        uint32_t e01 = e0 < e1v ? e0 : e1v, e23 = e2 < e3 ? e2 : e3;
        b->sel[i] = (uint8_t)(e01 <= e23 ? (e01 == e0 ? 0 : 1) : (e23 == e2 ? 2 : 3));
    }
}

static int evaluate(e1* s, coords c, int alt)
{
    c4 a = unpack565(c.lo, 0), b = unpack565(c.hi, 0);
    uint64_t re = a.r < b.r ? s->rlo[a.r] + s->rhi[b.r] : s->rhi[a.r] + s->rlo[b.r];
    uint64_t ge = a.g < b.g ? s->glo[a.g] + s->ghi[b.g] : s->ghi[a.g] + s->glo[b.g];
    uint64_t be = a.b < b.b ? s->blo[a.b] + s->bhi[b.b] : s->bhi[a.b] + s->blo[b.b];
    if (re + ge + be >= s->best.err) return 0;
    if (!alt && !tried_insert(s, (uint32_t)c.lo | ((uint32_t)c.hi << 16))) return 0;
    if (s->evaluate_hc) return eval_hc(s, c, alt);
    if (s->quality >= 3) return eval_uber(s, c, alt);
    return eval_fast(s, c, alt);
}

/* ------------------------------------------------------------- vectors */
static inline v3 v3s(float a, float b, float c) { v3 r = {{a, b, c}}; return r; }
static inline float dot3(v3 a, v3 b)
{
    float r = a.v[0] * b.v[0];
    r += a.v[1] * b.v[1];
    r += a.v[2] * b.v[2];
    return r;
}

static void compute_vectors(e1* s, v3 pw)
{
    s->mean = v3s(0, 0, 0);
    s->meanw = v3s(0, 0, 0);
    for (uint32_t i = 0; i < s->nu; i++) {
        const c4 c = s->uc[i].c;
        const float w = (float)s->uc[i].w;
        v3 n = v3s(c.r * 1.0f / 255.0f, c.g * 1.0f / 255.0f, c.b * 1.0f / 255.0f);
        v3 nw = v3s(pw.v[0] * n.v[0], pw.v[1] * n.v[1], pw.v[2] * n.v[2]);
        s->norm[i] = n;
        s->normw[i] = nw;
        for (int k = 0; k < 3; k++) {
            s->mean.v[k] += n.v[k] * w;
            s->meanw.v[k] += nw.v[k] * w;
        }
    }
    if (s->total_w) {
        const float inv = 1.0f / (float)s->total_w;
        for (int k = 0; k < 3; k++) {
            s->mean.v[k] *= inv;
            s->meanw.v[k] *= inv;
        }
    }
    for (uint32_t i = 0; i < s->nu; i++)
        for (int k = 0; k < 3; k++) {
            s->norm[i].v[k] -= s->mean.v[k];
            s->normw[i].v[k] -= s->meanw.v[k];
        }
}

static v3 compute_pca(e1* s, const v3* nc, v3 def)
{
    double cov[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t i = 0; i < s->nu; i++) {
        const float r = nc[i].v[0], g = nc[i].v[1], b = nc[i].v[2];
        if (s->uc[i].w > 1) {
            const double w = s->uc[i].w;
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
            r *= m; g *= m; b *= m;
        }
        double d = (vr - r) * (vr - r) + (vg - g) * (vg - g) + (vb - b) * (vb - b);
        vr = r; vg = g; vb = b;
        if (it > 2 && d < 1e-8) break;
    }
    double len = vr * vr + vg * vg + vb * vb;
    if (len < 1e-10) return def;
    len = 1.0f / sqrt(len);
    return v3s((float)(vr * len), (float)(vg * len), (float)(vb * len));
}

static void normalize_in_place(v3* a)
{
    double n = a->v[0] * a->v[0];
    n += a->v[1] * a->v[1];
    n += a->v[2] * a->v[2];
    if (n != 0) {
        const float k = (float)(1.0f / sqrt(n));
        for (int i = 0; i < 3; i++) a->v[i] *= k;
    }
}

/* ray-box intersection of crn_intersect.h:44-120 for the unit cube; returns 1 on success */
static int ray_unit_box(v3* coord, v3 org, v3 dir)
{
    int inside = 1, quad[3];
    float cand[3], maxt[3];
    for (int i = 0; i < 3; i++) {
        if (org.v[i] < 0.0f) { quad[i] = 1; cand[i] = 0.0f; inside = 0; }
        else if (org.v[i] > 1.0f) { quad[i] = 0; cand[i] = 1.0f; inside = 0; }
        else quad[i] = 2;
    }
    if (inside) { *coord = org; return 0; } /* cInside is not cSuccess */
    for (int i = 0; i < 3; i++)
        maxt[i] = (quad[i] != 2 && dir.v[i] != 0.0f) ? (cand[i] - org.v[i]) / dir.v[i] : -1.0f;
    int wp = 0;
    for (int i = 1; i < 3; i++)
        if (maxt[wp] < maxt[i]) wp = i;
    if (maxt[wp] < 0.0f) return 0;
    v3 c;
    for (int i = 0; i < 3; i++) {
        if (i != wp) {
            c.v[i] = org.v[i] + maxt[wp] * dir.v[i];
            if (c.v[i] < 0.0f || c.v[i] > 1.0f) return 0;
        } else {
            c.v[i] = cand[i];
        }
    }
    *coord = c;
    return 1;
}

/* ------------------------------------------------------------ refine */
static int refine(e1* s, int level)
{
    static const int w1tab[4] = {3, 0, 2, 1};
    static const int p0[4] = {0x00, 0x00, 0x02, 0x02};
    static const int p1[4] = {0x00, 0x09, 0x01, 0x04};
    static const int p2[4] = {0x09, 0x00, 0x04, 0x01};
    compute_selectors(s);
    double a0 = 0, a1 = 0, a2 = 0;
    double t1r = 0, t1g = 0, t1b = 0, t2r = 0, t2g = 0, t2b = 0;
    for (uint32_t i = 0; i < s->nu; i++) {
        const c4 c = s->uc[i].c;
        const double w = s->uc[i].w;
        const double r = c.r * w, g = c.g * w, b = c.b * w;
        const int step = s->best.sel[i] ^ 1;
        const int w1 = w1tab[step];
        a0 += p0[step] * w;
        a1 += p1[step] * w;
        a2 += p2[step] * w;
        t1r += w1 * r; t1g += w1 * g; t1b += w1 * b;
        t2r += r; t2g += g; t2b += b;
    }
    t2r = 3 * t2r - t1r;
    t2g = 3 * t2g - t1g;
    t2b = 3 * t2b - t1b;
    const double xx = a2, yy = a1, xy = a0;
    const double t = xx * yy - xy * xy;
    if (!yy || !xx || fabs(t) < .0000125f) return 0;
    const double frb = (3.0f * 31.0f / 255.0f) / t;
    const double fg = frb * (63.0f / 31.0f);
#define CL(v, hi) ((v) < 0 ? 0 : (v) > (hi) ? (hi) : (v))
    int e[2][3];
    e[0][0] = CL((int)((t1r * yy - t2r * xy) * frb + 0.5f), 31);
    e[0][1] = CL((int)((t1g * yy - t2g * xy) * fg + 0.5f), 63);
    e[0][2] = CL((int)((t1b * yy - t2b * xy) * frb + 0.5f), 31);
    e[1][0] = CL((int)((t2r * xx - t1r * xy) * frb + 0.5f), 31);
    e[1][1] = CL((int)((t2g * xx - t1g * xy) * fg + 0.5f), 63);
    e[1][2] = CL((int)((t2b * xx - t1b * xy) * frb + 0.5f), 31);
    int improved = 0;
    if (level == 0) {
        coords nc = {(uint16_t)((e[1][0] << 11) | (e[1][1] << 5) | e[1][2]),
                     (uint16_t)((e[0][0] << 11) | (e[0][1] << 5) | e[0][2])};
        improved |= evaluate(s, canon(nc), 0);
    } else if (level == 1) {
        for (int i = 0; i < 2; i++)
            for (int rr = -1; rr <= 1; rr++)
                for (int gr = -1; gr <= 1; gr++)
                    for (int br = -1; br <= 1; br++) {
                        int c[2][3];
                        memcpy(c, e, sizeof(c));
                        c[i][0] = CL(c[i][0] + rr, 31);
                        c[i][1] = CL(c[i][1] + gr, 63);
                        c[i][2] = CL(c[i][2] + br, 31);
                        coords nc = {pack565(c[0][0], c[0][1], c[0][2], 0), pack565(c[1][0], c[1][1], c[1][2], 0)};
                        improved |= evaluate(s, canon(nc), 0);
                    }
    } else {
        for (int orr = -1; orr <= 1; orr++)
            for (int ogr = -1; ogr <= 1; ogr++)
                for (int obr = -1; obr <= 1; obr++) {
                    int c[2][3];
                    memcpy(c, e, sizeof(c));
                    c[0][0] = CL(c[0][0] + orr, 31);
                    c[0][1] = CL(c[0][1] + ogr, 63);
                    c[0][2] = CL(c[0][2] + obr, 31);
                    for (int rr = -1; rr <= 1; rr++)
                        for (int gr = -1; gr <= 1; gr++)
                            for (int br = -1; br <= 1; br++) {
                                /* the second endpoint accumulates across iterations (crn_dxt1.cpp:668-672) */
                                c[1][0] = CL(c[1][0] + rr, 31);
                                c[1][1] = CL(c[1][1] + gr, 63);
                                c[1][2] = CL(c[1][2] + br, 31);
                                coords nc = {pack565(c[0][0], c[0][1], c[0][2], 0),
                                             pack565(c[1][0], c[1][1], c[1][2], 0)};
                                improved |= evaluate(s, canon(nc), 0);
                            }
                }
    }
#undef CL
    return improved;
}

/* --------------------------------------------------------- fast_random */
typedef struct { uint32_t jsr, jcong; } frng;

static uint32_t frng_u32(frng* r)
{
    r->jsr ^= r->jsr << 17;
    r->jsr ^= r->jsr >> 13;
    r->jsr ^= r->jsr << 5;
    r->jcong = 69069u * r->jcong + 1234567u;
    return r->jsr ^ r->jcong;
}

static float frng_f(frng* r, float l, float h)
{
    if (l >= h) return l;
    float v = (float)(l + (h - l) * (frng_u32(r) * (1.0 / 4294967296.0)));
    return v < l ? l : v > h ? h : v;
}

static int try_median4(e1* s, v3 lo, v3 hi)
{
    v3 means[4];
    if (s->nu <= 4) {
        for (int i = 0; i < 4; i++) means[i] = s->norm[(uint32_t)i < s->nu - 1 ? (uint32_t)i : s->nu - 1];
    } else {
        for (int k = 0; k < 3; k++) {
            means[0].v[k] = lo.v[k] - s->mean.v[k];
            means[3].v[k] = hi.v[k] - s->mean.v[k];
        }
        for (int k = 0; k < 3; k++) {
            means[1].v[k] = means[0].v[k] + (means[3].v[k] - means[0].v[k]) * (1.0f / 3.0f);
            means[2].v[k] = means[0].v[k] + (means[3].v[k] - means[0].v[k]) * (2.0f / 3.0f);
        }
        frng rm = {0xABCD917Au, 0x17F3DEADu};
        uint32_t rover = 0;
        float prev = 1.0e+37f;
        for (uint32_t it = 0; it < 8; it++) {
            v3 nm[4];
            float nw[4] = {0, 0, 0, 0};
            memset(nm, 0, sizeof(nm));
            float total = 0;
            for (uint32_t i = 0; i < s->nu; i++) {
                const v3 v = s->norm[i];
                float bd = 0;
                for (int k = 0; k < 3; k++) { float d = means[0].v[k] - v.v[k]; bd += d * d; }
                int bi = 0;
                for (int j = 1; j < 4; j++) {
                    float dd = 0;
                    for (int k = 0; k < 3; k++) { float d = means[j].v[k] - v.v[k]; dd += d * d; }
                    if (dd < bd) { bd = dd; bi = j; }
                }
                total += bd;
                const float w = (float)s->uc[i].w;
                for (int k = 0; k < 3; k++) nm[bi].v[k] += v.v[k] * w;
                nw[bi] += w;
            }
            uint32_t hidx = 0;
            float hw = 0;
            int empty = 0;
            for (int j = 0; j < 4; j++) {
                if (nw[j] > 0.0f) {
                    for (int k = 0; k < 3; k++) means[j].v[k] = nm[j].v[k] / nw[j];
                    if (nw[j] > hw) { hw = nw[j]; hidx = (uint32_t)j; }
                } else {
                    empty = 1;
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
                        means[j] = means[ri];
                        for (int k = 0; k < 3; k++) means[j].v[k] += frng_f(&rm, -.00196f, .00196f);
                    }
            }
        }
    }
    int improved = 0;
    for (int i = 0; i < 3; i++)
        for (int j = i + 1; j < 4; j++) {
            v3 a, b;
            for (int k = 0; k < 3; k++) {
                a.v[k] = means[i].v[k] + s->mean.v[k];
                b.v[k] = means[j].v[k] + s->mean.v[k];
            }
            coords c = {pack565((int)floor(.5f + a.v[0] * 31.0f), (int)floor(.5f + a.v[1] * 63.0f),
                                (int)floor(.5f + a.v[2] * 31.0f), 0),
                        pack565((int)floor(.5f + b.v[0] * 31.0f), (int)floor(.5f + b.v[1] * 63.0f),
                                (int)floor(.5f + b.v[2] * 31.0f), 0)};
            improved |= evaluate(s, canon(c), 0);
        }
    improved |= refine(s, s->quality == 4 ? 1 : 0);
    return improved;
}

/* ----------------------------------------------------- solid / combinatorial */
static int try_average_solid(e1* s)
{
    uint64_t tr = 0, tg = 0, tb = 0;
    uint32_t tw = 0;
    for (uint32_t i = 0; i < s->nu; i++) {
        const uint32_t w = s->uc[i].w;
        tw += w;
        tr += s->uc[i].c.r * (uint64_t)w;
        tg += s->uc[i].c.g * (uint64_t)w;
        tb += s->uc[i].c.b * (uint64_t)w;
    }
    const uint32_t half = tw >> 1;
    const uint32_t ar = (uint32_t)((tr + half) / tw), ag = (uint32_t)((tg + half) / tw), ab = (uint32_t)((tb + half) / tw);
    coords c = {(uint16_t)((om5[ar][0] << 11) | (om6[ag][0] << 5) | om5[ab][0]),
                (uint16_t)((om5[ar][1] << 11) | (om6[ag][1] << 5) | om5[ab][1])};
    int improved = evaluate(s, c, 0);
    if (s->use_alpha_blocks && s->best.err) {
        coords c3 = {(uint16_t)((om5_3[ar][0] << 11) | (om6_3[ag][0] << 5) | om5_3[ab][0]),
                     (uint16_t)((om5_3[ar][1] << 11) | (om6_3[ag][1] << 5) | om5_3[ab][1])};
        improved |= evaluate(s, c3, 0);
    }
    if (s->quality == 4) {
        for (uint32_t i = 0; i < s->nu; i++) {
            const uint32_t r = s->uc[i].c.r, g = s->uc[i].c.g, b = s->uc[i].c.b;
            if (r == ar && g == ag && b == ab) continue;
            coords d = {(uint16_t)((om5[r][0] << 11) | (om6[g][0] << 5) | om5[b][0]),
                        (uint16_t)((om5[r][1] << 11) | (om6[g][1] << 5) | om5[b][1])};
            improved |= evaluate(s, d, 0);
            if (s->use_alpha_blocks && s->best.err) {
                coords d3 = {(uint16_t)((om5_3[r][0] << 11) | (om6_3[g][0] << 5) | om5_3[b][0]),
                             (uint16_t)((om5_3[r][1] << 11) | (om6_3[g][1] << 5) | om5_3[b][1])};
                improved |= evaluate(s, d3, 0);
            }
        }
    }
    return improved;
}

static void comp_errors(e1* s, int comp, uint64_t err[4][256], uint64_t brem[4])
{
    uint64_t W[4] = {0}, WP2[4] = {0}, WPP[4] = {0};
    for (uint32_t i = 0; i < s->nu; i++) {
        const uint64_t p = comp == 0 ? s->uc[i].c.r : comp == 1 ? s->uc[i].c.g : s->uc[i].c.b;
        const uint64_t w = s->uc[i].w;
        const uint8_t sl = s->best.sel[i];
        W[sl] += w;
        WP2[sl] += w * p * 2;
        WPP[sl] += w * p * p;
    }
    const uint32_t lim = comp == 1 ? 64 : 32;
    for (int sl = 0; sl < 2; sl++) {
        uint64_t best = err[sl][0] = WPP[sl];
        for (uint32_t c = 1; c < lim; c++) {
            const uint64_t p = comp == 1 ? ((c << 2) | (c >> 4)) & 0xFF : ((c << 3) | (c >> 2)) & 0xFF;
            err[sl][c] = W[sl] * p * p - WP2[sl] * p + WPP[sl];
            if (err[sl][c] < best) best = err[sl][c];
        }
        brem[sl] = best;
    }
    for (int sl = 2; sl < 4; sl++) {
        uint64_t best = err[sl][0] = WPP[sl], d = W[sl] - WP2[sl], dd = W[sl] << 1, e = WPP[sl] + d;
        for (uint32_t p = 1; p < 256; p++, d += dd, e += d) {
            err[sl][p] = e;
            if (e < best) best = e;
        }
        brem[sl] = best;
    }
    for (int sl = 3; sl; sl--) brem[sl - 1] += brem[sl];
}

static void optimize_comps(e1* s)
{
    compute_selectors(s);
    if (s->best.alpha_block || !s->best.err) return;
    const c4 slo = unpack565(s->best.c.lo, 1), shi = unpack565(s->best.c.hi, 1);
    static uint64_t err[4][256];
    uint64_t brem[4];
    for (int comp = 0; comp < 3; comp++) {
        uint32_t p0 = comp == 0 ? slo.r : comp == 1 ? slo.g : slo.b;
        uint32_t p1 = comp == 0 ? shi.r : comp == 1 ? shi.g : shi.b;
        c4 low = unpack565(s->best.c.lo, 0), high = unpack565(s->best.c.hi, 0);
        uint8_t* lc = comp == 0 ? &low.r : comp == 1 ? &low.g : &low.b;
        uint8_t* hc = comp == 0 ? &high.r : comp == 1 ? &high.g : &high.b;
        comp_errors(s, comp, err, brem);
        uint64_t best = err[0][*lc] + err[1][*hc] + err[2][(p0 * 2 + p1) / 3] + err[3][(p0 + p1 * 2) / 3];
        if (brem[0] >= best) continue;
        const uint32_t lim = comp == 1 ? 64 : 32;
        for (uint32_t c0 = 0; c0 < lim; c0++) {
            uint64_t e0 = err[0][c0];
            if (e0 + brem[1] >= best) continue;
            *lc = (uint8_t)c0;
            const uint16_t plow = pack565(low.r, low.g, low.b, 0);
            p0 = comp == 1 ? (((c0 << 2) | (c0 >> 4)) & 0xFF) : (((c0 << 3) | (c0 >> 2)) & 0xFF);
            for (uint32_t c1 = 0; c1 < lim; c1++) {
                uint64_t e = e0 + err[1][c1];
                if (e + brem[2] >= best) continue;
                p1 = comp == 1 ? (((c1 << 2) | (c1 >> 4)) & 0xFF) : (((c1 << 3) | (c1 >> 2)) & 0xFF);
                e += err[2][(p0 * 2 + p1) / 3];
                if (e + brem[3] >= best) continue;
                e += err[3][(p0 + p1 * 2) / 3];
                if (e >= best) continue;
                *hc = (uint8_t)c1;
                coords nc = {plow, pack565(high.r, high.g, high.b, 0)};
                if (!evaluate(s, nc, 0)) continue;
                if (!s->best.err) return;
                compute_selectors(s);
                comp_errors(s, comp, err, brem);
                best = err[0][c0] + err[1][c1] + err[2][(p0 * 2 + p1) / 3] + err[3][(p0 + p1 * 2) / 3];
                e0 = err[0][c0];
                if (e0 + brem[1] >= best) break;
            }
        }
    }
}

static c4 lerp_color(c4 a, c4 b, float f, int rounding)
{
    const float r = rounding ? 1.0f : 0.0f;
    c4 o;
    int v;
    v = (int)(r + ((float)a.r + ((float)b.r - (float)a.r) * f)); o.r = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
    v = (int)(r + ((float)a.g + ((float)b.g - (float)a.g) * f)); o.g = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
    v = (int)(r + ((float)a.b + ((float)b.b - (float)a.b) * f)); o.b = (uint8_t)(v < 0 ? 0 : v > 255 ? 255 : v);
    o.a = 255;
    return o;
}

static void try_combinatorial(e1* s)
{
    if (s->nu < 2 || s->nu > 4) return;
    c4 tmp[4 + 48];
    uint32_t n = 0;
    for (uint32_t i = 0; i < s->nu; i++) tmp[n++] = s->uc[i].c;
    if (s->nu == 2) {
        static const float fs[10] = {2.0f, 3.0f, .5f, 1.5f, -1.0f, 2.0f, -.5f, .5f, -2.0f, -1.0f};
        for (int k = 0; k < 2; k++)
            for (int q = 0; q < 2; q++) {
                const int r = q ^ 1;
                for (int f = 0; f < 10; f++) {
                    /* tmp[q] and tmp[r] are the original two colours: appends never move them */
                    c4 c = lerp_color(tmp[q], tmp[r], fs[f], k);
                    tmp[n++] = c;
                }
            }
    } else if (s->nu == 3) {
        static const float fs[4] = {1.5f, 2.0f / 3.0f, 1.0f / 3.0f, -.5f};
        for (int i = 0; i <= 2; i++)
            for (int j = 0; j <= 2; j++) {
                if (i == j) continue;
                for (int f = 0; f < 4; f++) tmp[n++] = lerp_color(tmp[i], tmp[j], fs[f], 1);
            }
    }
    uint16_t packed[64];
    uint32_t np = 0;
    for (uint32_t i = 0; i < n; i++) {
        const uint16_t p = pack565(tmp[i].r, tmp[i].g, tmp[i].b, 1);
        int dup = 0;
        for (uint32_t k = 0; k < np; k++)
            if (packed[k] == p) { dup = 1; break; }
        if (!dup) packed[np++] = p;
    }
    for (uint32_t i = 0; s->best.err && i + 1 < np; i++)
        for (uint32_t j = i + 1; s->best.err && j < np; j++) {
            coords c = {packed[i], packed[j]};
            evaluate(s, c, 0);
        }
    const uint64_t err = s->best.err;
    if (err) s->best.err = 1;
    for (uint32_t i = 0; s->best.err && i + 1 < np; i++)
        for (uint32_t j = i + 1; s->best.err && j < np; j++) {
            coords c = {packed[i], packed[j]};
            evaluate(s, c, 1);
        }
    if (s->best.err) s->best.err = err;
}

/* ------------------------------------------------------ optimize endpoints */
static const int16_t k_uber_probe[15] = {0, 1, 2, 3, 5, 7, 9, 10, 13, 15, 19, 27, 43, 59, 91};
static const int16_t k_better_probe[10] = {0, 1, 2, 3, 5, 9, 15, 19, 27, 43};
static const int16_t k_normal_probe[5] = {0, 1, 3, 5, 7};
static const int16_t k_fast_probe[4] = {0, 1, 2, 3};
static const int8_t k_adj[26][3] = {
    {-1, -1, -1}, {0, -1, -1}, {1, -1, -1}, {-1, 0, -1}, {0, 0, -1}, {1, 0, -1}, {-1, 1, -1}, {0, 1, -1},
    {1, 1, -1}, {-1, -1, 0}, {0, -1, 0}, {1, -1, 0}, {-1, 0, 0}, {1, 0, 0}, {-1, 1, 0}, {0, 1, 0},
    {1, 1, 0}, {-1, -1, 1}, {0, -1, 1}, {1, -1, 1}, {-1, 0, 1}, {0, 0, 1}, {1, 0, 1}, {-1, 1, 1},
    {0, 1, 1}, {1, 1, 1}};

static void sweep(const int16_t* tab, uint32_t range, v3 base, const v3 axis2[2], uint16_t* out, uint32_t* nout)
{
    int prev[2] = {-1, -1};
    uint32_t n = 0;
    v3 init = v3s(base.v[0] + .5f, base.v[1] + .5f, base.v[2] + .5f);
    for (uint32_t i = 0; i < range; i++) {
        const int ls = i ? 0 : 1;
        const int x = tab[i];
        for (int sg = ls; sg < 2; sg++) {
            float pr = init.v[0] + axis2[sg].v[0] * (float)x;
            float pg = init.v[1] + axis2[sg].v[1] * (float)x;
            float pb = init.v[2] + axis2[sg].v[2] * (float)x;
            int r = (int)floor(pr), g = (int)floor(pg), b = (int)floor(pb);
            r = r < 0 ? 0 : r > 31 ? 31 : r;
            g = g < 0 ? 0 : g > 63 ? 63 : g;
            b = b < 0 ? 0 : b > 31 ? 31 : b;
            const int packed = b | (g << 5) | (r << 11);
            if (packed != prev[sg]) {
                out[n++] = (uint16_t)packed;
                prev[sg] = packed;
            }
        }
    }
    *nout = n;
}

static void lattice(e1* s, int which /* 0 low, 1 high */)
{
    const uint16_t cur = which ? s->best.c.hi : s->best.c.lo;
    c4 lc = unpack565(cur, 0);
    for (int i = 0; i < 26; i++) {
        const int r = lc.r + k_adj[i][0];
        if (r < 0 || r > 31) continue;
        const int g = lc.g + k_adj[i][1];
        if (g < 0 || g > 63) continue;
        const int b = lc.b + k_adj[i][2];
        if (b < 0 || b > 31) continue;
        coords c = which ? (coords){s->best.c.lo, pack565(r, g, b, 0)} : (coords){pack565(r, g, b, 0), s->best.c.hi};
        evaluate(s, canon(c), 0);
    }
    if (s->quality == 4) {
        lc = unpack565(which ? s->best.c.hi : s->best.c.lo, 0);
        for (int a = 0; a < 3; a++) {
            const int limit = a == 1 ? 63 : 31;
            for (int sg = -2; sg <= 2; sg += 4) {
                int comp[3] = {lc.r, lc.g, lc.b};
                const int q = comp[a] + sg;
                if (q < 0 || q > limit) continue;
                comp[a] = q;
                coords c = which ? (coords){s->best.c.lo, pack565(comp[0], comp[1], comp[2], 0)}
                                 : (coords){pack565(comp[0], comp[1], comp[2], 0), s->best.c.hi};
                evaluate(s, canon(c), 0);
            }
        }
    }
}

static void optimize_endpoints(e1* s, v3 lo, v3 hi)
{
    const v3 olo = lo, ohi = hi;
    const int16_t* tab = k_uber_probe;
    uint32_t range = 15, passes = 4;
    float dist = .015625f;
    if (s->quality == 3) {
        tab = k_better_probe;
        range = 10;
        passes = 2;
    } else if (s->quality == 2) {
        tab = k_normal_probe;
        range = 5;
        passes = 2;
        dist = .027063293f;
    } else if (s->quality < 2) {
        tab = k_fast_probe;
        range = 4;
        passes = s->quality == 0 ? 1 : 2;
        dist = .027063293f;
    }
    if (s->quality >= 3) try_median4(s, olo, ohi);
    v3 ax[2];
    for (int k = 0; k < 3; k++) ax[1].v[k] = s->axis.v[k] * dist;
    ax[1].v[0] *= 31.0f;
    ax[1].v[1] *= 63.0f;
    ax[1].v[2] *= 31.0f;
    for (int k = 0; k < 3; k++) ax[0].v[k] = -ax[1].v[k];
    const float lim[3] = {31.0f, 63.0f, 31.0f};
    for (int k = 0; k < 3; k++) {
        float a = lo.v[k] * lim[k];
        lo.v[k] = a < 0.0f ? 0.0f : a > lim[k] ? lim[k] : a;
        float b = hi.v[k] * lim[k];
        hi.v[k] = b < 0.0f ? 0.0f : b > lim[k] ? lim[k] : b;
    }
    uint16_t plow[31], phigh[31];
    for (uint32_t pass = 0; pass < passes; pass++) {
        if (pass) {
            c4 l = unpack565(s->best.c.lo, 0), h = unpack565(s->best.c.hi, 0);
            lo = v3s(l.r, l.g, l.b);
            hi = v3s(h.r, h.g, h.b);
        }
        const uint64_t prev = s->best.err;
        if (!prev) break;
        uint32_t nl = 0, nh = 0;
        sweep(tab, range, lo, ax, plow, &nl);
        sweep(tab, range, hi, ax, phigh, &nh);
        for (uint32_t i = 0; i < nl; i++)
            for (uint32_t j = 0; j < nh; j++) {
                coords c = {plow[i], phigh[j]};
                evaluate(s, canon(c), 0);
            }
        if (s->quality >= 2) {
            lattice(s, 0);
            lattice(s, 1);
        }
        if (!s->best.err || (pass && s->best.err == prev)) break;
        if (s->quality >= 4) refine(s, 1);
    }
    if (s->quality >= 2) {
        if (s->best.err && !s->have_alpha) {
            int solid = 0;
            if (sel_all_equal(&s->best)) solid = try_average_solid(s);
            if (!solid && s->quality == 4) optimize_comps(s);
        }
        if (s->quality == 4 && s->best.err) try_combinatorial(s);
    }
}

static void handle_multicolor(e1* s)
{
    uint32_t npass = 1;
    v3 pw = v3s(1.0f, 1.0f, 1.0f);
    if (s->perceptual) {
        float red = 0, blue = 0, al = 0;
        for (uint32_t i = 0; i < s->nu; i++) {
            const c4 c = s->uc[i].c;
            const int l = (c.r + c.g + c.b + 1) / 3;
            const float lf = (float)l;
            const float scale = (float)s->uc[i].w / (lf > 1.0f ? lf : 1.0f);
            red += scale * c.r;
            blue += scale * c.b;
            al += l;
        }
        red /= s->total_w;
        blue /= s->total_w;
        al /= s->total_w;
        al = al * 16.0f / 255.0f;
        al = 1.0f < al ? 1.0f : al;
        float m = red > blue ? red : blue;
        float sat = m * 1.0f / 3.0f;
        sat = sat < 0.0f ? 0.0f : sat > 1.0f ? 1.0f : sat;
        const float p = al * powf(sat, 2.75f);
        if (p >= 1.0f) {
            npass = 1;
        } else {
            npass = 2;
            const v3 lum = v3s(.212f, .72f, .072f);
            for (int k = 0; k < 3; k++) pw.v[k] = lum.v[k] + (pw.v[k] - lum.v[k]) * p;
        }
    }
    for (uint32_t pass = 0; pass < npass; pass++) {
        compute_vectors(s, pw);
        s->axis = compute_pca(s, s->normw, v3s(.2837149f, 0.9540631f, 0.096277453f));
        s->axis.v[0] /= pw.v[0];
        s->axis.v[1] /= pw.v[1];
        s->axis.v[2] /= pw.v[2];
        normalize_in_place(&s->axis);
        if (npass > 1) {
            if (fabsf(s->axis.v[0]) >= .795f) pw = v3s(.424f, .6f, .072f);
            else if (fabsf(s->axis.v[2]) >= .795f) pw = v3s(.212f, .6f, .212f);
            else break;
        }
    }
    float l = 1e+9f, h = -1e+9f;
    for (uint32_t i = 0; i < s->nu; i++) {
        const float d = dot3(s->norm[i], s->axis);
        l = d < l ? d : l;
        h = h < d ? d : h;
    }
    v3 lo, hi;
    for (int k = 0; k < 3; k++) {
        lo.v[k] = s->mean.v[k] + l * s->axis.v[k];
        hi.v[k] = s->mean.v[k] + h * s->axis.v[k];
    }
    int in = 1;
    for (int k = 0; k < 3; k++)
        if (lo.v[k] < 0.0f || lo.v[k] > 1.0f) in = 0;
    if (!in) {
        v3 c;
        if (ray_unit_box(&c, lo, s->axis)) lo = c;
    }
    in = 1;
    for (int k = 0; k < 3; k++)
        if (hi.v[k] < 0.0f || hi.v[k] > 1.0f) in = 0;
    if (!in) {
        v3 c;
        v3 nd = v3s(-s->axis.v[0], -s->axis.v[1], -s->axis.v[2]);
        if (ray_unit_box(&c, hi, nd)) hi = c;
    }
    optimize_endpoints(s, lo, hi);
}

/* -------------------------------------------------------------- driver */
static const uint8_t k_inv_null[4] = {0, 1, 2, 3};
static const uint8_t k_inv_alpha[4] = {1, 0, 2, 3};
static const uint8_t k_inv_color[4] = {1, 0, 3, 2};

static int lookup_color(const e1* s, uint32_t key)
{
    uint32_t h = hash32(key) & s->cmask;
    while (s->ckeys[h] != 0) {
        if (s->ckeys[h] == key) return (int)s->cvals[h];
        h = (h + 1) & s->cmask;
    }
    return -1;
}

typedef struct {
    uint64_t error;
    uint16_t low, high;
    uint8_t alpha_block, reordered, alternate_rounding, enforce_selector, enforced_selector;
} orc_dxt1_result;

/* compute_internal (crn_dxt1.cpp:2081-2232) and return_solution */
static void compute_internal(const uint8_t* px, uint32_t n, int quality, int flags, uint8_t* selectors,
                             orc_dxt1_result* res, uint64_t* ndist)
{
    e1 S;
    memset(&S, 0, sizeof(S));
    e1* s = &S;
    s->quality = quality;
    s->perceptual_req = flags & 1;
    s->use_alpha_blocks = (flags >> 1) & 1;
    s->grayscale = (flags >> 2) & 1;
    s->force_alpha_blocks = (flags >> 3) & 1;
    s->have_alpha = (flags >> 4) & 1;
    s->alpha_thresh = (flags >> 8) & 255;
    s->evaluate_hc = quality == 4 && !s->have_alpha && !s->force_alpha_blocks && !s->use_alpha_blocks && !s->grayscale;
    s->perceptual = s->perceptual_req && !s->grayscale;
    uint32_t cap = 16;
    while (cap < 2 * n) cap <<= 1;
    s->cmask = cap - 1;
    s->ckeys = (uint32_t*)calloc(cap, 4);
    s->cvals = (uint32_t*)calloc(cap, 4);
    s->uc = (ucolor*)malloc(sizeof(ucolor) * n);
    s->norm = (v3*)malloc(sizeof(v3) * n);
    s->normw = (v3*)malloc(sizeof(v3) * n);
    s->best.sel = (uint8_t*)malloc(n);
    s->trial.sel = (uint8_t*)malloc(n);
    s->trial_sel = (uint8_t*)malloc(n);
    uint32_t scap = 1 << 14;
    s->smask = scap - 1;
    s->skeys = (uint32_t*)malloc(4 * scap);
    s->sused = (uint8_t*)calloc(scap, 1);
    sol_clear(&s->best);
    for (uint32_t i = 0; i < n; i++) {
        if (s->have_alpha && px[4 * i + 3] < s->alpha_thresh) continue;
        const uint32_t key = ((uint32_t)px[4 * i] | ((uint32_t)px[4 * i + 1] << 8) | ((uint32_t)px[4 * i + 2] << 16)) | 0xFF000000u;
        int k = lookup_color(s, key);
        if (k < 0) {
            uint32_t h = hash32(key) & s->cmask;
            while (s->ckeys[h]) h = (h + 1) & s->cmask;
            s->ckeys[h] = key;
            s->cvals[h] = s->nu;
            s->uc[s->nu].c.r = px[4 * i];
            s->uc[s->nu].c.g = px[4 * i + 1];
            s->uc[s->nu].c.b = px[4 * i + 2];
            s->uc[s->nu].c.a = 255;
            s->uc[s->nu].w = 1;
            s->nu++;
        } else {
            s->uc[k].w++;
        }
        s->total_w++;
    }
    /* error planes (crn_dxt1.cpp:2134-2203) */
    uint64_t rp[32][3], gp[64][3], bp[32][3];
    memset(rp, 0, sizeof(rp));
    memset(gp, 0, sizeof(gp));
    memset(bp, 0, sizeof(bp));
    for (uint32_t i = 0; i < s->nu; i++) {
        const c4 c = s->uc[i].c;
        const uint64_t w = s->uc[i].w;
        const uint32_t R = c.r, G = c.g, B = c.b;
        const uint32_t r = (R >> 3) + (((R & 7) > (R >> 5)) ? 1 : 0);
        const uint32_t g = (G >> 2) + (((G & 3) > (G >> 6)) ? 1 : 0);
        const uint32_t b = (B >> 3) + (((B & 7) > (B >> 5)) ? 1 : 0);
        rp[r][0] += w; rp[r][1] += w * R; rp[r][2] += w * R * R;
        gp[g][0] += w; gp[g][1] += w * G; gp[g][2] += w * G * G;
        bp[b][0] += w; bp[b][1] += w * B; bp[b][2] += w * B * B;
    }
    if (s->perceptual) {
        for (int c = 0; c < 32; c++)
            for (int k = 0; k < 3; k++) rp[c][k] *= 8;
        for (int c = 0; c < 64; c++)
            for (int k = 0; k < 3; k++) gp[c][k] *= 25;
    }
    for (int c = 1; c < 32; c++)
        for (int k = 0; k < 3; k++) {
            rp[c][k] += rp[c - 1][k];
            bp[c][k] += bp[c - 1][k];
        }
    for (int c = 1; c < 64; c++)
        for (int k = 0; k < 3; k++) gp[c][k] += gp[c - 1][k];
    for (uint32_t c = 0; c < 32; c++) {
        const uint64_t C = (uint8_t)((c << 3) | (c >> 2));
        s->rlo[c] = rp[c][2] + C * C * rp[c][0] - 2 * C * rp[c][1];
        s->rhi[c] = rp[31][2] + C * C * rp[31][0] - 2 * C * rp[31][1] - s->rlo[c];
        s->blo[c] = bp[c][2] + C * C * bp[c][0] - 2 * C * bp[c][1];
        s->bhi[c] = bp[31][2] + C * C * bp[31][0] - 2 * C * bp[31][1] - s->blo[c];
    }
    for (uint32_t c = 0; c < 64; c++) {
        const uint64_t C = (uint8_t)((c << 2) | (c >> 4));
        s->glo[c] = gp[c][2] + C * C * gp[c][0] - 2 * C * gp[c][1];
        s->ghi[c] = gp[63][2] + C * C * gp[63][0] - 2 * C * gp[63][1] - s->glo[c];
    }
    const int transparent = s->total_w != n;
    if (!s->nu) {
        memset(res, 0, sizeof(*res));
        res->alpha_block = 1;
        memset(selectors, 3, n);
        free(s->ckeys); free(s->cvals); free(s->uc); free(s->norm); free(s->normw);
        free(s->best.sel); free(s->trial.sel); free(s->trial_sel); free(s->skeys); free(s->sused);
        return;
    }
    if (s->nu == 1 && !transparent) {
        const uint32_t r = s->uc[0].c.r, g = s->uc[0].c.g, b = s->uc[0].c.b;
        coords c = {(uint16_t)((om5[r][0] << 11) | (om6[g][0] << 5) | om5[b][0]),
                    (uint16_t)((om5[r][1] << 11) | (om6[g][1] << 5) | om5[b][1])};
        evaluate(s, c, 0);
        if (s->use_alpha_blocks && s->best.err) {
            coords c3 = {(uint16_t)((om5_3[r][0] << 11) | (om6_3[g][0] << 5) | om5_3[b][0]),
                         (uint16_t)((om5_3[r][1] << 11) | (om6_3[g][1] << 5) | om5_3[b][1])};
            evaluate(s, c3, 0);
        }
    } else {
        handle_multicolor(s);
    }
    /* return_solution */
    compute_selectors(s);
    int inv;
    if (s->best.alpha_block) inv = s->best.c.lo > s->best.c.hi;
    else inv = s->best.c.lo < s->best.c.hi;
    res->alternate_rounding = (uint8_t)s->best.alt_round;
    res->enforce_selector = (uint8_t)s->best.enforce;
    res->enforced_selector = (uint8_t)s->best.enforced;
    res->reordered = (uint8_t)inv;
    res->low = inv ? s->best.c.hi : s->best.c.lo;
    res->high = inv ? s->best.c.lo : s->best.c.hi;
    const uint8_t* tab = inv ? (s->best.alpha_block ? k_inv_alpha : k_inv_color) : k_inv_null;
    for (uint32_t i = 0; i < n; i++) {
        if (s->have_alpha && px[4 * i + 3] < s->alpha_thresh) {
            selectors[i] = 3;
            continue;
        }
        const uint32_t key = ((uint32_t)px[4 * i] | ((uint32_t)px[4 * i + 1] << 8) | ((uint32_t)px[4 * i + 2] << 16)) | 0xFF000000u;
        selectors[i] = tab[s->best.sel[lookup_color(s, key)]];
    }
    res->alpha_block = (uint8_t)s->best.alpha_block;
    res->error = s->best.err;
    if (ndist) *ndist += s->ndist;
    free(s->ckeys); free(s->cvals); free(s->uc); free(s->norm); free(s->normw);
    free(s->best.sel); free(s->trial.sel); free(s->trial_sel); free(s->skeys); free(s->sused);
}

static uint32_t metric(int perceptual, int grayscale, const uint8_t* a, c4 b)
{
    const int dr = (int)a[0] - (int)b.r, dg = (int)a[1] - (int)b.g, db = (int)a[2] - (int)b.b;
    if (perceptual) return (uint32_t)(dr * dr * 8 + dg * dg * 25 + db * db);
    if (grayscale) {
        const int y0 = ((int)a[0] * 19595 + (int)a[1] * 38470 + (int)a[2] * 7471 + 32768) >> 16;
        const int y1 = ((int)b.r * 19595 + (int)b.g * 38470 + (int)b.b * 7471 + 32768) >> 16;
        return (uint32_t)((y0 - y1) * (y0 - y1));
    }
    return (uint32_t)(dr * dr + dg * dg + db * db);
}

/* try_alpha_as_black_optimization (crn_dxt1.cpp:2001-2079): texels whose
   r, g and b are all <= 4 are re-encoded as DXT1A transparent texels (palette
   entry 3, black) when the block has some but not only such colours; the
   3-colour result replaces the first one if its error, now counting those
   texels against black, is lower. */
static void alpha_as_black(const uint8_t* px, uint32_t n, int quality, int flags, uint8_t* selectors,
                           orc_dxt1_result* res, uint64_t* ndist)
{
    uint32_t nu = 0, ndark = 0;
    uint32_t* seen = (uint32_t*)malloc(4 * (size_t)n);
    for (uint32_t i = 0; i < n; i++) {
        const uint32_t key = (uint32_t)px[4 * i] | ((uint32_t)px[4 * i + 1] << 8) | ((uint32_t)px[4 * i + 2] << 16);
        uint32_t k = 0;
        while (k < nu && seen[k] != key) k++;
        if (k < nu) continue;
        seen[nu++] = key;
        if (px[4 * i] <= 4 && px[4 * i + 1] <= 4 && px[4 * i + 2] <= 4) ndark++;
    }
    free(seen);
    if (!ndark || ndark == nu) return;
    uint8_t* tp = (uint8_t*)malloc(4 * (size_t)n);
    uint8_t* tsel = (uint8_t*)malloc(n);
    memcpy(tp, px, 4 * (size_t)n);
    for (uint32_t i = 0; i < n; i++)
        if (tp[4 * i] <= 4 && tp[4 * i + 1] <= 4 && tp[4 * i + 2] <= 4) tp[4 * i + 3] = 0;
    orc_dxt1_result tr;
    compute_internal(tp, n, quality, flags | 16, tsel, &tr, ndist);
    c4 pal[4];
    pal[0] = unpack565(tr.low, 1);
    pal[1] = unpack565(tr.high, 1);
    pal[2].r = (uint8_t)((pal[0].r + pal[1].r) >> 1);
    pal[2].g = (uint8_t)((pal[0].g + pal[1].g) >> 1);
    pal[2].b = (uint8_t)((pal[0].b + pal[1].b) >> 1);
    pal[3].r = pal[3].g = pal[3].b = 0;
    const int gray = (flags >> 2) & 1, perceptual = (flags & 1) && !gray;
    uint64_t err = 0;
    for (uint32_t i = 0; i < n; i++) err += metric(perceptual, gray, tp + 4 * i, pal[tsel[i]]);
    if (err < res->error) {
        res->error = err;
        res->low = tr.low;
        res->high = tr.high;
        memcpy(selectors, tsel, n);
        res->alpha_block = 1;
    }
    free(tp);
    free(tsel);
}

/* quality 0 (superfast) .. 4 (uber); flags bit0 perceptual, bit1
 * use_alpha_blocks, bit2 grayscale_sampling, bit3 force_alpha_blocks, bit4
 * pixels_have_alpha (DXT1A: texels with alpha below bits 8..15 are
 * transparent), bit5 use_transparent_indices_for_black.  px: RGBA8 texels. */
int orc_dxt1_compute_count(const uint8_t* px, uint32_t n, int quality, int flags, uint8_t* selectors,
                           orc_dxt1_result* res, uint64_t* ndist)
{
    if (!px || !n || quality < 0 || quality > 4 || (flags & ~0xFF3F)) return 0;
    init_tables();
    compute_internal(px, n, quality, flags, selectors, res, ndist);
    if (((flags >> 1) & 1) && ((flags >> 5) & 1) && !((flags >> 4) & 1))
        alpha_as_black(px, n, quality, flags, selectors, res, ndist);
    return 1;
}

int orc_dxt1_compute(const uint8_t* px, uint32_t n, int quality, int flags, uint8_t* selectors, orc_dxt1_result* res)
{
    return orc_dxt1_compute_count(px, n, quality, flags, selectors, res, 0);
}

/* blocks of 16 RGBA8 texels; out_sel 16 bytes per block; out_res per block */
int orc_dxt1_blocks(const uint8_t* px, uint32_t nblocks, int quality, int flags, uint8_t* out_sel,
                    orc_dxt1_result* out_res)
{
    for (uint32_t b = 0; b < nblocks; b++)
        if (!orc_dxt1_compute(px + 64 * (size_t)b, 16, quality, flags, out_sel + 16 * (size_t)b, &out_res[b]))
            return 0;
    return 1;
}

/* the same, also returning the number of colour distances the evaluators
   computed (early exits included): the encode's algorithmic work count */
int orc_dxt1_blocks_count(const uint8_t* px, uint32_t nblocks, int quality, int flags, uint8_t* out_sel,
                          orc_dxt1_result* out_res, uint64_t* ndist)
{
    *ndist = 0;
    for (uint32_t b = 0; b < nblocks; b++)
        if (!orc_dxt1_compute_count(px + 64 * (size_t)b, 16, quality, flags, out_sel + 16 * (size_t)b, &out_res[b],
                                    ndist))
            return 0;
    return 1;
}
