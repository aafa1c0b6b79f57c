/*
 * dxt5a_oracle.c -- CPU restatement of crnlib's dxt5_endpoint_optimizer.
 *
 * TEST INFRASTRUCTURE ONLY (see crnd_oracle.c for the rule).
 *
 * Follows crnlib/crn_dxt5a.cpp:40-262:
 *   - unique values in first-appearance order with pixel counts (:53-77);
 *   - one unique value: both endpoints = value, selectors 0, error 0 (:79-88);
 *   - every unique pair (i < j) in that order (:93-103);
 *   - at quality >= Better with a non-zero error: a +-16 (uber) / +-8 box
 *     around the current best (first, second), whose centre is re-read on
 *     every step so it follows improvements, skipping pairs already visited
 *     in either order (:105-149);
 *   - evaluate_solution (:198-262): 8-value then (if both block types) 6-value
 *     palettes (crn_dxt.cpp:407-431); per unique value the first palette entry
 *     with the smallest weight*d^2 (stop at 0), running sum with an early exit
 *     once above the best; strictly better replaces; a zero error stops;
 *   - canonical reorder (:151-184): equal endpoints -> selectors 0;
 *     6-value blocks keep first <= second, 8-value blocks first > second,
 *     inverting selectors through the six/eight invert tables.
 *
 * Parity pins (tests/test_dxt5a_oracle.py): the reported error equals the
 * weighted SSE recomputed from the emitted endpoints and selectors; the
 * result is never below the exhaustive optimum over all 65,536 pairs and both
 * block types (and equals it when two unique values bracket every pixel);
 * canonical order.  The visiting order and tie-breaking are fixed only by
 * the reference source (parity unpinned beyond these checks).
 */
#include <stdint.h>
#include <string.h>

static const uint8_t k_six_invert[8] = {1, 0, 5, 4, 3, 2, 6, 7};
static const uint8_t k_eight_invert[8] = {1, 0, 7, 6, 5, 4, 3, 2};

static void values8(uint32_t* v, uint32_t l, uint32_t h)
{
    v[0] = l; v[1] = h;
    v[2] = (l * 6 + h) / 7; v[3] = (l * 5 + h * 2) / 7; v[4] = (l * 4 + h * 3) / 7;
    v[5] = (l * 3 + h * 4) / 7; v[6] = (l * 2 + h * 5) / 7; v[7] = (l + h * 6) / 7;
}

static void values6(uint32_t* v, uint32_t l, uint32_t h)
{
    v[0] = l; v[1] = h;
    v[2] = (l * 4 + h) / 5; v[3] = (l * 3 + h * 2) / 5; v[4] = (l * 2 + h * 3) / 5;
    v[5] = (l + h * 4) / 5; v[6] = 0; v[7] = 255;
}

typedef struct {
    uint32_t nu;
    uint8_t uval[16 * 16]; /* unique values (at most 256) */
    uint32_t uw[256];
    int use_both;
    uint64_t best_err;
    uint32_t best_first, best_second, best_type;
    uint8_t best_sel[256], trial_sel[256];
} e2_state;

static void evaluate(e2_state* s, uint32_t lo, uint32_t hi)
{
    for (uint32_t type = 0; type < (s->use_both ? 2u : 1u); type++) {
        uint32_t pal[8];
        if (!type) values8(pal, lo, hi);
        else values6(pal, lo, hi);
        uint64_t err = 0;
        for (uint32_t i = 0; i < s->nu; i++) {
            const int val = s->uval[i];
            const int w = (int)s->uw[i];
            uint32_t best = 0xFFFFFFFFu, bsel = 0;
            for (uint32_t j = 0; j < 8; j++) {
                int e = val - (int)pal[j];
                e = e * e * w;
                if ((uint32_t)e < best) {
                    best = (uint32_t)e;
                    bsel = j;
                    if (!best) break;
                }
            }
            s->trial_sel[i] = (uint8_t)bsel;
            err += best;
            if (err > s->best_err) break;
        }
        if (err < s->best_err) {
            s->best_err = err;
            s->best_first = lo;
            s->best_second = hi;
            s->best_type = type;
            memcpy(s->best_sel, s->trial_sel, s->nu);
            if (!err) break;
        }
    }
}

/* quality: 0 superfast .. 4 uber (crn_dxt_quality); returns 1 on success */
int orc_dxt5a_compute(const uint8_t* values, uint32_t n, int quality, int use_both, uint8_t* selectors,
                      uint64_t* error, uint8_t* first, uint8_t* second, uint8_t* block_type, uint8_t* reordered)
{
    static __thread e2_state st;
    e2_state* s = &st;
    int map[256];
    if (!n || !values) return 0;
    for (int i = 0; i < 256; i++) map[i] = -1;
    s->nu = 0;
    for (uint32_t i = 0; i < n; i++) {
        int k = map[values[i]];
        if (k < 0) {
            k = (int)s->nu++;
            map[values[i]] = k;
            s->uval[k] = values[i];
            s->uw[k] = 0;
        }
        s->uw[k]++;
    }
    if (s->nu == 1) {
        *block_type = 0;
        *reordered = 0;
        *error = 0;
        *first = *second = s->uval[0];
        memset(selectors, 0, n);
        return 1;
    }
    s->use_both = use_both;
    s->best_err = ~0ull;
    s->best_first = s->best_second = s->best_type = 0;
    for (uint32_t i = 0; i + 1 < s->nu; i++)
        for (uint32_t j = i + 1; j < s->nu; j++) evaluate(s, s->uval[i], s->uval[j]);

    if (quality >= 3 && s->best_err) {
        static __thread uint8_t visited[256 * 256];
        memset(visited, 0, sizeof(visited));
        const int probe = (quality == 4) ? 16 : 8;
        for (int dl = -probe; dl <= probe; dl++) {
            const int l = (int)s->best_first + dl;
            if (l < 0) continue;
            if (l > 255) break;
            for (int dh = -probe; dh <= probe; dh++) {
                const int h = (int)s->best_second + dh;
                if (h < 0) continue;
                if (h > 255) break;
                if (visited[l * 256 + h] || visited[h * 256 + l]) continue;
                visited[l * 256 + h] = 1;
                evaluate(s, (uint32_t)l, (uint32_t)h);
            }
        }
    }

    uint32_t f = s->best_first, sc = s->best_second;
    *reordered = 0;
    if (f == sc) {
        memset(s->best_sel, 0, s->nu);
    } else if (s->best_type) {
        if (f > sc) {
            uint32_t t = f; f = sc; sc = t;
            *reordered = 1;
            for (uint32_t i = 0; i < s->nu; i++) s->best_sel[i] = k_six_invert[s->best_sel[i]];
        }
    } else if (f <= sc) {
        uint32_t t = f; f = sc; sc = t;
        *reordered = 1;
        for (uint32_t i = 0; i < s->nu; i++) s->best_sel[i] = k_eight_invert[s->best_sel[i]];
    }
    for (uint32_t i = 0; i < n; i++) selectors[i] = s->best_sel[map[values[i]]];
    *error = s->best_err;
    *first = (uint8_t)f;
    *second = (uint8_t)sc;
    *block_type = (uint8_t)s->best_type;
    return 1;
}

/* Batch helper: nblocks blocks of 16 values each. out_sel: 16 bytes per
 * block; out_meta: error (u64) per block; out_ep: first | second << 8 |
 * type << 16 | reordered << 24 per block. */
int orc_dxt5a_blocks(const uint8_t* values, uint32_t nblocks, int quality, int use_both, uint8_t* out_sel,
                     uint64_t* out_err, uint32_t* out_ep)
{
    for (uint32_t b = 0; b < nblocks; b++) {
        uint8_t f, sc, t, r;
        if (!orc_dxt5a_compute(values + 16 * (size_t)b, 16, quality, use_both, out_sel + 16 * (size_t)b, &out_err[b],
                               &f, &sc, &t, &r))
            return 0;
        out_ep[b] = (uint32_t)f | ((uint32_t)sc << 8) | ((uint32_t)t << 16) | ((uint32_t)r << 24);
    }
    return 1;
}
