/*
 * hc_oracle.c -- CPU restatement of dxt_hc's colour selector assignment.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP selector
 * distance kernel in crunch2_amd/csrc; only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it.
 *
 * Restates dxt_hc::create_color_selector_codebook_task
 * (crnlib/crn_dxt_hc.cpp:1306-1360): for every block, the 16 x 4 table
 * E2[p][s] = color_distance(perceptual, pixel p, cluster colour s)
 * (crn_color.h:728-745, weights 8/25/1, alpha ignored); the pair/quad tables
 * E4 and E8; the error of every codebook selector as four E8 lookups; the
 * first selector with the strictly smallest error; E2 added into that
 * selector's 16 x 4 error totals and the selector marked used.  With
 * subblocks, pixels 0-7 take the colours of the first half-block's cluster
 * and pixels 8-15 those of the second.
 *
 * Also restates dxt_hc::create_alpha_selector_codebook_task
 * (crn_dxt_hc.cpp:1516-1586: E3 = squared value differences against the
 * cluster's 8 values, the pair table E6, eight E6 lookups per 3-bit codebook
 * selector, strict-< scan, totals taken against the refined values) and the
 * per-entry tournaments that end create_color_selector_codebook
 * (:1491-1503) and create_alpha_selector_codebook (:1702-1719).
 */
#include <stdint.h>
#include <string.h>

static uint32_t color_distance(int perceptual, const uint8_t* a, const uint8_t* b)
{
    const int dr = (int)a[0] - (int)b[0], dg = (int)a[1] - (int)b[1], db = (int)a[2] - (int)b[2];
    if (perceptual) return (uint32_t)(8 * dr * dr + 25 * dg * dg + db * db);
    return (uint32_t)(dr * dr + dg * dg + db * db);
}

/* blocks: n_blocks x 16 x RGBA8; endpoint_index: per block (or two per block
 * with subblocks) cluster index; cluster_colors: per cluster 4 x RGBA8;
 * selectors: 2 bits per pixel, pixel p at bits 2p..2p+1.
 * Outputs best_index[n_blocks], best_error[n_blocks] (either may be NULL);
 * details (n_selectors x 64, may be NULL) and used (n_selectors, may be
 * NULL) are accumulated into, not cleared. */
void orc_hc_color_selectors(const uint8_t* blocks, const uint32_t* endpoint_index, const uint8_t* cluster_colors,
                            uint32_t n_blocks, int subblocks, int perceptual, const uint32_t* selectors,
                            uint32_t n_selectors, uint32_t* best_index, uint32_t* best_error, uint32_t* details,
                            uint8_t* used)
{
    for (uint32_t b = 0; b < n_blocks; b++) {
        uint32_t E2[16][4], E4[8][16], E8[4][256];
        const uint8_t* px = blocks + (size_t)b * 64;
        for (uint32_t p = 0; p < 16; p++) {
            const uint32_t cl = subblocks ? endpoint_index[2 * b + (p >> 3)] : endpoint_index[b];
            for (uint32_t s = 0; s < 4; s++)
                E2[p][s] = color_distance(perceptual, px + 4 * p, cluster_colors + (size_t)cl * 16 + 4 * s);
        }
        for (uint32_t p = 0; p < 8; p++)
            for (uint32_t s = 0; s < 16; s++) E4[p][s] = E2[p << 1][s & 3] + E2[p << 1 | 1][s >> 2];
        for (uint32_t p = 0; p < 4; p++)
            for (uint32_t s = 0; s < 256; s++) E8[p][s] = E4[p << 1][s & 15] + E4[p << 1 | 1][s >> 4];
        uint32_t best = 0, best_err = 0xFFFFFFFFu;
        for (uint32_t s = 0; s < n_selectors; s++) {
            const uint32_t sel = selectors[s];
            const uint32_t e = E8[0][sel & 255] + E8[1][(sel >> 8) & 255] + E8[2][(sel >> 16) & 255] + E8[3][sel >> 24];
            if (e < best_err) {
                best_err = e;
                best = s;
            }
        }
        if (best_index) best_index[b] = best;
        if (best_error) best_error[b] = best_err;
        if (details)
            for (uint32_t p = 0; p < 16; p++)
                for (uint32_t s = 0; s < 4; s++) details[(size_t)best * 64 + 4 * p + s] += E2[p][s];
        if (used) used[best] = 1;
    }
}

/* block_vals: n_blocks x 16; values / scatter_values: per cluster 8 values;
 * selectors: 3 bits per pixel, pixel p at bits 3p.  Same outputs as above with
 * details n_selectors x 128 (16 pixels x 8 values). */
void orc_hc_alpha_selectors(const uint8_t* block_vals, const uint32_t* endpoint_index, const uint8_t* values,
                            const uint8_t* scatter_values, uint32_t n_blocks, const uint64_t* selectors,
                            uint32_t n_selectors, uint32_t* best_index, uint32_t* best_error, uint32_t* details,
                            uint8_t* used)
{
    for (uint32_t b = 0; b < n_blocks; b++) {
        uint32_t E3[16][8], E6[8][64];
        const uint8_t* bv = block_vals + (size_t)b * 16;
        const uint8_t* cv = values + (size_t)endpoint_index[b] * 8;
        for (uint32_t p = 0; p < 16; p++)
            for (uint32_t s = 0; s < 8; s++) {
                const int d = (int)bv[p] - (int)cv[s];
                E3[p][s] = (uint32_t)(d * d);
            }
        for (uint32_t p = 0; p < 8; p++)
            for (uint32_t s = 0; s < 64; s++) E6[p][s] = E3[p << 1][s & 7] + E3[p << 1 | 1][s >> 3];
        uint32_t best = 0, best_err = 0xFFFFFFFFu;
        for (uint32_t s = 0; s < n_selectors; s++) {
            const uint64_t sel = selectors[s];
            uint32_t e = 0;
            for (uint32_t q = 0; q < 8; q++) e += E6[q][(sel >> (6 * q)) & 63];
            if (e < best_err) {
                best_err = e;
                best = s;
            }
        }
        if (best_index) best_index[b] = best;
        if (best_error) best_error[b] = best_err;
        if (details) {
            const uint8_t* sv = (scatter_values ? scatter_values : values) + (size_t)endpoint_index[b] * 8;
            for (uint32_t p = 0; p < 16; p++)
                for (uint32_t s = 0; s < 8; s++) {
                    const int d = (int)bv[p] - (int)sv[s];
                    details[(size_t)best * 128 + 8 * p + s] += (uint32_t)(d * d);
                }
        }
        if (used) used[best] = 1;
    }
}

/* details: n x 16 x (1 << bits) totals; out: u32 (bits 2) or u64 (bits 3). */
void orc_hc_selector_refine(const uint32_t* details, uint32_t n, int bits, void* out)
{
    for (uint32_t i = 0; i < n; i++) {
        if (bits == 2) {
            uint32_t sel = 0;
            for (uint32_t p = 0; p < 16; p++) {
                const uint32_t* e = details + (size_t)i * 64 + 4 * p;
                const uint32_t s03 = e[3] < e[0] ? 3 : 0, s12 = e[2] < e[1] ? 2 : 1;
                sel |= (e[s12] < e[s03] ? s12 : s03) << (2 * p);
            }
            ((uint32_t*)out)[i] = sel;
        } else {
            uint64_t sel = 0;
            for (uint32_t p = 0; p < 16; p++) {
                const uint32_t* e = details + (size_t)i * 128 + 8 * p;
                const uint32_t s07 = e[7] < e[0] ? 7 : 0, s12 = e[2] < e[1] ? 2 : 1;
                const uint32_t s34 = e[4] < e[3] ? 4 : 3, s56 = e[6] < e[5] ? 6 : 5;
                const uint32_t s02 = e[s12] < e[s07] ? s12 : s07, s36 = e[s56] < e[s34] ? s56 : s34;
                sel |= (uint64_t)(e[s36] < e[s02] ? s36 : s02) << (3 * p);
            }
            ((uint64_t*)out)[i] = sel;
        }
    }
}
