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
