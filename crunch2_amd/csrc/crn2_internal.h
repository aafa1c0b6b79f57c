// crn2_internal.h -- declarations shared by the host-side sources of libcrn2.
#pragma once
#include <stdint.h>

#include "../../include/crn2.h"

#ifdef __cplusplus
extern "C" {
#endif

// Quantized blocks handed to the CRN writer.  Index arrays cover every block
// of the even-padded block grid of every level, levels in order, faces
// stacked within a level (rows of face f follow the rows of face f-1), rows
// in raster order -- the order in which the transcoder walks a level stream.
typedef struct crn2_pack_input {
    uint32_t width, height, levels, faces, format;
    uint32_t userdata0, userdata1;
    // colour palettes (DXT1, DXT5)
    const uint32_t* color_endpoints;        // c0 | c1 << 16, RGB565 each
    uint32_t n_color_endpoints;
    const uint32_t* color_selectors_linear; // 2 bits per texel, 0 = c0 .. 3 = c1
    uint32_t n_color_selectors;
    // alpha palettes (DXT5, DXT5A, DXN)
    const uint32_t* alpha_endpoints;        // a0 | a1 << 8
    uint32_t n_alpha_endpoints;
    const uint64_t* alpha_selectors_linear; // 3 bits per texel, 0 = a0 .. 7 = a1
    uint32_t n_alpha_selectors;
    // per-block palette indices
    const uint32_t* color_endpoint_idx;
    const uint32_t* color_selector_idx;
    const uint32_t* alpha0_endpoint_idx;
    const uint32_t* alpha0_selector_idx;
    const uint32_t* alpha1_endpoint_idx;
    const uint32_t* alpha1_selector_idx;
    // non-zero: reorder every palette as crnlib's compressor does before
    // packing (crn_comp::optimize_color / optimize_alpha and the selector
    // orders); zero: keep the given order
    uint32_t order_palettes;
} crn2_pack_input;

int crn2_pack(const crn2_pack_input* in, uint8_t** out, uint32_t* out_size);

// Test hook of the transcoder's per-prefix code lengths (kLutSize = 2,048
// bytes out) for the prefix code with the given code sizes.
int crn2_prefix_lengths(const uint8_t* code_sizes, uint32_t n, uint8_t* out);

#ifdef __cplusplus
}
#endif
