/*
 * crn2.h -- C ABI of libcrn2, the MI355X (gfx950) implementation of crnlib's
 * texture-coding hot path.
 *
 * Every entry point is extern "C" and takes plain pointers and sizes.  Host
 * entry points accept host memory; the *_device entry points accept device
 * pointers (hipMalloc'ed) and a hipStream_t passed as void*.
 * Return values: 0 (CRN2_OK) on success, a negative CRN2_E* code on failure,
 * except where a function is documented to return a count, size or handle.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   crn2_unpack_begin            <- crnd::crnd_unpack_begin        inc/crn_defs.h:174, inc/crn_decomp.h:4404
 *   crn2_get_data                <- crnd::crnd_get_data            inc/crn_defs.h:178, inc/crn_decomp.h:4422
 *   crn2_unpack_level            <- crnd::crnd_unpack_level        inc/crn_defs.h:188, inc/crn_decomp.h:4441
 *   crn2_unpack_level_segmented  <- crnd::crnd_unpack_level_segmented inc/crn_defs.h:196, inc/crn_decomp.h:4457
 *   crn2_unpack_end              <- crnd::crnd_unpack_end          inc/crn_defs.h:205, inc/crn_decomp.h:4474
 *   crn2_validate_file           <- crnd::crnd_validate_file       inc/crn_defs.h:139, inc/crn_decomp.h:2672
 *   crn2_get_texture_info        <- crnd::crnd_get_texture_info    inc/crn_defs.h:143, inc/crn_decomp.h:2737
 *   crn2_get_level_info          <- crnd::crnd_get_level_info      inc/crn_defs.h:147, inc/crn_decomp.h:2761
 *   crn2_get_level_data          <- crnd::crnd_get_level_data      inc/crn_defs.h:211, inc/crn_decomp.h:2790
 *   crn2_get_bytes_per_dxt_block <- crnd::crnd_get_bytes_per_dxt_block inc/crn_defs.h:136, inc/crn_decomp.h:2651
 *   crn_get_width/height/levels/dxt_format/bytes_per_block/uncompressed_size,
 *   crn_decompress               <- the extern "C" transcode exports of emscripten/crunch_lib.cpp:32-120
 */
#ifndef CRN2_H
#define CRN2_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CRN2_OK 0
#define CRN2_EINVAL (-1)       /* invalid argument (bad pointer, size, pitch, level) */
#define CRN2_ENOMEM (-2)       /* host or device allocation failed */
#define CRN2_EUNSUPPORTED (-3) /* format outside DXT1/DXT5(+swizzles)/DXT5A/DXN */
#define CRN2_ECORRUPT (-4)     /* stream decodes to an out-of-range palette index */
#define CRN2_EHIP (-5)         /* a HIP runtime call failed */
#define CRN2_ENODEV (-6)       /* no gfx950 device / kernels not loadable */

/* crn_format values of inc/crnlib.h:74-110 */
#define CRN2_FMT_DXT1 0
#define CRN2_FMT_DXT3 1
#define CRN2_FMT_DXT5 2
#define CRN2_FMT_DXT5_CCXY 3
#define CRN2_FMT_DXT5_XGXR 4
#define CRN2_FMT_DXT5_XGBR 5
#define CRN2_FMT_DXT5_AGBR 6
#define CRN2_FMT_DXN_XY 7
#define CRN2_FMT_DXN_YX 8
#define CRN2_FMT_DXT5A 9

/* ------------------------------------------------------------------ info */
typedef struct crn2_texture_info {
    uint32_t width, height, levels, faces, bytes_per_block, userdata0, userdata1, format;
} crn2_texture_info;

typedef struct crn2_level_info {
    uint32_t width, height, faces, blocks_x, blocks_y, bytes_per_block, format;
} crn2_level_info;

/* 1 if the header and data CRC16s match and the header fields are in range. */
int crn2_validate_file(const void* data, uint32_t size);
int crn2_get_texture_info(const void* data, uint32_t size, crn2_texture_info* info);
/* bytes_per_block follows crn_decomp.h:2784 (8 only for DXT1 and DXT5A). */
int crn2_get_level_info(const void* data, uint32_t size, uint32_t level, crn2_level_info* info);
/* Pointer to the level's compressed bytes inside `data`; *level_size receives its size. */
const void* crn2_get_level_data(const void* data, uint32_t size, uint32_t level, uint32_t* level_size);
uint32_t crn2_get_bytes_per_dxt_block(uint32_t format);

/* ------------------------------------------------------------ transcoder */
typedef struct crn2_unpack_ctx crn2_unpack_ctx;

/* Parses the header and the five prefix-code models, decodes the four
 * palettes and copies file, tables and palettes to the device.  `data` must stay valid until
 * crn2_unpack_end (as crnd_unpack_begin requires).  NULL on failure. */
crn2_unpack_ctx* crn2_unpack_begin(const void* data, uint32_t size);

/* Same, for a file that already lives in device memory (`data_dev`); the
 * host copy `data` is still read for the header and models. */
crn2_unpack_ctx* crn2_unpack_begin_device(const void* data, const void* data_dev, uint32_t size, void* stream);

int crn2_get_data(crn2_unpack_ctx* ctx, const void** data, uint32_t* size);

/* Transcodes one level into the host buffers dst[0..faces-1] with the given
 * row pitch (0 = minimal; otherwise a multiple of 4 and >= blocks_x * bpb);
 * dst_size >= pitch * blocks_y.  Argument checks follow crn_decomp.h:4441-4455
 * and 3563-3575. */
int crn2_unpack_level(crn2_unpack_ctx* ctx, void** dst, uint32_t dst_size, uint32_t row_pitch, uint32_t level);

/* Segmented variant: the level's compressed bytes are given by the caller. */
int crn2_unpack_level_segmented(crn2_unpack_ctx* ctx, const void* src, uint32_t src_size, void** dst,
                                uint32_t dst_size, uint32_t row_pitch, uint32_t level);

/* Device variant: dst_dev[f] are device pointers; work is queued on `stream`
 * (a hipStream_t, NULL = default stream) and the call returns once queued. */
int crn2_unpack_level_device(crn2_unpack_ctx* ctx, void** dst_dev, uint32_t dst_size, uint32_t row_pitch,
                             uint32_t level, void* stream);

/* Transcodes levels [first, first + count) of every face into one device
 * buffer, levels in order, faces in order within a level, minimal pitch. */
int crn2_unpack_levels_device(crn2_unpack_ctx* ctx, void* dst_dev, uint64_t dst_size, uint32_t first,
                              uint32_t count, void* stream);

/* Bytes written by crn2_unpack_levels_device for the given level range. */
uint64_t crn2_levels_size(crn2_unpack_ctx* ctx, uint32_t first, uint32_t count);

int crn2_unpack_end(crn2_unpack_ctx* ctx);

/* ------------------------------------------------------ batched transcode */
/* A plan that transcodes every level of n textures in one parse launch and
 * one write launch (the reference transcodes one level per call; there is no
 * counterpart for the batch).  Texture i is written to dst_dev[i] as
 * crn2_unpack_levels_device(ctxs[i], dst_dev[i], .., 0, levels) would write
 * it.  The plan shares each context's scratch: two runs that touch the same
 * context must not overlap. */
typedef struct crn2_batch crn2_batch;
int crn2_batch_create(crn2_unpack_ctx* const* ctxs, uint32_t n, void* const* dst_dev, const uint64_t* dst_size,
                      crn2_batch** out);
int crn2_batch_run(crn2_batch* batch, void* stream);
/* Same launches, captured into a HIP graph on the first call (private capture
 * stream) and replayed on `stream` afterwards; for callers that run the same
 * plan every frame.  Profiling (crn2_profile_enable) does not see graph runs. */
int crn2_batch_run_graph(crn2_batch* b, void* stream);
/* Waits for the stream and returns CRN2_ECORRUPT if any stream of the last
 * runs decoded an out-of-range palette index (clearing the flag). */
int crn2_batch_status(crn2_batch* batch, void* stream);
int crn2_batch_destroy(crn2_batch* batch);

/* Stage timing with HIP events on the launch stream of crn2_batch_run and the
 * unpack calls (stage 0 = transition tables, 1 = stream walk, 2 / 3 = symbol
 * decode of even / odd rows, 4 = palette-index chunks, 5 = palette indices and
 * block write): enable, run, then read total ms and launch count. */
void crn2_profile_enable(int on);
void crn2_profile_reset(void);
int crn2_profile_read(int kernel, double* total_ms, uint64_t* launches);

/* -------------------------------------- emscripten-compatible C exports */
unsigned int crn_get_width(void* src, unsigned int src_size);
unsigned int crn_get_height(void* src, unsigned int src_size);
unsigned int crn_get_levels(void* src, unsigned int src_size);
unsigned int crn_get_dxt_format(void* src, unsigned int src_size);
unsigned int crn_get_bytes_per_block(void* src, unsigned int src_size);
unsigned int crn_get_uncompressed_size(void* src, unsigned int src_size, unsigned int level);
/* Levels [firstLevel, firstLevel + levelCount) of face 0, packed back to back
 * at minimal pitch, sizes computed from width >> level exactly as
 * emscripten/crunch_lib.cpp:91-120 does (no max(1, ...)). */
void crn_decompress(void* src, unsigned int src_size, void* dst, unsigned int dst_size, unsigned int firstLevel,
                    unsigned int levelCount);

/* ------------------------------------------- DXT5 alpha endpoint search */
/* crnlib's dxt5_endpoint_optimizer::compute (crnlib/crn_dxt5a.h:72,
 * crnlib/crn_dxt5a.cpp:40-196) for many 16-texel blocks at once, on the
 * device.  values: 16 bytes per block (16-byte aligned); quality 0..4
 * (crn_dxt_quality); use_both = m_use_both_block_types.  Per block out:
 * selectors48 (u64: 3-bit DXT5 selector of texel i at bit 3*i), errors (u64,
 * results::m_error), meta (u32: first | second << 8 | block_type << 16 |
 * reordered << 24). */
int crn2_dxt5a_blocks_device(const void* values, uint32_t nblocks, int quality, int use_both, void* selectors48,
                             void* errors, void* meta, void* stream);

/* Plain DXT5A/DXN-channel encode of an RGBA8 image in device memory, block by
 * block as dxt_image does for cAlphaDXT5 elements (crnlib/crn_dxt_image.cpp:
 * 330-344, 1491-1523): 4x4 texels with edge clamp, channel `comp`.  Writes one
 * 8-byte DXT5 alpha block per block at blocks + i * block_stride_u64 (u64
 * units; 2 interleaves the X and Y blocks of DXN); errors optional (u64). */
int crn2_encode_dxt5a_image_device(const void* rgba, uint32_t width, uint32_t height, uint32_t comp, int quality,
                                   int use_both, void* blocks, uint32_t block_stride_u64, void* errors, void* stream);

/* --------------------------------------------- DXT1 endpoint search */
/* crnlib's dxt1_endpoint_optimizer::compute (crnlib/crn_dxt1.h:223,
 * crnlib/crn_dxt1.cpp:2234-2246) for many 16-texel blocks at once, on the
 * device, endpoint caching off.  rgba: 16 RGBA8 texels per block (64 bytes,
 * 16-byte aligned); quality 0 (superfast) .. 4 (uber), crn_dxt_quality;
 * flags: bit0 perceptual, bit1 use_alpha_blocks (3-colour blocks), bit2
 * grayscale_sampling, bit3 force_alpha_blocks, bit4 pixels_have_alpha (DXT1A:
 * texels with alpha below the threshold in bits 8..15 are transparent,
 * selector 3), bit5 use_transparent_indices_for_black
 * (try_alpha_as_black_optimization, crn_dxt1.cpp:2001-2079), bit6 sets bit4
 * per block when a texel's alpha is below the threshold (dxt_image's DXT1A
 * rule, crn_dxt_image.cpp:1439-1450); a block without opaque texels gives
 * error 0, endpoints 0 and all selectors 3.  out: 24 bytes per block = {u64 results::m_error,
 * u32 m_low_color | m_high_color << 16, u32 selectors (2-bit DXT1 code of
 * texel i at bit 2i), u32 m_alpha_block | m_reordered << 1 |
 * m_alternate_rounding << 2 | m_enforce_selector << 3 | m_enforced_selector << 4,
 * u32 0}. */
int crn2_dxt1_blocks_device(const void* rgba, uint32_t nblocks, int quality, int flags, void* out, void* stream);

/* Plain DXT1 encode of an RGBA8 image in device memory (width x height,
 * rows of width texels, 4-byte aligned), block by block as dxt_image does for
 * cColorDXT1 elements (crnlib/crn_dxt_image.cpp:283-349 gather with edge
 * clamp, :1436-1484 set_block_pixels): per block the 8-byte dxt1_block (low
 * colour, high colour, 2-bit selectors row by row) at blocks + i *
 * block_stride_u64 (u64 units, blocks in raster order), and errors (optional,
 * u64 results::m_error).  quality and flags as crn2_dxt1_blocks_device. */
int crn2_encode_dxt1_image_device(const void* rgba, uint32_t width, uint32_t height, int quality, int flags,
                                  void* blocks, uint32_t block_stride_u64, void* errors, void* stream);

/* One whole level in any of the DXTn formats (CRN2_FMT_DXT1 .. DXT5A), the
 * unclustered dxt_image path: the element layout of dxt_image::init_internal
 * (crnlib/crn_dxt_image.cpp:120-200) filled by the GPU searches; the output is
 * the level's DXTn blocks in raster order at minimal pitch.  quality 0..4
 * (crn_dxt_quality); flags: bit0 perceptual, bit1 use both block types
 * (3-colour DXT1 blocks for DXT1 only, 6-value alpha blocks), bit2 grayscale
 * sampling, bit3 DXT1A for transparency (DXT1: blocks with a texel of alpha
 * below the threshold in bits 8..15 -- crn_comp_params::m_dxt1a_alpha_threshold,
 * 128 by default -- are 3-colour blocks with those texels transparent;
 * cCRNCompFlagDXT1AForTransparency), bit4 transparent indices for black
 * (DXT1; cCRNCompFlagUseTransparentIndicesForBlack). */
int crn2_encode_image_device(const void* rgba, uint32_t width, uint32_t height, uint32_t format, int quality,
                             int flags, void* blocks, void* stream);

/* Host-memory form for any number of 4x4 blocks (16 RGBA8 texels each,
 * r first): the work of crn_compress_block (inc/crnlib.h:679,
 * crnlib/crnlib.cpp:345-410) for nblocks blocks in one call; out receives
 * nblocks DXTn blocks of the format.  Synchronous. */
int crn2_compress_blocks(const uint32_t* pixels, uint32_t nblocks, uint32_t format, int quality, int flags,
                         void* out);

/* A texture set encoded in two launches (no reference counterpart for the
 * batch; the per-block work is that of crn2_encode_image_device): every
 * item is one level image in device memory (rgba: width x height texels, 4-byte
 * aligned) and the device buffer its DXTn blocks go to (8-byte aligned,
 * minimal pitch, the element layout of crn2_encode_image_device); formats
 * DXT1, DXT5, DXT5A, DXN_XY, DXN_YX.  All colour elements of the set run in
 * one launch, all alpha elements in a second, each thread taking elements
 * with a grid stride.  quality 0..4, flags as crn2_encode_image_device.  With
 * the autotuner on (crn2_autotune_mode(1)), the first run times six launch
 * shapes (64 / 128 / 256 threads per workgroup, one thread per element or a
 * persistent grid of 8 x 256 threads per CU) on the first 131,072 elements of
 * each launch and keeps the fastest.  crn2_encode_set_run_graph captures the
 * two launches into a HIP graph on its first call and replays it afterwards
 * (the items' buffers must stay where they were). */
typedef struct crn2_encode_item {
    const void* rgba;
    uint32_t width, height, format;
    void* blocks;
} crn2_encode_item;
typedef struct crn2_encode_set crn2_encode_set;
int crn2_encode_set_create(const crn2_encode_item* items, uint32_t n, int quality, int flags, crn2_encode_set** out);
int crn2_encode_set_run(crn2_encode_set* set, void* stream);
int crn2_encode_set_run_graph(crn2_encode_set* set, void* stream);
/* Elements per launch and the chosen launch shapes (-1: default, not tuned). */
int crn2_encode_set_info(const crn2_encode_set* set, uint64_t* colour_elements, uint64_t* alpha_elements,
                         int* colour_shape, int* alpha_shape);
int crn2_encode_set_destroy(crn2_encode_set* set);

/* ------------------------------------------------------------ DXTn decode */
/* crn_decompress_block (inc/crnlib.h:690, crnlib/crnlib.cpp:420-543) for a
 * whole level on the device: blocks (rows of (width+3)/4 blocks,
 * block_row_pitch bytes apart, 0 = minimal) of the format decoded to RGBA8
 * texels (rgba: height rows of width texels, rgba_row_pitch bytes apart, 0 =
 * width * 4).  Texels past the image edge are not written.  The swizzled DXT5
 * formats decode as DXT5; unlike the reference, DXT5 does not fall through
 * into the DXN case (crnlib.cpp:485-503). */
int crn2_decode_dxt_device(const void* blocks, uint32_t format, uint32_t width, uint32_t height,
                           uint32_t block_row_pitch, void* rgba, uint32_t rgba_row_pitch, void* stream);
/* The squared error of the decoded level against src_rgba, per channel
 * (r, g, b, a), added into sse[4] (device, u64): the sums of
 * error_metrics::compute (crnlib/crn_image_utils.cpp:1048-1120). */
int crn2_dxt_error_device(const void* blocks, uint32_t format, uint32_t width, uint32_t height,
                          uint32_t block_row_pitch, const void* src_rgba, uint32_t src_row_pitch, uint64_t* sse,
                          void* stream);
/* Host-memory form: nblocks consecutive blocks to 16 texels each.  Synchronous. */
int crn2_decompress_blocks(const void* blocks, uint32_t nblocks, uint32_t format, uint32_t* rgba);

/* The same search for pixel clusters of any size on the path dxt_hc takes for
 * its endpoint clusters (crn_dxt_hc.cpp:663-754 calling
 * dxt1_endpoint_optimizer::compute with uber quality and no 3-colour blocks;
 * crn_dxt1.cpp:1759-1869 evaluators), one wave per cluster.  rgba: all
 * clusters' texels (RGBA8 words, alpha ignored), cluster c owning texels
 * [offsets[c], offsets[c+1]); flags bit0 perceptual, bit1 m_use_alpha_blocks
 * (the uber evaluator with 3-colour blocks, crn_dxt1.cpp:1370-1592, as qdxt1
 * runs it, crn_qdxt1.cpp:616-626), bits 8..15 with bit1: the DXT1A alpha
 * threshold (a cluster with a texel below it searches its opaque texels'
 * colours as 3-colour blocks, transparent texels on selector 3).  out: 24 bytes per
 * cluster = {u64 m_error, u32 m_low_color | m_high_color << 16, u32 unique
 * colours, u32 m_reordered | m_alpha_block << 1 | m_alternate_rounding << 2 | m_enforce_selector
 * << 3 | m_enforced_selector << 4 | 256 when the cluster has more than 4096
 * unique colours (not encoded), u32 0}; selectors (optional): the 2-bit
 * DXT1 code of every texel.  Allocates stream-ordered scratch (not for graph
 * capture). */
int crn2_dxt1_clusters_device(const void* rgba, const uint32_t* offsets, uint32_t nclusters, int flags, void* out,
                              uint8_t* selectors, void* stream);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): crn2_dxt1_clusters_device with the same arguments and
 * results, bit for bit, that allocates nothing.  Its scratch (4 counters and
 * 4 lists of cluster indices) is a fixed array in device memory of 1 MiB per
 * GPU, held until the process ends, so nclusters is at most 65,536
 * (CRN2_EINVAL above).  The four cluster size classes run in turn on `stream`
 * (no helper streams).  Every call on one GPU uses that GPU's one array:
 * calls that overlap on one GPU (other streams, other threads) race and must
 * be serialised by the caller. */
int crn2_dxt1_clusters_noalloc_device(const void* rgba, const uint32_t* offsets, uint32_t nclusters, int flags,
                                      void* out, uint8_t* selectors, void* stream);

/* ------------------------------------ dxt_hc colour selector assignment */
/* dxt_hc::create_color_selector_codebook_task (crnlib/crn_dxt_hc.cpp:1306-1360)
 * for all blocks at once, on the device.  blocks: n_blocks x 16 RGBA8 texels;
 * endpoint_index: the cluster of each block (two per block, one per 4x2 half,
 * when subblocks is non-zero); cluster_colors: per cluster its 4 colours
 * (RGBA8) in selector order (cluster::color_values); selectors: the codebook,
 * 2 bits per texel, texel p at bits 2p..2p+1; perceptual: 8/25/1 channel
 * weights, else 1/1/1 (crn_color.h:728-745).  Writes, per block, the first
 * codebook entry with the smallest error (best_index) and that error
 * (best_error, optional); adds each block's 16 x 4 distance table into
 * details[best_index * 64 + 4 * texel + colour] (optional, u32, accumulated,
 * not cleared) and sets used[best_index] = 1 (optional).  All pointers are
 * device memory; runs on `stream`. */
int crn2_hc_color_selectors_device(const void* blocks, const uint32_t* endpoint_index, const void* cluster_colors,
                                   uint32_t n_blocks, int subblocks, int perceptual, const uint32_t* selectors,
                                   uint32_t n_selectors, uint32_t* best_index, uint32_t* best_error,
                                   uint32_t* details, uint8_t* used, void* stream);

/* dxt_hc::create_alpha_selector_codebook_task (crnlib/crn_dxt_hc.cpp:1516-1586)
 * for one alpha channel and all blocks at once, on the device.  block_vals:
 * n_blocks x 16 values of the channel; endpoint_index: the alpha cluster of
 * each block; values: per cluster its 8 values in selector order
 * (alpha_cluster::alpha_values), used to choose the selector;
 * scatter_values: per cluster the values the totals are taken against
 * (refined_alpha_values when the refiner replaced the endpoints, else the same
 * as values; NULL means values); selectors: the codebook, 3 bits per texel,
 * texel p at bits 3p..3p+2.  Writes per block the first codebook entry with
 * the smallest error (best_index) and that error (best_error, optional); adds
 * the block's 16 x 8 table into details[best_index * 128 + 8 * texel + value]
 * (optional, u32, accumulated, not cleared: call once per channel for DXN)
 * and sets used[best_index] = 1 (optional). */
int crn2_hc_alpha_selectors_device(const void* block_vals, const uint32_t* endpoint_index, const void* values,
                                   const void* scatter_values, uint32_t n_blocks, const uint64_t* selectors,
                                   uint32_t n_selectors, uint32_t* best_index, uint32_t* best_error,
                                   uint32_t* details, uint8_t* used, void* stream);

/* The last step of dxt_hc::create_color_selector_codebook (bits 2,
 * crnlib/crn_dxt_hc.cpp:1491-1503) and create_alpha_selector_codebook (bits 3,
 * :1702-1719): each codebook entry becomes, per texel, the code with the
 * smallest accumulated error in details (n_selectors x 16 x 4 or x 8 u32),
 * ties settled by the reference's pairwise tournament.  Writes selectors as
 * u32 (bits 2) or u64 (bits 3), texel p at bits p*bits. */
int crn2_hc_selector_refine_device(const uint32_t* details, uint32_t n_selectors, int bits, void* selectors,
                                   void* stream);

/* ------------------------------------------------- VQ tree clusterizer */
/* tree_clusterizer<vecNF>::generate_codebook (crnlib/crn_tree_clusterizer.h:
 * 89-171, split_node :265-570) as dxt_hc calls it: top-down splits of the
 * highest-variance node, max_splits - 1 of them at most, on the device.
 * vectors: n x dim floats (dim 2: alpha endpoints, 6: colour endpoints, 16:
 * selectors);
 * weights: n.  Writes the codebook (up to max_splits x dim floats, leaves in
 * node order), its size (*codebook_size, host memory) and, optionally, the
 * codebook index of each vector (vector_leaf, n).  Float sums run in a fixed
 * chunked order (documented with the kernels), and with one thread's
 * behaviour where the reference splits work across its helper threads.
 * Returns when the result is complete; allocates stream-ordered scratch (not
 * for graph capture).  CRN2_EINVAL for zero vectors, splits or total weight;
 * CRN2_EUNSUPPORTED for another dim. */
int crn2_tree_codebook_device(const float* vectors, const uint32_t* weights, uint32_t n, uint32_t dim,
                              uint32_t max_splits, float* codebook, uint32_t* codebook_size, uint32_t* vector_leaf,
                              void* stream);

/* ------------------------------------------------ dxt_hc, end to end */
/* dxt_hc::compress (crnlib/crn_dxt_hc.cpp:97-316; bool dxt_hc::compress(
 * color_quad_u8 (*)[16], vector<endpoint_indices_details>&,
 * vector<selector_indices_details>&, vector<uint32>& color_endpoints,
 * vector<uint32>& alpha_endpoints, vector<uint32>& color_selectors,
 * vector<uint64>& alpha_selectors, const params&), crn_dxt_hc.h:176-184) on
 * the device, every stage above in order.  blocks: num_blocks x 16 RGBA8
 * (device); level_desc: per level {first_block, num_blocks, block_width} and
 * level_weight per level (host; dxt_hc::params::m_levels); num_faces;
 * has_color (DXT1, DXT5) and num_alpha / alpha_comp (host; DXT5 {3}, DXT5A
 * {c}, DXN_XY {0, 1}, DXN_YX {1, 0}) as params::m_format implies;
 * codebook_sizes (host): colour endpoints, colour selectors, alpha endpoints,
 * alpha selectors; tile_params (host, optional): colour and alpha PSNR
 * derating and the colour / alpha ratio (default 2, 2, 3).  Writes (device)
 * endpoint_indices (num_blocks x 4 u16: colour, alpha0, alpha1, reference
 * code), selector_indices (num_blocks x 3 u16), the four palettes (capacity:
 * their codebook sizes; the unused ones may be NULL) and counts[4] (host:
 * entries written to each).  The reference's helper-thread splits are
 * restated for one thread.  Returns CRN2_EUNSUPPORTED when a colour cluster
 * holds more than 4,096 distinct colours.  Allocates stream-ordered scratch
 * (not for graph capture). */
int crn2_hc_compress_device(const void* blocks, uint32_t num_blocks, uint32_t nlevels, const uint32_t* level_desc,
                            const float* level_weight, uint32_t num_faces, int has_color, uint32_t num_alpha,
                            const uint32_t* alpha_comp, int perceptual, const uint32_t* codebook_sizes,
                            const float* tile_params, uint16_t* endpoint_indices, uint16_t* selector_indices,
                            uint32_t* color_endpoints, uint32_t* alpha_endpoints, uint32_t* color_selectors,
                            uint64_t* alpha_selectors, uint32_t* counts, void* stream);

/* --------------------------------------------- dxt_hc adaptive tiles */
/* dxt_hc::determine_tiles_task (crnlib/crn_dxt_hc.cpp:386-561, with the
 * derating tables of dxt_hc::compress :122-142) for every level at once, on
 * the device.  blocks: num_blocks x 16 RGBA8 texels (all levels, device
 * memory); level_desc: per level {first_block, num_blocks, block_width}
 * (host memory; widths and per-face heights in blocks must be even, else
 * CRN2_EUNSUPPORTED); has_color: DXT1 / DXT5 colour blocks; num_alpha and
 * alpha_comp (host): the alpha channels scored (DXT5: {3}; DXN: {0, 1});
 * perceptual; the colour and alpha PSNR deratings and the colour / alpha
 * weighting ratio (dxt_hc::params, defaults 2, 2, 3).  Writes per block its
 * encoding (0..7) and tile slot (tile_indices); per slot (num_blocks of them)
 * tile_src = top-left block << 4 | tile (0..8), or 0xFFFFFFFF when unused,
 * and tile_vec = 10 floats: the colour endpoints (palettize_color, 6) and
 * per alpha channel 2 (palettize_alpha).  Returns when done; allocates
 * stream-ordered scratch (not for graph capture). */
int crn2_hc_tiles_device(const void* blocks, uint32_t num_blocks, uint32_t nlevels, const uint32_t* level_desc,
                         uint32_t num_faces, int has_color, uint32_t num_alpha, const uint32_t* alpha_comp,
                         int perceptual, float color_derating, float alpha_derating, float ratio,
                         uint8_t* block_encodings, uint32_t* tile_indices, uint32_t* tile_src, float* tile_vec,
                         void* stream);

/* ----------------------------------------- dxt_hc endpoint clustering */
/* dxt_hc::determine_color_endpoints (crnlib/crn_dxt_hc.cpp:888-972; dim 6,
 * the tiles' colour endpoint vectors with weights (uint)(texels * level
 * weight)) and determine_alpha_endpoints (:1160-1239; dim 2, every alpha
 * channel's tile endpoints, weights = texels) up to the cluster of every tile,
 * on the device: the (vector, weight) pairs sorted and merged, the tree
 * clusterizer's codebook of at most codebook_size entries, and every input
 * vector's nearest entry (lowest index on ties).  Writes cluster (n),
 * codebook (optional, codebook_size x dim capacity) and *n_clusters (host
 * memory).  Vectors hold no NaN or negative zero.  Returns when done;
 * allocates stream-ordered scratch (not for graph capture). */
int crn2_hc_endpoint_clusters_device(const float* vectors, const uint32_t* weights, uint32_t n, uint32_t dim,
                                     uint32_t codebook_size, uint32_t* cluster, float* codebook,
                                     uint32_t* n_clusters, void* stream);

/* ------------------------------------------- dxt_hc selector codebooks */
/* dxt_hc::create_color_selector_codebook (crnlib/crn_dxt_hc.cpp:1379-1504) on
 * the device: the blocks' selector words (block_selectors[b] = 16 selectors,
 * texel 0 in the top bits, << 32 | weight, as crn2_hc_color_endpoints_device
 * writes them) merged, the tree clusterizer's codebook of at most
 * codebook_size entries, every block assigned its best entry against its
 * cluster's colours (endpoint_index, cluster_colors as for
 * crn2_hc_color_selectors_device), and every entry refined to its texels'
 * best codes.  Writes selectors (codebook_size capacity; texel p at bits 2p),
 * *n_selectors (host memory), selector_index per block and used per entry.
 * Returns when done; allocates stream-ordered scratch (not for graph
 * capture). */
int crn2_hc_color_selector_codebook_device(const uint64_t* block_selectors, uint32_t n_blocks, uint32_t codebook_size,
                                           const void* blocks, const uint32_t* endpoint_index,
                                           const void* cluster_colors, int perceptual, uint32_t* selectors,
                                           uint32_t* n_selectors, uint32_t* selector_index, uint8_t* used,
                                           void* stream);

/* dxt_hc::create_alpha_selector_codebook (crnlib/crn_dxt_hc.cpp:1588-1720) for
 * one (DXT5, DXT5A) or two (DXN) alpha channels: block_selectors, block_vals,
 * endpoint_index and selector_index hold n_channels consecutive arrays of
 * n_blocks entries (words = 16 3-bit selectors, texel 0 on top, << 16 |
 * weight; 16 values per block); alpha_values / refined_values per alpha
 * cluster as for crn2_hc_alpha_selectors_device.  Writes selectors (u64, texel
 * p at bits 3p), *n_selectors (host memory), selector_index and used. */
int crn2_hc_alpha_selector_codebook_device(const uint64_t* block_selectors, uint32_t n_blocks, uint32_t n_channels,
                                           uint32_t codebook_size, const void* block_vals,
                                           const uint32_t* endpoint_index, const void* alpha_values,
                                           const void* refined_values, uint64_t* selectors, uint32_t* n_selectors,
                                           uint32_t* selector_index, uint8_t* used, void* stream);

/* ------------------------------------- dxt_hc colour endpoint codebook */
/* dxt_hc::determine_color_endpoint_codebook_task (crnlib/crn_dxt_hc.cpp:663-754)
 * with dxt_endpoint_refiner::refine (crnlib/crn_dxt_endpoint_refiner.cpp:
 * 36-144, 203-302) for all clusters at once, on the device.  cluster_px: every
 * cluster's texels (RGBA8 words), cluster c owning [cluster_ofs[c],
 * cluster_ofs[c+1]) of total_px; block_cluster[b]: the cluster whose block list
 * holds block b (0xFFFFFFFF: none); blocks: nblocks x 16 RGBA8; block_weights
 * (float) and block_encodings (0..7) per block (dxt_hc's m_block_weights /
 * m_block_encodings); perceptual: 8/25/1 weights.  Writes per cluster
 * endpoints[c] = first | second << 16 (the uber search's endpoints, replaced
 * by the refiner's when their error is lower) and color_values (4 RGBA8
 * words, the search's colours in linear order); per block of a cluster
 * block_selectors[b] = 16 selectors (texel 0 in the top bits) << 32 | weight;
 * *not_encoded = clusters with more than 4096 distinct colours (left zero).
 * Allocates stream-ordered scratch (not for graph capture). */
int crn2_hc_color_endpoints_device(const void* cluster_px, const uint32_t* cluster_ofs, uint32_t nclusters,
                                   uint32_t total_px, const uint32_t* block_cluster, uint32_t nblocks,
                                   const void* blocks, const float* block_weights, const uint8_t* block_encodings,
                                   int perceptual, uint32_t* endpoints, void* color_values, uint64_t* block_selectors,
                                   uint32_t* not_encoded, void* stream);

/* dxt_hc::determine_alpha_endpoint_codebook_task (crnlib/crn_dxt_hc.cpp:1014-1130,
 * without ETC colour blocks) with the DXT5 refiner
 * (crn_dxt_endpoint_refiner.cpp:36-201) for one alpha channel and all
 * clusters at once.  cluster_vals: every cluster's alpha values (one byte per
 * texel), cluster c owning [cluster_ofs[c], cluster_ofs[c+1]) of total_vals;
 * block_cluster[b]: the cluster whose list holds block b (0xFFFFFFFF: none);
 * block_vals: nblocks x 16 alpha values of that channel; block_encodings per
 * block.  Writes per cluster endpoints[c] = first | second << 8 | refined <<
 * 16, alpha_values and refined_values (8 bytes each: the search's and the
 * final endpoints' values in linear order); per block of a cluster
 * block_selectors[b] = 16 3-bit selectors (texel 0 on top) << 16 | weight.
 * Allocates stream-ordered scratch (not for graph capture). */
int crn2_hc_alpha_endpoints_device(const void* cluster_vals, const uint32_t* cluster_ofs, uint32_t nclusters,
                                   uint32_t total_vals, const uint32_t* block_cluster, uint32_t nblocks,
                                   const void* block_vals, const uint8_t* block_encodings, uint32_t* endpoints,
                                   void* alpha_values, void* refined_values, uint64_t* block_selectors, void* stream);

/* ------------------------------------------------------------- autotuner */
/* No reference counterpart (the reference's task pool splits work evenly,
 * crn_dxt_image.cpp:480-485).  Kernels with several launch shapes of identical
 * results (crn2_hc_color_selectors_device: waves per workgroup;
 * crn2_dxt1_blocks_device, crn2_dxt5a_blocks_device: workgroup size) time
 * every shape on the first call per size class (log2 of the work size) when
 * mode is 1, and reuse the fastest afterwards; mode 0 (default) runs the
 * default shape and forgets earlier choices.  kernel: 0 selector assignment,
 * 1 DXT1, 2 DXT5A. */
int crn2_autotune_mode(int mode);
/* The shape chosen for `work` (n_blocks * n_selectors for kernel 0, n_blocks
 * otherwise) and its time per launch in ms; CRN2_EINVAL if none yet. */
int crn2_autotune_choice(int kernel, uint64_t work, int* shape, float* ms);

/* ---------------------------------------- learned DXT1 endpoint proposal */
/* No reference counterpart: a small network proposes DXT1 endpoints that an
 * integer search refines; crnlib's dxt1_endpoint_optimizer
 * (crnlib/crn_dxt1.cpp:2234-2246) stays the fallback and the exact block
 * error judges every result.  Model: 57 features -> 128 -> 128 -> 7, tanh-form
 * GELU, bfloat16 weights and layer inputs, float32 sums; parameters theta
 * (24,839 float32) in the order W1 [57][128], b1, W2 [128][128], b2,
 * W3 [128][7], b3.  All pointers device memory, work queued on `stream`. */
/* Features of 16-texel RGBA8 blocks: 48 texel channels / 255 - 0.5, then the
 * 3x3 colour covariance x 4 (57 floats per block). */
int crn2_mlp_features_device(const void* rgba, uint32_t n, float* feat, void* stream);
/* C(m,n) [+]= sum_k A(m,k) B(k,n) in float32 with A(m,k) = A[m*sam + k*sak],
 * B(k,n) = B[k*sbk + n*sbn], C row-major (ldc); round_a / round_b round the
 * operands to bfloat16 on load; ksplit > 1 splits K and adds atomically into
 * C (which must start cleared). */
int crn2_gemm_f32_device(uint32_t M, uint32_t N, uint32_t K, const float* A, int64_t sam, int64_t sak, const float* B,
                         int64_t sbk, int64_t sbn, float* C, int64_t ldc, int accumulate, int round_a, int round_b,
                         uint32_t ksplit, void* stream);
/* Z += bias (per column); A = bfloat16(gelu(Z)) if use_gelu else A = Z. */
int crn2_mlp_bias_act_device(float* Z, float* A, const float* bias, uint32_t rows, uint32_t cols, int use_gelu,
                             void* stream);
/* dZ = dA * gelu'(Z). */
int crn2_mlp_gelu_back_device(const float* dA, const float* Z, float* dZ, uint64_t n, void* stream);
/* Loss of outputs Y (rows x 7) against targets T (rows x 6): batch mean of
 * squared error of outputs 0-5 plus 0.05 x logistic loss of output 6 against
 * labels (one float 0/1 per row), or with labels NULL against "every rounded
 * endpoint component within one step"; writes dY, adds the loss into *loss. */
int crn2_mlp_loss_device(const float* Y, const float* T, const float* labels, uint32_t rows, float* dY, float* loss,
                         void* stream);
/* Adam step `step` (1-based) on n parameters. */
int crn2_adam_device(float* w, const float* g, float* m, float* v, uint32_t n, float lr, float b1, float b2, float eps,
                     uint32_t step, void* stream);
/* Network outputs for n RGBA8 blocks on the bfloat16 matrix cores: y holds
 * 8 floats per block (outputs 0-6, one pad). */
int crn2_mlp_infer_device(const void* rgba, uint32_t n, const float* theta, float* y, void* stream);
/* Refines start endpoints by +-1 steps of one endpoint channel while the
 * exact error drops (at most max_iters steps per start) and keeps the best.
 * starts: bit 0 = the proposal y (8 floats per block as crn2_mlp_infer_device
 * writes, rounded to RGB565), bit 1 = the block's brightest and darkest
 * texels (luma 2r + 4g + b), bit 2 = the extremes along the texels' principal
 * axis; bit 3 = after each descent, refit both endpoints by least squares to
 * the selectors (refine_solution, crnlib/crn_dxt1.cpp:525-698) and descend
 * again while that lowers the error; bit 4 = then move the kept result to
 * the best of its 728 neighbours with every endpoint component changed by
 * -1 / 0 / +1 (refine_solution's both-ends lattice, crn_dxt1.cpp:601-643)
 * while that lowers the error, descending again after each move.
 * out: 4 u32 per block = {c0 | c1 << 16
 * (c0 > c1: four colours), selectors (2 bits per texel), perceptual error
 * (8/25/1) under the standard DXT1 decode, largest step count of a start
 * (max_iters: stopped before converging)}. */
int crn2_dxt1_refine_device(const void* rgba, uint32_t n, const float* y, uint32_t max_iters, uint32_t starts,
                            uint32_t* out, void* stream);
/* The same error for given endpoint words (c0 | c1 << 16 at
 * endpoints[i * stride_u32]) with the best selectors (written to selectors
 * when it is not null). */
int crn2_dxt1_block_error_device(const void* rgba, uint32_t n, const uint32_t* endpoints, uint32_t stride_u32,
                                 uint32_t* err, uint32_t* selectors, void* stream);

/* ------------------------------------------- host-memory (synchronous) */
/* The same work as the _device entry points of the same name, for callers
 * with host buffers: inputs staged to the device, results copied back before
 * the call returns (these back the C++ facades of include/cxx). */
int crn2_dxt1_blocks(const void* rgba, uint32_t nblocks, int quality, int flags, void* results);
int crn2_dxt1_clusters(const void* rgba, const uint32_t* offsets, uint32_t nclusters, int flags, void* results,
                       uint8_t* selectors);
int crn2_dxt5a_blocks(const void* values, uint32_t nblocks, int quality, int use_both, void* selectors48,
                      void* errors, void* meta);
/* One whole level (rgba: width x height texels) to DXTn blocks as
 * crn2_encode_image_device. */
int crn2_encode_image(const void* rgba, uint32_t width, uint32_t height, uint32_t format, int quality, int flags,
                      void* blocks);

/* ------------------------------------------------------------ whole files */
/* The CRN path of crn_compress (inc/crnlib.h:609, crnlib/crn_comp.cpp:
 * 521-742): levels[f * nlevels + l] = RGBA8 level l of face f (max(1, width
 * >> l) x max(1, height >> l)); the blocks of every level (grid padded to a
 * multiple of 8 texels, edge clamp, level weights min(12, 1.3^l)) go through
 * the GPU dxt_hc (crn2_hc_compress_device with codebook_sizes and
 * tile_params as there) and the project's CRN writer.  Formats DXT1, DXT5,
 * DXT5A, DXN_XY, DXN_YX.  *out is freed with crn2_free. */
int crn2_compress_crn(const uint8_t* const* levels, uint32_t width, uint32_t height, uint32_t nlevels, uint32_t faces,
                      uint32_t format, int perceptual, const uint32_t* codebook_sizes, const float* tile_params,
                      uint32_t userdata0, uint32_t userdata1, uint8_t** out, uint32_t* out_size);
/* The plain DDS path of crn_compress: every level of every face through
 * crn2_encode_image, written by crn2_write_dds. */
int crn2_compress_dds(const uint8_t* const* levels, uint32_t width, uint32_t height, uint32_t nlevels,
                      uint32_t faces, uint32_t format, int quality, int flags, uint8_t** out, uint32_t* out_size);
/* The clustered DDS path of crn_compress below quality 255 (dds_comp::
 * convert_to_dxt, crnlib/crn_dds_comp.cpp:148-205; qdxt_pack_init /
 * qdxt_pack, crn_mipmapped_texture.cpp:2310-2560): every face's levels
 * through crn2_qdxt1_* (DXT1, DXT5's colour with the 0.75 codebook curve)
 * and crn2_qdxt5_* (DXT5's alpha, DXT5A, DXN's two channels), written by
 * crn2_write_dds.  The colour and each alpha channel run at once, each on
 * a stream and host thread of its own (qdxt_pack runs them one after the
 * other; the results are the same).  flags: bit0 perceptual, bit1 both
 * block types, bit2 hierarchical.  Formats DXT1, DXT5, DXT5A, DXN_XY,
 * DXN_YX (CRN2_EUNSUPPORTED otherwise, and where crn2_qdxt1_pack is).
 * *out is freed with crn2_free. */
int crn2_compress_dds_clustered(const uint8_t* const* levels, uint32_t width, uint32_t height, uint32_t nlevels,
                                uint32_t faces, uint32_t format, uint32_t quality_level, int dxt_quality, int flags,
                                uint32_t alpha_threshold, uint8_t** out, uint32_t* out_size);
/* A DDS file as mipmapped_texture::write_dds writes a DXTn texture
 * (crnlib/crn_mipmapped_texture.cpp:921-1140): level_data[f * levels + l] =
 * that level's blocks at minimal pitch.  *out is freed with crn2_free. */
int crn2_write_dds(uint32_t width, uint32_t height, uint32_t levels, uint32_t faces, uint32_t format,
                   const void* const* level_data, uint8_t** out, uint32_t* out_size);
/* crn_decompress_crn_to_dds (inc/crnlib.h:620, crnlib/crnlib.cpp:269-291):
 * every level and face transcoded on the GPU, then written as a DDS file. */
int crn2_crn_to_dds(const void* crn, uint32_t size, uint8_t** out, uint32_t* out_size);

/* ------------------------------------------------------- mip generation */
/* mipmapped_texture::generate_mipmaps (crnlib/crn_mipmapped_texture.cpp:
 * 2140-2220): level l (max(1, w >> l) x max(1, h >> l)) resampled from level 0
 * by image_utils::resample (crnlib/crn_image_utils.cpp:478-865) with crnlib's
 * Resampler (crnlib/crn_resampler.cpp) -- same levels, bytes for bytes.
 * filter: the crn_mip_filter values (inc/crnlib.h:437-449). */
#define CRN2_MIP_BOX 0
#define CRN2_MIP_TENT 1
#define CRN2_MIP_LANCZOS4 2
#define CRN2_MIP_MITCHELL 3
#define CRN2_MIP_KAISER 4
#define CRN2_MIP_SRGB 1        /* gamma filtering: R, G, B through the source-gamma tables */
#define CRN2_MIP_WRAP 2        /* tiled: contributors wrap around the edges (else clamp) */
#define CRN2_MIP_ALPHA 4       /* alpha is valid and resampled (else 255 in levels >= 1) */
#define CRN2_MIP_RENORMALIZE 8 /* image_utils::renorm_normal_map on every generated level */
#define CRN2_MIP_THREADED 16   /* the multithreaded resampler's order (X pass then Y pass on
                                  every level), which crnlib uses with helper threads;
                                  else the single-thread Resampler's per-level choice */
typedef struct crn2_mip_params {
    uint32_t filter;       /* CRN2_MIP_BOX .. CRN2_MIP_KAISER */
    uint32_t flags;        /* CRN2_MIP_* flags */
    float filter_scale;    /* crn_mipmap_params::m_blurriness (0.9) */
    float source_gamma;    /* resample_params::m_source_gamma (2.2: generate_mipmaps keeps it) */
    uint32_t max_levels;   /* total levels, level 0 included; 0 = no limit */
    uint32_t min_mip_size; /* halve while either side is above this (1) */
} crn2_mip_params;
/* Level count of generate_mipmaps (crn_mipmapped_texture.cpp:2145-2157). */
uint32_t crn2_mip_level_count(uint32_t width, uint32_t height, uint32_t min_mip_size, uint32_t max_levels);
/* A reusable plan for one level-0 size: contributor lists and gamma tables on
 * the device, scratch for the two passes.  Sides up to 16,384
 * (CRNLIB_RESAMPLER_MAX_DIMENSION). */
typedef struct crn2_mip_plan crn2_mip_plan;
int crn2_mip_plan_create(uint32_t width, uint32_t height, const crn2_mip_params* params, crn2_mip_plan** out);
/* levels, bytes of the whole chain (RGBA8, levels 0..n-1 back to back), the
 * byte offset of each level (n entries) and 1 where the level ran Y then X
 * (n entries, 0 for level 0); any pointer may be NULL. */
int crn2_mip_plan_info(const crn2_mip_plan* plan, uint32_t* levels, uint64_t* chain_bytes, uint64_t* level_offsets,
                       uint8_t* y_first);
/* rgba (device, RGBA8 width x height, tightly packed) -> out (device,
 * chain_bytes): level 0 copied, levels 1..n-1 generated, all on `stream`.
 * Allocates nothing; calls on one plan must not overlap. */
int crn2_mip_plan_run(crn2_mip_plan* plan, const void* rgba, void* out, uint64_t out_size, void* stream);
int crn2_mip_plan_destroy(crn2_mip_plan* plan);
/* Stage timing of a plan's runs with HIP events on the run stream (each run
 * then waits for its events): on != 0 switches it on (totals reset); ms[0..2]
 * = total ms of the level-0 planes, pass 1 and pass 2 since then. */
int crn2_mip_plan_profile(crn2_mip_plan* plan, int on, double* ms, uint64_t* runs);
/* One-shot form: plan, run on `stream`, wait, free. */
int crn2_generate_mipmaps_device(const void* rgba, uint32_t width, uint32_t height, const crn2_mip_params* params,
                                 void* out, uint64_t out_size, void* stream);
/* Host-memory form (the path of crn_compress): rgba and out in host memory. */
int crn2_generate_mipmaps(const void* rgba, uint32_t width, uint32_t height, const crn2_mip_params* params,
                          void* out, uint64_t out_size);
/* Host copy of the contributor list of one axis (Resampler::make_clist,
 * crn_resampler.cpp:117-416): offs[dst + 1] and up to cap (pixel, weight)
 * pairs; returns the pair count, or a negative CRN2_E* code. */
int crn2_mip_contributors(uint32_t src, uint32_t dst, int wrap, uint32_t filter, float filter_scale, uint32_t* offs,
                          uint32_t* pix, float* weight, uint32_t cap);
/* The 256-entry sRGB-to-linear and 8,192-entry linear-to-sRGB tables of
 * resample_single_thread (crn_image_utils.cpp:514-545). */
int crn2_mip_gamma_tables(float gamma, float* srgb_to_linear, uint8_t* linear_to_srgb);

/* ------------------------------------------- clustered DXT1 (qdxt1) */
/* crnlib's qdxt1 (crnlib/crn_qdxt1.h:36-229, crn_qdxt1.cpp): DXT1 blocks
 * whose endpoints and selectors are shared across clusters of similar
 * blocks, the DDS path of crn_compress below the top quality.  blocks
 * (device): the dxt_pixel_blocks of every level, 16 RGBA8 texels each (r
 * first), levels back to back; level_desc (host): first block, width and
 * height in blocks per level (qdxt1_params::m_mip_desc).  init
 * (qdxt1::init, :80-430) computes the endpoint training vectors, grows the
 * endpoint tree to 65,535 leaves and counts the distinct dxt_fast selector
 * words; pack (qdxt1::pack, :876-1023) writes one 8-byte dxt1_block per
 * block (low colour, high colour, selector bytes) to out_blocks (device).
 * Endpoint caching is off in the endpoint search.  Supported: quality 255
 * at any dxt quality but superfast; below 255, uber quality with or without
 * 3-colour blocks (DXT1A transparency included) -- CRN2_EUNSUPPORTED
 * otherwise.  Streams: init and pack return with their work on `stream`
 * done; the handle's device buffers are stream-ordered allocations
 * (hipMallocAsync) that never make a call wait for other streams, so
 * handles on different streams run concurrently.  Calls on one handle must
 * not overlap; destroy waits for the stream of the handle's last call, so
 * call it before destroying that stream. */
#define CRN2_QDXT_PERCEPTUAL 1
#define CRN2_QDXT_ALPHA_BLOCKS 2   /* qdxt1_params::m_use_alpha_blocks */
#define CRN2_QDXT_HIERARCHICAL 4   /* qdxt1_params::m_hierarchical */
typedef struct crn2_qdxt1_params {
    uint32_t quality_level;          /* 0..255 (cCRNMaxQualityLevel) */
    int dxt_quality;                 /* crn_dxt_quality */
    int flags;                       /* CRN2_QDXT_* */
    uint32_t dxt1a_alpha_threshold;  /* texels below are transparent (with alpha blocks) */
} crn2_qdxt1_params;
typedef struct crn2_qdxt1 crn2_qdxt1;
int crn2_qdxt1_init(const void* blocks, uint32_t nblocks, uint32_t nlevels, const uint32_t* level_desc,
                    int hierarchical, crn2_qdxt1** out, void* stream);
/* quality_power_mul: qdxt_pack's codebook-curve factor (1; 0.75 for DXT5's
 * colour, 1.5 for DXT5_CCxY). */
int crn2_qdxt1_pack(crn2_qdxt1* q, const crn2_qdxt1_params* params, float quality_power_mul, void* out_blocks,
                    void* stream);
/* Endpoint codebook size (tree leaves), selector budget (distinct dxt_fast
 * selector words + 128), and the cluster counts of the last pack. */
int crn2_qdxt1_info(const crn2_qdxt1* q, uint32_t* codebook_size, uint32_t* max_selector_clusters,
                    uint32_t* endpoint_clusters, uint32_t* selector_clusters);
/* Host copies of init's training vectors (6 floats per block), weights and
 * per-block tile encodings (255: not hierarchical); any pointer may be NULL. */
int crn2_qdxt1_training(const crn2_qdxt1* q, float* vecs, uint32_t* weights, uint8_t* encodings);
int crn2_qdxt1_destroy(crn2_qdxt1* q);
/* crnlib's qdxt5 (crnlib/crn_qdxt5.h:36-228, crn_qdxt5.cpp): clustered DXT5
 * alpha blocks of one channel (comp_index: 3 for DXT5's alpha, 0 / 1 for
 * DXN's X and Y), the alpha half of the DDS path below the top quality.
 * blocks / level_desc as crn2_qdxt1_init; init (qdxt5::init, :80-428): 2-D
 * endpoint training vectors, the endpoint tree to 65,535 leaves, the
 * distinct dxt_fast alpha selector words; pack (qdxt5::pack, :843-935)
 * writes one 8-byte DXT5 alpha block per block (first, second, 48 bits of
 * 3-bit selectors, texel 0 lowest) to out_blocks (device).  dxt_quality
 * superfast .. uber (superfast: dxt_fast's alpha compressor over each
 * endpoint cluster, crn_qdxt5.cpp:532-541).  Streams as crn2_qdxt1. */
#define CRN2_QDXT5_BOTH_BLOCK_TYPES 1   /* qdxt5_params::m_use_both_block_types */
typedef struct crn2_qdxt5_params {
    uint32_t quality_level;   /* 0..255 */
    int dxt_quality;          /* crn_dxt_quality */
    int flags;                /* CRN2_QDXT5_BOTH_BLOCK_TYPES */
} crn2_qdxt5_params;
typedef struct crn2_qdxt5 crn2_qdxt5;
int crn2_qdxt5_init(const void* blocks, uint32_t nblocks, uint32_t nlevels, const uint32_t* level_desc,
                    int hierarchical, uint32_t comp_index, crn2_qdxt5** out, void* stream);
int crn2_qdxt5_pack(crn2_qdxt5* q, const crn2_qdxt5_params* params, void* out_blocks, void* stream);
int crn2_qdxt5_info(const crn2_qdxt5* q, uint32_t* codebook_size, uint32_t* max_selector_clusters,
                    uint32_t* endpoint_clusters, uint32_t* selector_clusters);
/* Host copies of init's training vectors (2 floats per block), weights and
 * tile encodings; any pointer may be NULL. */
int crn2_qdxt5_training(const crn2_qdxt5* q, float* vecs, uint32_t* weights, uint8_t* encodings);
int crn2_qdxt5_destroy(crn2_qdxt5* q);
/* Stage timing of the clusterizer's tree splits (no reference counterpart):
 * on != 0 resets and starts it (HIP events around every split launch, a
 * device counter of vector passes); on == 0 stops it and writes per tree
 * dimension (2, 6, 16) the summed launch ms, launches and vector passes (one
 * vector read by one pass of a split: covariance, PCA sides, each 2-means
 * round, the partition).  Any output pointer may be NULL. */
int crn2_qdxt_profile(int on, double* ms, uint64_t* launches, uint64_t* passes);
/* clusterizer<vec<dim,float>> (crnlib/crn_clusterizer.h) over n vectors in
 * device memory (dim 2, 6 or 16; weights u32): grown best first to max_size
 * leaves, then retrieve_clusters(max_clusters; 0 = every leaf).  idx (host,
 * n) and ofs (host, n + 1) receive the clusters' vector lists (ascending
 * within a cluster, clusters in depth-first order). */
int crn2_vq_clusters_device(const float* vecs, const uint32_t* weights, uint32_t n, uint32_t dim, uint32_t max_size,
                            uint32_t max_clusters, uint32_t* idx, uint32_t* ofs, uint32_t* nclusters,
                            uint32_t* codebook_size, void* stream);
/* threaded_clusterizer::create_clusters (crnlib/crn_threaded_clusterizer.h:
 * 71-181) for vectors in device memory; results as crn2_vq_clusters_device. */
int crn2_vq_threaded_clusters_device(const float* vecs, const uint32_t* weights, uint32_t n, uint32_t dim,
                                     uint32_t max_clusters, uint32_t* idx, uint32_t* ofs, uint32_t* nclusters,
                                     void* stream);

/* ---------------------------------------------- writer and test textures */
/* CRC16 of crn_decomp.h:2048-2069. */
uint16_t crn2_crc16(const void* p, uint32_t len, uint16_t crc);
/* Seeded synthetic RGBA8 texture: kind 0 = RGB, 1 = RGB + radial alpha, 2 = normal map. */
int crn2_synth_rgba(uint32_t w, uint32_t h, uint64_t seed, int kind, uint8_t* out);
/* 2x2 box filter with rounding, w x h -> max(1,w/2) x max(1,h/2). */
int crn2_downsample(const uint8_t* src, uint32_t w, uint32_t h, uint8_t* dst);
/* Quantizes RGBA8 levels (levels[f * nlevels + l]) to palettes of at most
 * palette_sizes = {colour endpoints, colour selectors, alpha endpoints, alpha
 * selectors} entries and writes a .crn; *out is freed with crn2_free. */
int crn2_compress_rgba(const uint8_t* const* levels, uint32_t width, uint32_t height, uint32_t nlevels,
                       uint32_t faces, uint32_t format, const uint32_t* palette_sizes, uint8_t** out,
                       uint32_t* out_size);
void crn2_free(void* p);

#ifdef __cplusplus
}
#endif

#endif /* CRN2_H */
