/*
 * crnlib.h (libcrn2) -- crnlib's public compression API with the reference's
 * C++ signatures (inc/crnlib.h), for callers that compiled against it.
 *
 * Put include/cxx first on the include path in place of the reference's inc/
 * and link libcrn2.so.  The enums and parameter structs below keep the
 * reference's names, values and member order (inc/crnlib.h:47-575), so code
 * that fills a crn_comp_params keeps compiling and the mangled names of the
 * functions match the reference's.  What runs behind each function:
 *
 *   crn_compress (both forms) ........ GPU: DDS output = the plain per-block
 *       searches of every level (crn2_encode_image_device); CRN output = the
 *       GPU dxt_hc (crn2_hc_compress_device) and the project's CRN writer
 *       (valid .crn, transcodes bit-exactly, but not byte-identical to the
 *       reference's writer).  With crn_mipmap_params, levels are generated
 *       on the GPU as mipmapped_texture::generate_mipmaps makes them
 *       (crn2_generate_mipmaps, bit-exact) when the mode asks for it.
 *       DDS below quality 255 is the clustered path (qdxt1 / qdxt5 on the
 *       GPU, crn2_compress_dds_clustered) for DXT1, DXT5, DXT5A and DXN at
 *       uber quality.  A target
 *       bitrate with DDS output, ETC formats, DXT3 in CRN, scaling, cropping
 *       and a renormalized top level return NULL.
 *   crn_decompress_crn_to_dds ........ GPU transcode of every level and face
 *       into a DDS file laid out as mipmapped_texture::write_dds writes it
 *       (crnlib/crn_mipmapped_texture.cpp:921-1140).
 *   crn_create_block_compressor / crn_compress_block / crn_free_block_compressor
 *       GPU searches, one block per call (crn2_compress_blocks for many).
 *   crn_decompress_block ............. GPU decode (crn2_decompress_blocks);
 *       DXT5 does not fall through into the DXN case as the reference does
 *       (crnlib/crnlib.cpp:485-503).
 *   crn_free_block, format helpers ... host.
 * Endpoint caching (cCRNCompFlagDisableEndpointCaching clear) is not
 * emulated: results equal the reference's with caching disabled.
 */
#ifndef CRN2_CXX_CRNLIB_H
#define CRN2_CXX_CRNLIB_H

#include <cstddef>

typedef unsigned char crn_uint8;
typedef unsigned short crn_uint16;
typedef unsigned int crn_uint32;
typedef signed char crn_int8;
typedef signed short crn_int16;
typedef signed int crn_int32;
typedef unsigned int crn_bool;

enum crn_file_type { cCRNFileTypeCRN = 0, cCRNFileTypeDDS, cCRNFileTypeForceDWORD = 0xFFFFFFFF };

enum crn_format {
    cCRNFmtInvalid = -1,
    cCRNFmtDXT1 = 0,
    cCRNFmtFirstValid = cCRNFmtDXT1,
    cCRNFmtDXT3,
    cCRNFmtDXT5,
    cCRNFmtDXT5_CCxY,
    cCRNFmtDXT5_xGxR,
    cCRNFmtDXT5_xGBR,
    cCRNFmtDXT5_AGBR,
    cCRNFmtDXN_XY,
    cCRNFmtDXN_YX,
    cCRNFmtDXT5A,
    cCRNFmtETC1,
    cCRNFmtETC2,
    cCRNFmtETC2A,
    cCRNFmtETC1S,
    cCRNFmtETC2AS,
    cCRNFmtTotal,
    cCRNFmtForceDWORD = 0xFFFFFFFF
};

enum crn_limits {
    cCRNMaxLevelResolution = 4096,
    cCRNMinPaletteSize = 8,
    cCRNMaxPaletteSize = 8192,
    cCRNMaxFaces = 6,
    cCRNMaxLevels = 16,
    cCRNMaxHelperThreads = 15,
    cCRNMinQualityLevel = 0,
    cCRNMaxQualityLevel = 255
};

enum crn_comp_flags {
    cCRNCompFlagPerceptual = 1,
    cCRNCompFlagHierarchical = 2,
    cCRNCompFlagQuick = 4,
    cCRNCompFlagUseBothBlockTypes = 8,
    cCRNCompFlagUseTransparentIndicesForBlack = 16,
    cCRNCompFlagDisableEndpointCaching = 32,
    cCRNCompFlagManualPaletteSizes = 64,
    cCRNCompFlagDXT1AForTransparency = 128,
    cCRNCompFlagGrayscaleSampling = 256,
    cCRNCompFlagDebugging = 0x80000000,
    cCRNCompFlagForceDWORD = 0xFFFFFFFF
};

enum crn_dxt_quality {
    cCRNDXTQualitySuperFast,
    cCRNDXTQualityFast,
    cCRNDXTQualityNormal,
    cCRNDXTQualityBetter,
    cCRNDXTQualityUber,
    cCRNDXTQualityTotal,
    cCRNDXTQualityForceDWORD = 0xFFFFFFFF
};

enum crn_dxt_compressor_type {
    cCRNDXTCompressorCRN,
    cCRNDXTCompressorCRNF,
    cCRNDXTCompressorRYG,
    cCRNTotalDXTCompressors,
    cCRNDXTCompressorForceDWORD = 0xFFFFFFFF
};

typedef crn_bool (*crn_progress_callback_func)(crn_uint32 phase_index, crn_uint32 total_phases,
                                               crn_uint32 subphase_index, crn_uint32 total_subphases,
                                               void* pUser_data_ptr);

struct crn_comp_params {
    crn_comp_params() { clear(); }
    void clear()
    {
        m_size_of_obj = sizeof(*this);
        m_file_type = cCRNFileTypeCRN;
        m_faces = 1;
        m_width = 0;
        m_height = 0;
        m_levels = 1;
        m_format = cCRNFmtDXT1;
        m_flags = cCRNCompFlagPerceptual | cCRNCompFlagHierarchical | cCRNCompFlagUseBothBlockTypes;
        for (crn_uint32 f = 0; f < cCRNMaxFaces; f++)
            for (crn_uint32 l = 0; l < cCRNMaxLevels; l++) m_pImages[f][l] = NULL;
        m_target_bitrate = 0.0f;
        m_quality_level = cCRNMaxQualityLevel;
        m_dxt1a_alpha_threshold = 128;
        m_dxt_quality = cCRNDXTQualityUber;
        m_dxt_compressor_type = cCRNDXTCompressorCRN;
        m_alpha_component = 3;
        m_crn_adaptive_tile_color_psnr_derating = 2.0f;
        m_crn_adaptive_tile_alpha_psnr_derating = 2.0f;
        m_crn_color_endpoint_palette_size = 0;
        m_crn_color_selector_palette_size = 0;
        m_crn_alpha_endpoint_palette_size = 0;
        m_crn_alpha_selector_palette_size = 0;
        m_num_helper_threads = 0;
        m_userdata0 = 0;
        m_userdata1 = 0;
        m_pProgress_func = NULL;
        m_pProgress_func_data = NULL;
    }
    /* the argument checks of inc/crnlib.h:322-344 */
    bool check() const
    {
        const bool pal_bad = (m_crn_color_endpoint_palette_size && (m_crn_color_endpoint_palette_size < 8 ||
                                                                    m_crn_color_endpoint_palette_size > 8192)) ||
                             (m_crn_color_selector_palette_size && (m_crn_color_selector_palette_size < 8 ||
                                                                    m_crn_color_selector_palette_size > 8192)) ||
                             (m_crn_alpha_endpoint_palette_size && (m_crn_alpha_endpoint_palette_size < 8 ||
                                                                    m_crn_alpha_endpoint_palette_size > 8192)) ||
                             (m_crn_alpha_selector_palette_size && (m_crn_alpha_selector_palette_size < 8 ||
                                                                    m_crn_alpha_selector_palette_size > 8192));
        return !(m_file_type > cCRNFileTypeDDS || m_quality_level > cCRNMaxQualityLevel ||
                 m_dxt1a_alpha_threshold > 255 || (m_faces != 1 && m_faces != 6) || m_width < 1 ||
                 m_width > cCRNMaxLevelResolution || m_height < 1 || m_height > cCRNMaxLevelResolution ||
                 m_levels < 1 || m_levels > cCRNMaxLevels || m_format < cCRNFmtDXT1 || m_format >= cCRNFmtTotal ||
                 pal_bad || m_alpha_component > 3 || m_num_helper_threads > cCRNMaxHelperThreads ||
                 m_dxt_quality > cCRNDXTQualityUber || m_dxt_compressor_type >= cCRNTotalDXTCompressors);
    }
    bool get_flag(crn_comp_flags flag) const { return (m_flags & flag) != 0; }
    void set_flag(crn_comp_flags flag, bool val)
    {
        m_flags &= ~flag;
        if (val) m_flags |= flag;
    }

    crn_uint32 m_size_of_obj;
    crn_file_type m_file_type;
    crn_uint32 m_faces;
    crn_uint32 m_width;
    crn_uint32 m_height;
    crn_uint32 m_levels;
    crn_format m_format;
    crn_uint32 m_flags;
    const crn_uint32* m_pImages[cCRNMaxFaces][cCRNMaxLevels];
    float m_target_bitrate;
    crn_uint32 m_quality_level;
    crn_uint32 m_dxt1a_alpha_threshold;
    crn_dxt_quality m_dxt_quality;
    crn_dxt_compressor_type m_dxt_compressor_type;
    crn_uint32 m_alpha_component;
    float m_crn_adaptive_tile_color_psnr_derating;
    float m_crn_adaptive_tile_alpha_psnr_derating;
    crn_uint32 m_crn_color_endpoint_palette_size;
    crn_uint32 m_crn_color_selector_palette_size;
    crn_uint32 m_crn_alpha_endpoint_palette_size;
    crn_uint32 m_crn_alpha_selector_palette_size;
    crn_uint32 m_num_helper_threads;
    crn_uint32 m_userdata0;
    crn_uint32 m_userdata1;
    crn_progress_callback_func m_pProgress_func;
    void* m_pProgress_func_data;
};

enum crn_mip_mode {
    cCRNMipModeUseSourceOrGenerateMips,
    cCRNMipModeUseSourceMips,
    cCRNMipModeGenerateMips,
    cCRNMipModeNoMips,
    cCRNMipModeTotal,
    cCRNModeForceDWORD = 0xFFFFFFFF
};

enum crn_mip_filter {
    cCRNMipFilterBox,
    cCRNMipFilterTent,
    cCRNMipFilterLanczos4,
    cCRNMipFilterMitchell,
    cCRNMipFilterKaiser,
    cCRNMipFilterTotal,
    cCRNMipFilterForceDWORD = 0xFFFFFFFF
};

enum crn_scale_mode {
    cCRNSMDisabled,
    cCRNSMAbsolute,
    cCRNSMRelative,
    cCRNSMLowerPow2,
    cCRNSMNearestPow2,
    cCRNSMNextPow2,
    cCRNSMTotal,
    cCRNSMForceDWORD = 0xFFFFFFFF
};

struct crn_mipmap_params {
    crn_mipmap_params() { clear(); }
    void clear()
    {
        m_size_of_obj = sizeof(*this);
        m_mode = cCRNMipModeUseSourceOrGenerateMips;
        m_filter = cCRNMipFilterKaiser;
        m_gamma_filtering = true;
        m_gamma = 2.2f;
        m_blurriness = .9f;
        m_renormalize = false;
        m_rtopmip = false;
        m_tiled = false;
        m_max_levels = cCRNMaxLevels;
        m_min_mip_size = 1;
        m_scale_mode = cCRNSMDisabled;
        m_scale_x = 1.0f;
        m_scale_y = 1.0f;
        m_window_left = 0;
        m_window_top = 0;
        m_window_right = 0;
        m_window_bottom = 0;
        m_clamp_scale = false;
        m_clamp_width = 0;
        m_clamp_height = 0;
    }
    bool check() const { return true; }

    crn_uint32 m_size_of_obj;
    crn_mip_mode m_mode;
    crn_mip_filter m_filter;
    crn_bool m_gamma_filtering;
    float m_gamma;
    float m_blurriness;
    crn_uint32 m_max_levels;
    crn_uint32 m_min_mip_size;
    crn_bool m_renormalize;
    crn_bool m_rtopmip;
    crn_bool m_tiled;
    crn_scale_mode m_scale_mode;
    float m_scale_x;
    float m_scale_y;
    crn_uint32 m_window_left;
    crn_uint32 m_window_top;
    crn_uint32 m_window_right;
    crn_uint32 m_window_bottom;
    crn_bool m_clamp_scale;
    crn_uint32 m_clamp_width;
    crn_uint32 m_clamp_height;
};

typedef void* (*crn_realloc_func)(void* p, size_t size, size_t* pActual_size, bool movable, void* pUser_data);
typedef size_t (*crn_msize_func)(void* p, void* pUser_data);
void crn_set_memory_callbacks(crn_realloc_func pRealloc, crn_msize_func pMSize, void* pUser_data);

void crn_free_block(void* pBlock);

void* crn_compress(const crn_comp_params& comp_params, crn_uint32& compressed_size,
                   crn_uint32* pActual_quality_level = NULL, float* pActual_bitrate = NULL);
void* crn_compress(const crn_comp_params& comp_params, const crn_mipmap_params& mip_params,
                   crn_uint32& compressed_size, crn_uint32* pActual_quality_level = NULL,
                   float* pActual_bitrate = NULL);

void* crn_decompress_crn_to_dds(const void* pCRN_file_data, crn_uint32& file_size);

crn_uint32 crn_get_format_fourcc(crn_format fmt);
crn_uint32 crn_get_format_bits_per_texel(crn_format fmt);
crn_uint32 crn_get_bytes_per_dxt_block(crn_format fmt);
crn_format crn_get_fundamental_dxt_format(crn_format fmt);

const char* crn_get_file_type_ext(crn_file_type file_type);
const char* crn_get_format_string(crn_format fmt);
const char* crn_get_dxt_quality_string(crn_dxt_quality q);

typedef void* crn_block_compressor_context_t;
crn_block_compressor_context_t crn_create_block_compressor(const crn_comp_params& params);
void crn_compress_block(crn_block_compressor_context_t pContext, const crn_uint32* pPixels, void* pDst_block);
void crn_free_block_compressor(crn_block_compressor_context_t pContext);
bool crn_decompress_block(const void* pSrc_block, crn_uint32* pDst_pixels, crn_format crn_fmt);

#define CRNLIB_VERSION 104

#endif  // CRN2_CXX_CRNLIB_H
