// crnlib_facade.cpp -- the reference's C++ entry points (crnd:: transcoder,
// crnlib's compression API, the two endpoint optimizer classes) with their
// mangled names, each a thin layer over the C ABI of crn2.h; declarations in
// include/cxx.  Error convention of the reference: bool / NULL, no throw.
//
//   crnd_* ..................... inc/crn_defs.h:81-221, inc/crn_decomp.h:2567-2868, 4404-4487
//   crn_compress, crn_decompress_crn_to_dds, crn_free_block, block API,
//   crn_decompress_block, format and string helpers ... inc/crnlib.h:586-690,
//                                crnlib/crnlib.cpp:72-543
//   dxt1_endpoint_optimizer::compute  crnlib/crn_dxt1.h:223, crn_dxt1.cpp:2234-2246
//   dxt5_endpoint_optimizer::compute  crnlib/crn_dxt5a.h:72, crn_dxt5a.cpp:40-196
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/cxx/crn_decomp.h"
#include "../../include/cxx/crn_dxt_optimizers.h"
#include "../../include/cxx/crnlib.h"
#include "crn2_internal.h"
#include "crnd_host.h"

namespace {
crnd::crnd_realloc_func g_crnd_realloc = nullptr;
crnd::crnd_msize_func g_crnd_msize = nullptr;
void* g_crnd_user = nullptr;
crn_realloc_func g_crn_realloc = nullptr;
crn_msize_func g_crn_msize = nullptr;
void* g_crn_user = nullptr;

bool is_dxt(crn_format f) { return f >= cCRNFmtDXT1 && f <= cCRNFmtDXT5A; }
}  // namespace

namespace crnd {

void crnd_set_memory_callbacks(crnd_realloc_func pRealloc, crnd_msize_func pMSize, void* pUser_data)
{
    // recorded; the context's host memory comes from the C++ allocator
    if (!pRealloc || !pMSize) {
        g_crnd_realloc = nullptr;
        g_crnd_msize = nullptr;
        g_crnd_user = nullptr;
        return;
    }
    g_crnd_realloc = pRealloc;
    g_crnd_msize = pMSize;
    g_crnd_user = pUser_data;
}

uint32 crnd_crn_format_to_fourcc(crn_format fmt)
{
    auto cc = [](char a, char b, char c, char d) {
        return (uint32)(uint8)a | ((uint32)(uint8)b << 8) | ((uint32)(uint8)c << 16) | ((uint32)(uint8)d << 24);
    };
    switch (fmt) {
        case cCRNFmtDXT1: return cc('D', 'X', 'T', '1');
        case cCRNFmtDXT3: return cc('D', 'X', 'T', '3');
        case cCRNFmtDXT5: return cc('D', 'X', 'T', '5');
        case cCRNFmtDXN_XY: return cc('A', '2', 'X', 'Y');
        case cCRNFmtDXN_YX: return cc('A', 'T', 'I', '2');
        case cCRNFmtDXT5A: return cc('A', 'T', 'I', '1');
        case cCRNFmtDXT5_CCxY: return cc('C', 'C', 'x', 'Y');
        case cCRNFmtDXT5_xGxR: return cc('x', 'G', 'x', 'R');
        case cCRNFmtDXT5_xGBR: return cc('x', 'G', 'B', 'R');
        case cCRNFmtDXT5_AGBR: return cc('A', 'G', 'B', 'R');
        case cCRNFmtETC1: return cc('E', 'T', 'C', '1');
        case cCRNFmtETC2: return cc('E', 'T', 'C', '2');
        case cCRNFmtETC2A: return cc('E', 'T', '2', 'A');
        case cCRNFmtETC1S: return cc('E', 'T', '1', 'S');
        case cCRNFmtETC2AS: return cc('E', '2', 'A', 'S');
        default: return 0;
    }
}

crn_format crnd_get_fundamental_dxt_format(crn_format fmt)
{
    switch (fmt) {
        case cCRNFmtDXT5_CCxY: case cCRNFmtDXT5_xGxR: case cCRNFmtDXT5_xGBR: case cCRNFmtDXT5_AGBR:
            return cCRNFmtDXT5;
        default: return fmt;
    }
}

uint32 crnd_get_crn_format_bits_per_texel(crn_format fmt)
{
    switch (fmt) {
        case cCRNFmtDXT1: case cCRNFmtDXT5A: case cCRNFmtETC1: case cCRNFmtETC2: case cCRNFmtETC1S:
            return 4;
        case cCRNFmtDXT3: case cCRNFmtDXT5: case cCRNFmtDXN_XY: case cCRNFmtDXN_YX: case cCRNFmtDXT5_CCxY:
        case cCRNFmtDXT5_xGxR: case cCRNFmtDXT5_xGBR: case cCRNFmtDXT5_AGBR: case cCRNFmtETC2A: case cCRNFmtETC2AS:
            return 8;
        default: return 0;
    }
}

uint32 crnd_get_bytes_per_dxt_block(crn_format fmt) { return (crnd_get_crn_format_bits_per_texel(fmt) << 4) >> 3; }

bool crnd_validate_file(const void* pData, uint32 data_size, crn_file_info* pFile_info)
{
    if (pFile_info) {
        if (pFile_info->m_struct_size != sizeof(crn_file_info)) return false;
        memset(reinterpret_cast<uint8*>(pFile_info) + sizeof(uint32), 0, sizeof(crn_file_info) - sizeof(uint32));
    }
    if (!pData || !crn2_validate_file(pData, data_size)) return false;
    if (pFile_info) {
        crn2::Header h;
        if (!crn2::parse_header(static_cast<const uint8_t*>(pData), data_size, h)) return false;
        pFile_info->m_actual_data_size = h.data_size;
        pFile_info->m_header_size = h.header_size;
        pFile_info->m_total_palette_size = h.pal_size[0] + h.pal_size[1] + h.pal_size[2] + h.pal_size[3];
        pFile_info->m_tables_size = h.tables_size;
        pFile_info->m_levels = h.levels;
        for (uint32 i = 0; i < h.levels && i < (uint32)cCRNMaxLevels; i++) {
            const uint32 next = (i + 1 < h.levels) ? h.level_ofs[i + 1] : h.data_size;
            pFile_info->m_level_compressed_size[i] = next - h.level_ofs[i];
        }
        pFile_info->m_color_endpoint_palette_entries = h.pal_num[0];
        pFile_info->m_color_selector_palette_entries = h.pal_num[1];
        pFile_info->m_alpha_endpoint_palette_entries = h.pal_num[2];
        pFile_info->m_alpha_selector_palette_entries = h.pal_num[3];
    }
    return true;
}

bool crnd_get_texture_info(const void* pData, uint32 data_size, crn_texture_info* pInfo)
{
    if (!pData || !pInfo || pInfo->m_struct_size != sizeof(crn_texture_info)) return false;
    crn2_texture_info t;
    if (crn2_get_texture_info(pData, data_size, &t) != CRN2_OK) return false;
    pInfo->m_width = t.width;
    pInfo->m_height = t.height;
    pInfo->m_levels = t.levels;
    pInfo->m_faces = t.faces;
    pInfo->m_bytes_per_block = t.bytes_per_block;
    pInfo->m_userdata0 = t.userdata0;
    pInfo->m_userdata1 = t.userdata1;
    pInfo->m_format = static_cast<crn_format>(t.format);
    return true;
}

bool crnd_get_level_info(const void* pData, uint32 data_size, uint32 level_index, crn_level_info* pLevel_info)
{
    if (!pData || !pLevel_info || pLevel_info->m_struct_size != sizeof(crn_level_info)) return false;
    crn2_level_info li;
    if (crn2_get_level_info(pData, data_size, level_index, &li) != CRN2_OK) return false;
    pLevel_info->m_width = li.width;
    pLevel_info->m_height = li.height;
    pLevel_info->m_faces = li.faces;
    pLevel_info->m_blocks_x = li.blocks_x;
    pLevel_info->m_blocks_y = li.blocks_y;
    pLevel_info->m_bytes_per_block = li.bytes_per_block;
    pLevel_info->m_format = static_cast<crn_format>(li.format);
    return true;
}

crnd_unpack_context crnd_unpack_begin(const void* pData, uint32 data_size)
{
    return crn2_unpack_begin(pData, data_size);
}

bool crnd_get_data(crnd_unpack_context pContext, const void** ppData, uint32* pData_size)
{
    return crn2_get_data(static_cast<crn2_unpack_ctx*>(pContext), ppData, pData_size) == CRN2_OK;
}

bool crnd_unpack_level(crnd_unpack_context pContext, void** ppDst, uint32 dst_size_in_bytes,
                       uint32 row_pitch_in_bytes, uint32 level_index)
{
    return crn2_unpack_level(static_cast<crn2_unpack_ctx*>(pContext), ppDst, dst_size_in_bytes, row_pitch_in_bytes,
                             level_index) == CRN2_OK;
}

bool crnd_unpack_level_segmented(crnd_unpack_context pContext, const void* pSrc, uint32 src_size_in_bytes,
                                 void** ppDst, uint32 dst_size_in_bytes, uint32 row_pitch_in_bytes,
                                 uint32 level_index)
{
    return crn2_unpack_level_segmented(static_cast<crn2_unpack_ctx*>(pContext), pSrc, src_size_in_bytes, ppDst,
                                       dst_size_in_bytes, row_pitch_in_bytes, level_index) == CRN2_OK;
}

bool crnd_unpack_end(crnd_unpack_context pContext)
{
    return pContext && crn2_unpack_end(static_cast<crn2_unpack_ctx*>(pContext)) == CRN2_OK;
}

const void* crnd_get_level_data(const void* pData, uint32 data_size, uint32 level_index, uint32* pSize)
{
    return crn2_get_level_data(pData, data_size, level_index, pSize);
}

uint32 crnd_get_segmented_file_size(const void* pData, uint32 data_size)
{
    crn2::Header h;
    if (!pData || !crn2::parse_header(static_cast<const uint8_t*>(pData), data_size, h)) return 0;
    uint32 size = h.header_size;
    for (int i = 0; i < 4; i++) size = std::max(size, h.pal_ofs[i] + h.pal_size[i]);
    return std::max(size, h.tables_ofs + h.tables_size);
}

bool crnd_create_segmented_file(const void* pData, uint32 data_size, void* pBase_data, uint base_data_size)
{
    crn2::Header h;
    if (!pData || !pBase_data || !crn2::parse_header(static_cast<const uint8_t*>(pData), data_size, h)) return false;
    if (h.flags & 1) return false;   // already segmented
    const uint32 n = crnd_get_segmented_file_size(pData, data_size);
    if (base_data_size < n || n > data_size) return false;
    uint8* b = static_cast<uint8*>(pBase_data);
    memmove(b, pData, n);
    // header fields are big-endian: flags at 19, data size at 6, CRCs at 10 (data) and 4 (header)
    const uint32 flags = h.flags | 1u;
    b[19] = (uint8)(flags >> 8);
    b[20] = (uint8)flags;
    b[6] = (uint8)(n >> 24);
    b[7] = (uint8)(n >> 16);
    b[8] = (uint8)(n >> 8);
    b[9] = (uint8)n;
    const uint16 dcrc = crn2_crc16(b + h.header_size, n - h.header_size, 0);
    b[10] = (uint8)(dcrc >> 8);
    b[11] = (uint8)dcrc;
    const uint16 hcrc = crn2_crc16(b + 6, h.header_size - 6, 0);
    b[4] = (uint8)(hcrc >> 8);
    b[5] = (uint8)hcrc;
    return true;
}

}  // namespace crnd

// ------------------------------------------------------------------ crnlib

void crn_set_memory_callbacks(crn_realloc_func pRealloc, crn_msize_func pMSize, void* pUser_data)
{
    // recorded; blocks returned by this library come from malloc and go back through crn_free_block
    g_crn_realloc = pRealloc;
    g_crn_msize = pMSize;
    g_crn_user = pUser_data;
}

void crn_free_block(void* pBlock) { crn2_free(pBlock); }

crn_uint32 crn_get_format_fourcc(crn_format fmt) { return crnd::crnd_crn_format_to_fourcc(fmt); }
crn_uint32 crn_get_format_bits_per_texel(crn_format fmt) { return crnd::crnd_get_crn_format_bits_per_texel(fmt); }
crn_uint32 crn_get_bytes_per_dxt_block(crn_format fmt) { return crnd::crnd_get_bytes_per_dxt_block(fmt); }
crn_format crn_get_fundamental_dxt_format(crn_format fmt) { return crnd::crnd_get_fundamental_dxt_format(fmt); }

const char* crn_get_file_type_ext(crn_file_type file_type)
{
    switch (file_type) {
        case cCRNFileTypeDDS: return "dds";
        case cCRNFileTypeCRN: return "crn";
        default: return "?";
    }
}

const char* crn_get_format_string(crn_format fmt)
{
    switch (fmt) {
        case cCRNFmtDXT1: return "DXT1";
        case cCRNFmtDXT3: return "DXT3";
        case cCRNFmtDXT5: return "DXT5";
        case cCRNFmtDXT5_CCxY: return "DXT5_CCxY";
        case cCRNFmtDXT5_xGBR: return "DXT5_xGBR";
        case cCRNFmtDXT5_AGBR: return "DXT5_AGBR";
        case cCRNFmtDXT5_xGxR: return "DXT5_xGxR";
        case cCRNFmtDXN_XY: return "DXN_XY";
        case cCRNFmtDXN_YX: return "DXN_YX";
        case cCRNFmtDXT5A: return "DXT5A";
        case cCRNFmtETC1: return "ETC1";
        case cCRNFmtETC2: return "ETC2";
        case cCRNFmtETC2A: return "ETC2A";
        case cCRNFmtETC1S: return "ETC1S";
        case cCRNFmtETC2AS: return "ETC2AS";
        default: return "?";
    }
}

const char* crn_get_dxt_quality_string(crn_dxt_quality q)
{
    switch (q) {
        case cCRNDXTQualitySuperFast: return "SuperFast";
        case cCRNDXTQualityFast: return "Fast";
        case cCRNDXTQualityNormal: return "Normal";
        case cCRNDXTQualityBetter: return "Better";
        case cCRNDXTQualityUber: return "Uber";
        default: return "?";
    }
}

namespace {

// crn2 encoder flags of a crn_comp_params: bit0 perceptual, bit1 both block
// types, bit2 grayscale sampling, bit3 DXT1A for transparency with the alpha
// threshold in bits 8..15, bit4 transparent indices for black (DXT1 only)
int encoder_flags(const crn_comp_params& p)
{
    int f = (p.get_flag(cCRNCompFlagPerceptual) ? 1 : 0) | (p.get_flag(cCRNCompFlagUseBothBlockTypes) ? 2 : 0) |
            (p.get_flag(cCRNCompFlagGrayscaleSampling) ? 4 : 0);
    if (crnd::crnd_get_fundamental_dxt_format(p.m_format) == cCRNFmtDXT1) {
        if (p.get_flag(cCRNCompFlagDXT1AForTransparency)) f |= 8 | (int)((p.m_dxt1a_alpha_threshold & 255) << 8);
        if (p.get_flag(cCRNCompFlagUseTransparentIndicesForBlack)) f |= 16;
    }
    return f;
}

// the parameter combinations the GPU path covers (see include/cxx/crnlib.h)
bool gpu_supported(const crn_comp_params& p)
{
    return is_dxt(p.m_format) && p.m_dxt1a_alpha_threshold <= 255;
}

// palette sizes of crn_comp::quantize_images (crnlib/crn_comp.cpp:527-577)
void palette_sizes(const crn_comp_params& p, uint32_t* out)
{
    if (p.get_flag(cCRNCompFlagManualPaletteSizes)) {
        const crn_uint32 v[4] = {p.m_crn_color_endpoint_palette_size, p.m_crn_color_selector_palette_size,
                                 p.m_crn_alpha_endpoint_palette_size, p.m_crn_alpha_selector_palette_size};
        for (int i = 0; i < 4; i++) out[i] = std::min<uint32_t>(std::max<uint32_t>(v[i], 8), 8192);
        return;
    }
    uint32_t max_entries = ((p.m_width + 3) / 4) * ((p.m_height + 3) / 4);
    max_entries = std::min<uint32_t>(std::max<uint32_t>(max_entries, 8), 8192);
    const float quality = std::min(std::max((float)p.m_quality_level / 255.0f, 0.0f), 1.0f);
    float cmul = 1.0f, amul = 1.0f;
    if (p.m_format == cCRNFmtDXT5_CCxY) {
        cmul = 3.5f;
        amul = .35f;
    } else if (p.m_format == cCRNFmtDXT5) {
        cmul = .75f;
    }
    auto size = [&](float lo, float q) {
        const float v = .5f + (lo + ((float)max_entries - lo) * q);
        const uint32_t u = v <= 0.0f ? 0u : (uint32_t)v;
        return std::min<uint32_t>(std::max<uint32_t>(u, 8), 8192);
    };
    out[0] = size(64.0f, powf(quality, 1.8f * cmul));
    out[1] = size(96.0f, powf(quality, 1.65f * cmul));
    out[2] = size(24.0f, powf(quality, 2.1f * amul));
    out[3] = size(48.0f, powf(quality, 1.65f * amul));
}

void* compress_levels(const crn_comp_params& p, crn_uint32 levels, crn_uint32& compressed_size,
                      crn_uint32* pActual_quality_level, float* pActual_bitrate)
{
    std::vector<const uint8_t*> lv((size_t)p.m_faces * levels);
    for (crn_uint32 f = 0; f < p.m_faces; f++)
        for (crn_uint32 l = 0; l < levels; l++) {
            if (!p.m_pImages[f][l]) return nullptr;
            lv[f * levels + l] = reinterpret_cast<const uint8_t*>(p.m_pImages[f][l]);
        }
    uint8_t* out = nullptr;
    uint32_t size = 0;
    int rc;
    if (p.m_file_type == cCRNFileTypeDDS) {
        // a bitrate target (dds_comp's quality search, crn_dds_comp.cpp:279-370) is not covered
        if (p.m_target_bitrate > 0.0f) return nullptr;
        if (p.m_quality_level < cCRNMaxQualityLevel) {
            // clustered DDS (crn_dds_comp.cpp:158-205)
            const int f = (p.get_flag(cCRNCompFlagPerceptual) ? 1 : 0) |
                          (p.get_flag(cCRNCompFlagUseBothBlockTypes) ? 2 : 0) |
                          (p.get_flag(cCRNCompFlagHierarchical) ? 4 : 0);
            rc = crn2_compress_dds_clustered(lv.data(), p.m_width, p.m_height, levels, p.m_faces,
                                             (uint32_t)p.m_format, p.m_quality_level, (int)p.m_dxt_quality, f,
                                             p.m_dxt1a_alpha_threshold, &out, &size);
        } else {
            rc = crn2_compress_dds(lv.data(), p.m_width, p.m_height, levels, p.m_faces, (uint32_t)p.m_format,
                                   (int)p.m_dxt_quality, encoder_flags(p), &out, &size);
        }
    } else {
        uint32_t cb[4];
        palette_sizes(p, cb);
        const float tile[3] = {p.m_crn_adaptive_tile_color_psnr_derating, p.m_crn_adaptive_tile_alpha_psnr_derating,
                               3.0f};
        rc = crn2_compress_crn(lv.data(), p.m_width, p.m_height, levels, p.m_faces, (uint32_t)p.m_format,
                               p.get_flag(cCRNCompFlagPerceptual) ? 1 : 0, cb, tile, p.m_userdata0, p.m_userdata1,
                               &out, &size);
    }
    if (rc != CRN2_OK || !out) return nullptr;
    compressed_size = size;
    if (pActual_quality_level) *pActual_quality_level = p.m_quality_level;
    if (pActual_bitrate) {
        double texels = 0;
        for (crn_uint32 l = 0; l < levels; l++)
            texels += (double)std::max<crn_uint32>(1, p.m_width >> l) * std::max<crn_uint32>(1, p.m_height >> l);
        *pActual_bitrate = (float)(size * 8.0 / (texels * p.m_faces));
    }
    return out;
}

}  // namespace

void* crn_compress(const crn_comp_params& comp_params, crn_uint32& compressed_size, crn_uint32* pActual_quality_level,
                   float* pActual_bitrate)
{
    compressed_size = 0;
    if (pActual_quality_level) *pActual_quality_level = 0;
    if (pActual_bitrate) *pActual_bitrate = 0.0f;
    if (!comp_params.check() || !gpu_supported(comp_params)) return nullptr;
    return compress_levels(comp_params, comp_params.m_levels, compressed_size, pActual_quality_level,
                           pActual_bitrate);
}

void* crn_compress(const crn_comp_params& comp_params, const crn_mipmap_params& mip_params,
                   crn_uint32& compressed_size, crn_uint32* pActual_quality_level, float* pActual_bitrate)
{
    compressed_size = 0;
    if (pActual_quality_level) *pActual_quality_level = 0;
    if (pActual_bitrate) *pActual_bitrate = 0.0f;
    if (!comp_params.check() || !mip_params.check() || !gpu_supported(comp_params)) return nullptr;
    // scaling, cropping and the renormalized top level are not built
    if (mip_params.m_scale_mode != cCRNSMDisabled || mip_params.m_clamp_scale || mip_params.m_window_right ||
        mip_params.m_window_bottom || (mip_params.m_renormalize && mip_params.m_rtopmip))
        return nullptr;
    crn_uint32 levels = comp_params.m_levels;
    bool generate = false;
    switch (mip_params.m_mode) {
        case cCRNMipModeNoMips: levels = 1; break;
        case cCRNMipModeUseSourceMips: break;
        case cCRNMipModeUseSourceOrGenerateMips: generate = levels == 1; break;
        case cCRNMipModeGenerateMips: generate = true; break;
        default: return nullptr;
    }
    if (!generate) {
        levels = std::min<crn_uint32>(levels, std::max<crn_uint32>(1, mip_params.m_max_levels));
        return compress_levels(comp_params, levels, compressed_size, pActual_quality_level, pActual_bitrate);
    }
    // mipmapped_texture::generate_mipmaps from level 0 of every face on the GPU
    // (crn_texture_comp.cpp:546-572): alpha is resampled when any given texel
    // is below 255 (crn_texture_comp.cpp:298-323); the resampler's source gamma
    // stays 2.2 whatever m_gamma says; crnlib's multithreaded resampler (X then
    // Y on every level) runs when helper threads are asked for
    bool alpha = false;
    for (crn_uint32 f = 0; f < comp_params.m_faces && !alpha; f++)
        for (crn_uint32 l = 0; l < comp_params.m_levels && !alpha; l++) {
            const crn_uint32* px = comp_params.m_pImages[f][l];
            if (!px) return nullptr;
            const size_t n = (size_t)std::max<crn_uint32>(1, comp_params.m_width >> l) *
                             std::max<crn_uint32>(1, comp_params.m_height >> l);
            for (size_t i = 0; i < n && !alpha; i++) alpha = (px[i] >> 24) < 255;
        }
    crn2_mip_params mp;
    mp.filter = (uint32_t)mip_params.m_filter;
    mp.flags = (mip_params.m_gamma_filtering ? CRN2_MIP_SRGB : 0) | (mip_params.m_tiled ? CRN2_MIP_WRAP : 0) |
               (alpha ? CRN2_MIP_ALPHA : 0) | (mip_params.m_renormalize ? CRN2_MIP_RENORMALIZE : 0) |
               (comp_params.m_num_helper_threads > 0 ? CRN2_MIP_THREADED : 0);
    mp.filter_scale = mip_params.m_blurriness;
    mp.source_gamma = 2.2f;
    mp.max_levels = mip_params.m_max_levels;
    mp.min_mip_size = mip_params.m_min_mip_size;
    levels = crn2_mip_level_count(comp_params.m_width, comp_params.m_height, mp.min_mip_size, mp.max_levels);
    if (!levels || levels > cCRNMaxLevels) return nullptr;
    uint64_t chain = 0;
    for (crn_uint32 l = 0; l < levels; l++)
        chain += 4ull * std::max<crn_uint32>(1, comp_params.m_width >> l) *
                 std::max<crn_uint32>(1, comp_params.m_height >> l);
    std::vector<std::vector<uint8_t>> store(comp_params.m_faces, std::vector<uint8_t>(chain));
    crn_comp_params q(comp_params);
    q.m_levels = levels;
    for (crn_uint32 f = 0; f < comp_params.m_faces; f++) {
        if (crn2_generate_mipmaps(comp_params.m_pImages[f][0], comp_params.m_width, comp_params.m_height, &mp,
                                  store[f].data(), chain) != CRN2_OK)
            return nullptr;
        uint64_t ofs = 0;
        for (crn_uint32 l = 0; l < levels; l++) {
            q.m_pImages[f][l] = reinterpret_cast<const crn_uint32*>(store[f].data() + ofs);
            ofs += 4ull * std::max<crn_uint32>(1, comp_params.m_width >> l) *
                   std::max<crn_uint32>(1, comp_params.m_height >> l);
        }
    }
    return compress_levels(q, levels, compressed_size, pActual_quality_level, pActual_bitrate);
}

void* crn_decompress_crn_to_dds(const void* pCRN_file_data, crn_uint32& file_size)
{
    uint8_t* out = nullptr;
    uint32_t size = 0;
    if (crn2_crn_to_dds(pCRN_file_data, file_size, &out, &size) != CRN2_OK) {
        file_size = 0;
        return nullptr;
    }
    file_size = size;
    return out;
}

namespace {
struct BlockCompressor {
    uint32_t format;
    int quality, flags;
};
}  // namespace

crn_block_compressor_context_t crn_create_block_compressor(const crn_comp_params& params)
{
    const crn_format f = crnd::crnd_get_fundamental_dxt_format(params.m_format);
    if (!is_dxt(f) || !gpu_supported(params)) return nullptr;
    BlockCompressor* c = new (std::nothrow) BlockCompressor;
    if (!c) return nullptr;
    c->format = (uint32_t)f;
    c->quality = (int)params.m_dxt_quality;
    // 3-colour DXT1 blocks only for DXT1 itself (crn_dxt_image.cpp:1464-1467)
    c->flags = encoder_flags(params);
    return c;
}

void crn_compress_block(crn_block_compressor_context_t pContext, const crn_uint32* pPixels, void* pDst_block)
{
    const BlockCompressor* c = static_cast<const BlockCompressor*>(pContext);
    if (c && pPixels && pDst_block) (void)crn2_compress_blocks(pPixels, 1, c->format, c->quality, c->flags, pDst_block);
}

void crn_free_block_compressor(crn_block_compressor_context_t pContext)
{
    delete static_cast<BlockCompressor*>(pContext);
}

bool crn_decompress_block(const void* pSrc_block, crn_uint32* pDst_pixels, crn_format crn_fmt)
{
    if (!pSrc_block || !pDst_pixels || !is_dxt(crn_fmt)) return false;
    return crn2_decompress_blocks(pSrc_block, 1, (uint32_t)crn_fmt, pDst_pixels) == CRN2_OK;
}

// --------------------------------------------------- endpoint optimizers

namespace crnlib {

namespace {
// crn2_dxt1_blocks / crn2_dxt1_clusters result record (24 bytes)
struct Dxt1Rec {
    uint64_t error;
    uint32_t lo_hi, selectors_or_unique, flags, pad;
};
static_assert(sizeof(Dxt1Rec) == 24, "record");

bool dxt1_block_path(const dxt1_endpoint_optimizer::params& p)
{
    return p.m_num_pixels == 16 && p.m_quality <= cCRNDXTQualityUber && p.m_dxt1a_alpha_threshold <= 255;
}

bool dxt1_cluster_path(const dxt1_endpoint_optimizer::params& p)
{
    return p.m_num_pixels > 0 && p.m_quality == cCRNDXTQualityUber && !p.m_pixels_have_alpha &&
           !p.m_use_transparent_indices_for_black && !p.m_use_alpha_blocks && !p.m_grayscale_sampling &&
           !p.m_force_alpha_blocks;
}

int dxt1_flags(const dxt1_endpoint_optimizer::params& p)
{
    return (p.m_perceptual ? 1 : 0) | (p.m_use_alpha_blocks ? 2 : 0) | (p.m_grayscale_sampling ? 4 : 0) |
           (p.m_force_alpha_blocks ? 8 : 0) | (p.m_pixels_have_alpha ? 16 : 0) |
           (p.m_use_transparent_indices_for_black ? 32 : 0) | (int)((p.m_dxt1a_alpha_threshold & 255) << 8);
}

void dxt1_fill(const Dxt1Rec& o, bool block, dxt1_endpoint_optimizer::results& r)
{
    r.m_error = o.error;
    r.m_low_color = (uint16)(o.lo_hi & 0xFFFF);
    r.m_high_color = (uint16)(o.lo_hi >> 16);
    if (block) {
        r.m_alpha_block = o.flags & 1;
        r.m_reordered = (o.flags >> 1) & 1;
    } else {
        r.m_alpha_block = false;
        r.m_reordered = o.flags & 1;
    }
    r.m_alternate_rounding = (o.flags >> 2) & 1;
    r.m_enforce_selector = (o.flags >> 3) & 1;
    r.m_enforced_selector = (uint8)((o.flags >> 4) & 3);
}
}  // namespace

dxt1_endpoint_optimizer::dxt1_endpoint_optimizer() : m_reserved(nullptr) {}

bool dxt1_endpoint_optimizer::compute(const params& p, results& r) { return compute_batch(&p, &r, 1); }

bool dxt1_endpoint_optimizer::compute_batch(const params* p, results* r, uint n)
{
    if (!p || !r) return false;
    if (!n) return true;
    for (uint i = 0; i < n; i++)
        if (!p[i].m_pPixels) return false;
    const bool block = dxt1_block_path(p[0]);
    if (!block && !dxt1_cluster_path(p[0])) return false;
    for (uint i = 1; i < n; i++)
        if (dxt1_flags(p[i]) != dxt1_flags(p[0]) || p[i].m_quality != p[0].m_quality ||
            (block ? !dxt1_block_path(p[i]) : !dxt1_cluster_path(p[i])))
            return false;
    std::vector<Dxt1Rec> out(n);
    if (block) {
        std::vector<uint32_t> px((size_t)n * 16);
        for (uint i = 0; i < n; i++) memcpy(&px[(size_t)i * 16], p[i].m_pPixels, 64);
        if (crn2_dxt1_blocks(px.data(), n, (int)p[0].m_quality, dxt1_flags(p[0]), out.data()) != CRN2_OK)
            return false;
        for (uint i = 0; i < n; i++) {
            dxt1_fill(out[i], true, r[i]);
            if (r[i].m_pSelectors)
                for (int t = 0; t < 16; t++) r[i].m_pSelectors[t] = (uint8)((out[i].selectors_or_unique >> (2 * t)) & 3);
        }
        return true;
    }
    std::vector<uint32_t> ofs(n + 1, 0);
    for (uint i = 0; i < n; i++) ofs[i + 1] = ofs[i] + p[i].m_num_pixels;
    std::vector<uint32_t> px(ofs[n]);
    for (uint i = 0; i < n; i++) memcpy(&px[ofs[i]], p[i].m_pPixels, (size_t)p[i].m_num_pixels * 4);
    std::vector<uint8_t> sel(ofs[n]);
    if (crn2_dxt1_clusters(px.data(), ofs.data(), n, p[0].m_perceptual ? 1 : 0, out.data(), sel.data()) != CRN2_OK)
        return false;
    for (uint i = 0; i < n; i++) {
        if (out[i].flags & 256) return false;   // more distinct colours than the one-wave search holds
        dxt1_fill(out[i], false, r[i]);
        if (r[i].m_pSelectors) memcpy(r[i].m_pSelectors, &sel[ofs[i]], p[i].m_num_pixels);
    }
    return true;
}

dxt5_endpoint_optimizer::dxt5_endpoint_optimizer() : m_reserved(nullptr) {}

bool dxt5_endpoint_optimizer::compute(const params& p, results& r) { return compute_batch(&p, &r, 1); }

bool dxt5_endpoint_optimizer::compute_batch(const params* p, results* r, uint n)
{
    if (!p || !r) return false;
    if (!n) return true;
    for (uint i = 0; i < n; i++)
        if (!p[i].m_pPixels || p[i].m_num_pixels != 16 || p[i].m_comp_index > 3 || p[i].m_quality > cCRNDXTQualityUber ||
            p[i].m_quality != p[0].m_quality || p[i].m_use_both_block_types != p[0].m_use_both_block_types)
            return false;
    std::vector<uint8_t> v((size_t)n * 16);
    for (uint i = 0; i < n; i++) {
        const uint8_t* px = reinterpret_cast<const uint8_t*>(p[i].m_pPixels);
        for (int t = 0; t < 16; t++) v[(size_t)i * 16 + t] = px[4 * t + p[i].m_comp_index];
    }
    std::vector<uint64_t> sel(n), err(n);
    std::vector<uint32_t> meta(n);
    if (crn2_dxt5a_blocks(v.data(), n, (int)p[0].m_quality, p[0].m_use_both_block_types ? 1 : 0, sel.data(),
                          err.data(), meta.data()) != CRN2_OK)
        return false;
    for (uint i = 0; i < n; i++) {
        r[i].m_error = err[i];
        r[i].m_first_endpoint = (uint8)(meta[i] & 0xFF);
        r[i].m_second_endpoint = (uint8)((meta[i] >> 8) & 0xFF);
        r[i].m_block_type = (uint8)((meta[i] >> 16) & 0xFF);
        r[i].m_reordered = ((meta[i] >> 24) & 1) != 0;
        if (r[i].m_pSelectors)
            for (int t = 0; t < 16; t++) r[i].m_pSelectors[t] = (uint8)((sel[i] >> (3 * t)) & 7);
    }
    return true;
}

}  // namespace crnlib
