// crn_api.cpp -- header queries and the emscripten-compatible C exports.
//
// crn2_validate_file      <- crnd_validate_file    (inc/crn_decomp.h:2672-2735)
// crn2_get_texture_info   <- crnd_get_texture_info (inc/crn_decomp.h:2737-2759)
// crn2_get_level_info     <- crnd_get_level_info   (inc/crn_decomp.h:2761-2788)
// crn2_get_level_data     <- crnd_get_level_data   (inc/crn_decomp.h:2790-2817)
// crn_get_* / crn_decompress <- emscripten/crunch_lib.cpp:44-120
#include <algorithm>
#include <cstring>

#include "crn2_internal.h"
#include "crnd_host.h"

using namespace crn2;

static uint32_t max_mips(uint32_t w, uint32_t h)
{
    if ((w | h) == 0) return 0;
    uint32_t n = 1;
    while (w > 1 || h > 1) {
        w >>= 1;
        h >>= 1;
        n++;
    }
    return n;
}

extern "C" {

uint32_t crn2_get_bytes_per_dxt_block(uint32_t format)
{
    switch (format) {
    case CRN2_FMT_DXT1: case CRN2_FMT_DXT5A: case 10: case 11: case 13: return 8;
    case CRN2_FMT_DXT3: case CRN2_FMT_DXT5: case CRN2_FMT_DXN_XY: case CRN2_FMT_DXN_YX: case CRN2_FMT_DXT5_CCXY:
    case CRN2_FMT_DXT5_XGXR: case CRN2_FMT_DXT5_XGBR: case CRN2_FMT_DXT5_AGBR: case 12: case 14: return 16;
    }
    return 0;
}

int crn2_validate_file(const void* data, uint32_t size)
{
    Header h;
    const uint8_t* d = static_cast<const uint8_t*>(data);
    if (!d || size < 62 || !parse_header(d, size, h)) return 0;
    if (crn2_crc16(d + 6, h.header_size - 6, 0) != h.header_crc) return 0;
    if (crn2_crc16(d + h.header_size, h.data_size - h.header_size, 0) != h.data_crc) return 0;
    if (h.faces != 1 && h.faces != 6) return 0;
    if (h.width < 1 || h.width > 4096 || h.height < 1 || h.height > 4096) return 0;
    if (h.levels < 1 || h.levels > max_mips(h.width, h.height)) return 0;
    if (h.format > 14) return 0;
    return 1;
}

int crn2_get_texture_info(const void* data, uint32_t size, crn2_texture_info* info)
{
    Header h;
    if (!info || !parse_header(static_cast<const uint8_t*>(data), size, h)) return CRN2_EINVAL;
    info->width = h.width;
    info->height = h.height;
    info->levels = h.levels;
    info->faces = h.faces;
    info->format = h.format;
    // crn_decomp.h:2754: 8 bytes for DXT1, DXT5A and the ETC1-class formats
    info->bytes_per_block = (h.format == 0 || h.format == 9 || h.format == 10 || h.format == 11 || h.format == 13) ? 8 : 16;
    info->userdata0 = h.userdata0;
    info->userdata1 = h.userdata1;
    return CRN2_OK;
}

int crn2_get_level_info(const void* data, uint32_t size, uint32_t level, crn2_level_info* info)
{
    Header h;
    if (!info || size < 62 || !parse_header(static_cast<const uint8_t*>(data), size, h) || level >= h.levels)
        return CRN2_EINVAL;
    info->width = std::max<uint32_t>(1, h.width >> level);
    info->height = std::max<uint32_t>(1, h.height >> level);
    info->faces = h.faces;
    info->blocks_x = (info->width + 3) >> 2;
    info->blocks_y = (info->height + 3) >> 2;
    // crn_decomp.h:2784 reports 8 bytes only for DXT1 and DXT5A
    info->bytes_per_block = (h.format == CRN2_FMT_DXT1 || h.format == CRN2_FMT_DXT5A) ? 8 : 16;
    info->format = h.format;
    return CRN2_OK;
}

const void* crn2_get_level_data(const void* data, uint32_t size, uint32_t level, uint32_t* level_size)
{
    if (level_size) *level_size = 0;
    Header h;
    if (size < 62 || !parse_header(static_cast<const uint8_t*>(data), size, h) || level >= h.levels) return nullptr;
    if (level_size) {
        const uint32_t next = (level + 1 < h.levels) ? h.level_ofs[level + 1] : size;
        *level_size = next - h.level_ofs[level];
    }
    return static_cast<const uint8_t*>(data) + h.level_ofs[level];
}

static crn2_texture_info tex_info(void* src, unsigned int src_size)
{
    crn2_texture_info t;
    memset(&t, 0, sizeof(t));
    crn2_get_texture_info(src, src_size, &t);
    return t;
}

unsigned int crn_get_width(void* src, unsigned int src_size) { return tex_info(src, src_size).width; }
unsigned int crn_get_height(void* src, unsigned int src_size) { return tex_info(src, src_size).height; }
unsigned int crn_get_levels(void* src, unsigned int src_size) { return tex_info(src, src_size).levels; }
unsigned int crn_get_dxt_format(void* src, unsigned int src_size) { return tex_info(src, src_size).format; }

unsigned int crn_get_bytes_per_block(void* src, unsigned int src_size)
{
    return crn2_get_bytes_per_dxt_block(tex_info(src, src_size).format);
}

unsigned int crn_get_uncompressed_size(void* src, unsigned int src_size, unsigned int level)
{
    const crn2_texture_info t = tex_info(src, src_size);
    const uint32_t w = t.width >> level, h = t.height >> level;
    return ((w + 3) >> 2) * crn2_get_bytes_per_dxt_block(t.format) * ((h + 3) >> 2);
}

void crn_decompress(void* src, unsigned int src_size, void* dst, unsigned int dst_size, unsigned int firstLevel,
                    unsigned int levelCount)
{
    (void)dst_size;
    const crn2_texture_info t = tex_info(src, src_size);
    uint32_t w = t.width >> firstLevel, h = t.height >> firstLevel;
    const uint32_t bpb = crn2_get_bytes_per_dxt_block(t.format);
    crn2_unpack_ctx* c = crn2_unpack_begin(src, src_size);
    uint8_t* out = static_cast<uint8_t*>(dst);
    for (unsigned int l = firstLevel; l < firstLevel + levelCount; l++) {
        const uint32_t pitch = ((w + 3) >> 2) * bpb, total = pitch * ((h + 3) >> 2);
        void* faces[1] = {out};
        if (c) crn2_unpack_level(c, faces, total, pitch, l);
        out += total;
        w >>= 1;
        h >>= 1;
    }
    if (c) crn2_unpack_end(c);
}

}  // extern "C"
