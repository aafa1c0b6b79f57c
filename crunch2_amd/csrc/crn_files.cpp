// crn_files.cpp -- DDS files around the GPU transcoder and encoders.
//
// crn2_write_dds follows mipmapped_texture::write_dds for DXTn textures
// (crnlib/crn_mipmapped_texture.cpp:921-1140): "DDS " and a 124-byte
// DDSURFACEDESC2 (inc/dds_defs.h:81-158) with width, height, pixel format,
// caps, mip count and cube flags, a linear size of the rounded-up top level,
// then every face's levels in order, each level's blocks at minimal pitch.
// The pixel format is the FOURCC crnd_crn_format_to_fourcc gives
// (inc/crn_decomp.h:2567-2606) with write_dds's special cases: DXN_XY
// ('A2XY') is written as FOURCC 'ATI2' with 'A2XY' in dwRGBBitCount, the
// swizzled DXT5 formats as 'DXT5' with their own FOURCC there.
// crn2_crn_to_dds is crn_decompress_crn_to_dds (crnlib/crnlib.cpp:269-291):
// read_crn_from_memory's per-level transcode (crn_mipmapped_texture.cpp:
// 2687-2813, row pitch = blocks_x * bytes per block), then the writer.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "crn2_internal.h"

namespace {

uint32_t fourcc(char a, char b, char c, char d)
{
    return (uint32_t)(uint8_t)a | ((uint32_t)(uint8_t)b << 8) | ((uint32_t)(uint8_t)c << 16) |
           ((uint32_t)(uint8_t)d << 24);
}

// FOURCC of crnd_crn_format_to_fourcc for the DXTn formats (0 otherwise)
uint32_t format_fourcc(uint32_t f)
{
    switch (f) {
        case CRN2_FMT_DXT1: return fourcc('D', 'X', 'T', '1');
        case CRN2_FMT_DXT3: return fourcc('D', 'X', 'T', '3');
        case CRN2_FMT_DXT5: return fourcc('D', 'X', 'T', '5');
        case CRN2_FMT_DXN_XY: return fourcc('A', '2', 'X', 'Y');
        case CRN2_FMT_DXN_YX: return fourcc('A', 'T', 'I', '2');
        case CRN2_FMT_DXT5A: return fourcc('A', 'T', 'I', '1');
        case CRN2_FMT_DXT5_CCXY: return fourcc('C', 'C', 'x', 'Y');
        case CRN2_FMT_DXT5_XGXR: return fourcc('x', 'G', 'x', 'R');
        case CRN2_FMT_DXT5_XGBR: return fourcc('x', 'G', 'B', 'R');
        case CRN2_FMT_DXT5_AGBR: return fourcc('A', 'G', 'B', 'R');
        default: return 0;
    }
}

uint32_t bytes_per_block(uint32_t f) { return (f == CRN2_FMT_DXT1 || f == CRN2_FMT_DXT5A) ? 8 : 16; }

void put32(uint8_t* p, uint32_t v)
{
    p[0] = (uint8_t)v;
    p[1] = (uint8_t)(v >> 8);
    p[2] = (uint8_t)(v >> 16);
    p[3] = (uint8_t)(v >> 24);
}

size_t level_bytes(uint32_t width, uint32_t height, uint32_t l, uint32_t format)
{
    const uint32_t w = std::max<uint32_t>(1, width >> l), h = std::max<uint32_t>(1, height >> l);
    return (size_t)((w + 3) >> 2) * ((h + 3) >> 2) * bytes_per_block(format);
}

}  // namespace

extern "C" {

int crn2_write_dds(uint32_t width, uint32_t height, uint32_t levels, uint32_t faces, uint32_t format,
                   const void* const* level_data, uint8_t** out, uint32_t* out_size)
{
    if (!out || !out_size) return CRN2_EINVAL;
    *out = nullptr;
    *out_size = 0;
    if (!level_data || !width || !height || !levels || levels > 16 || (faces != 1 && faces != 6)) return CRN2_EINVAL;
    if (format > CRN2_FMT_DXT5A) return CRN2_EUNSUPPORTED;
    size_t total = 128;
    for (uint32_t f = 0; f < faces; f++)
        for (uint32_t l = 0; l < levels; l++) {
            if (!level_data[f * levels + l]) return CRN2_EINVAL;
            total += level_bytes(width, height, l, format);
        }
    if (total > 0xFFFFFFFFull) return CRN2_EINVAL;
    uint8_t* p = static_cast<uint8_t*>(malloc(total));
    if (!p) return CRN2_ENOMEM;
    memset(p, 0, 128);
    memcpy(p, "DDS ", 4);
    uint8_t* d = p + 4;   // DDSURFACEDESC2
    uint32_t flags = 0x1 | 0x2 | 0x4 | 0x1000;          // caps, height, width, pixel format
    uint32_t caps = 0x1000, caps2 = 0;                   // texture
    if (levels > 1) {
        flags |= 0x20000;                                // mipmap count
        caps |= 0x400000 | 0x8;                          // mipmap, complex
    }
    if (faces > 1) {
        caps |= 0x8;
        caps2 |= 0x200 | 0x400 | 0x800 | 0x1000 | 0x2000 | 0x4000 | 0x8000;   // cubemap, all six faces
    }
    uint32_t cc = format_fourcc(format), bitcount = 0;
    if (format == CRN2_FMT_DXN_XY) {
        cc = fourcc('A', 'T', 'I', '2');
        bitcount = fourcc('A', '2', 'X', 'Y');
    } else if (format >= CRN2_FMT_DXT5_CCXY && format <= CRN2_FMT_DXT5_AGBR) {
        bitcount = cc;
        cc = fourcc('D', 'X', 'T', '5');
    }
    const uint32_t bpp = bytes_per_block(format) == 8 ? 4 : 8;
    flags |= 0x80000;                                    // linear size
    put32(d + 0, 124);
    put32(d + 4, flags);
    put32(d + 8, height);
    put32(d + 12, width);
    put32(d + 16, (uint32_t)((((uint64_t)(width + 3) & ~3ull) * ((height + 3) & ~3u) * bpp) >> 3));
    if (levels > 1) put32(d + 24, levels);
    put32(d + 72, 32);                                   // pixel format: size, flags, FOURCC, bit count
    put32(d + 76, 0x4);
    put32(d + 80, cc);
    put32(d + 84, bitcount);
    put32(d + 104, caps);
    put32(d + 108, caps2);
    size_t o = 128;
    for (uint32_t f = 0; f < faces; f++)
        for (uint32_t l = 0; l < levels; l++) {
            const size_t n = level_bytes(width, height, l, format);
            memcpy(p + o, level_data[f * levels + l], n);
            o += n;
        }
    *out = p;
    *out_size = (uint32_t)total;
    return CRN2_OK;
}

int crn2_crn_to_dds(const void* crn, uint32_t size, uint8_t** out, uint32_t* out_size)
{
    if (!out || !out_size) return CRN2_EINVAL;
    *out = nullptr;
    *out_size = 0;
    crn2_texture_info t;
    if (!crn || crn2_get_texture_info(crn, size, &t) != CRN2_OK) return CRN2_EINVAL;
    if (!format_fourcc(t.format) || t.format > CRN2_FMT_DXT5A) return CRN2_EUNSUPPORTED;
    if (t.faces != 1 && t.faces != 6) return CRN2_EINVAL;
    crn2_unpack_ctx* ctx = crn2_unpack_begin(crn, size);
    if (!ctx) return CRN2_ECORRUPT;
    std::vector<std::vector<uint8_t>> data((size_t)t.faces * t.levels);
    std::vector<const void*> ptrs(data.size());
    int rc = CRN2_OK;
    for (uint32_t l = 0; l < t.levels && rc == CRN2_OK; l++) {
        const uint32_t w = std::max<uint32_t>(1, t.width >> l), h = std::max<uint32_t>(1, t.height >> l);
        const uint32_t pitch = ((w + 3) >> 2) * bytes_per_block(t.format), bytes = pitch * ((h + 3) >> 2);
        void* dst[6];
        for (uint32_t f = 0; f < t.faces; f++) {
            data[f * t.levels + l].resize(bytes);
            dst[f] = data[f * t.levels + l].data();
            ptrs[f * t.levels + l] = dst[f];
        }
        rc = crn2_unpack_level(ctx, dst, bytes, pitch, l);
    }
    crn2_unpack_end(ctx);
    if (rc != CRN2_OK) return rc;
    return crn2_write_dds(t.width, t.height, t.levels, t.faces, t.format, ptrs.data(), out, out_size);
}

int crn2_compress_dds(const uint8_t* const* levels, uint32_t width, uint32_t height, uint32_t nlevels, uint32_t faces,
                      uint32_t format, int quality, int flags, uint8_t** out, uint32_t* out_size)
{
    if (!out || !out_size) return CRN2_EINVAL;
    *out = nullptr;
    *out_size = 0;
    if (!levels || !width || !height || !nlevels || nlevels > 16 || (faces != 1 && faces != 6)) return CRN2_EINVAL;
    if (format > CRN2_FMT_DXT5A) return CRN2_EUNSUPPORTED;
    std::vector<std::vector<uint8_t>> data((size_t)faces * nlevels);
    std::vector<const void*> ptrs(data.size());
    for (uint32_t f = 0; f < faces; f++)
        for (uint32_t l = 0; l < nlevels; l++) {
            const uint32_t w = std::max<uint32_t>(1, width >> l), h = std::max<uint32_t>(1, height >> l);
            std::vector<uint8_t>& v = data[f * nlevels + l];
            v.resize(level_bytes(width, height, l, format));
            if (!levels[f * nlevels + l]) return CRN2_EINVAL;
            const int rc = crn2_encode_image(levels[f * nlevels + l], w, h, format, quality, flags, v.data());
            if (rc != CRN2_OK) return rc;
            ptrs[f * nlevels + l] = v.data();
        }
    return crn2_write_dds(width, height, nlevels, faces, format, ptrs.data(), out, out_size);
}

}  // extern "C"
