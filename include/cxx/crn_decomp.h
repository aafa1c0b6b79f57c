/*
 * crn_decomp.h (libcrn2) -- the crnd transcoder API with the reference's C++
 * signatures, for callers that compiled against crnlib's inc/crn_decomp.h.
 *
 * Put include/cxx (this directory) first on the include path in place of the
 * reference's inc/, link libcrn2.so, and do not compile the reference's
 * header-only transcoder: every function below is defined in libcrn2 with
 * C++ linkage, so the mangled names are those the reference's declarations
 * produce (inc/crn_defs.h:81-221), and the work runs on the GPU through the C
 * ABI of crn2.h.  Only declarations are given here; the file-format structs
 * of the reference header (crn_header, crn_packed_uint) are not part of the
 * call interface and are not restated.
 *
 * Interface parity (types, struct layouts, argument meaning, bool/NULL error
 * convention) follows inc/crn_defs.h:41-221.  Behaviour differences:
 *   - crnd_unpack_level rejects level_index >= levels (the reference reads
 *     past m_level_ofs);
 *   - a stream that decodes to an out-of-range palette index fails the call
 *     instead of reading outside the palette;
 *   - crnd_set_memory_callbacks is accepted and recorded, but the context's
 *     host memory comes from the C++ allocator (device memory from HIP).
 */
#ifndef CRN2_CXX_CRN_DECOMP_H
#define CRN2_CXX_CRN_DECOMP_H

#include <stddef.h>

#include "crnlib.h"

namespace crnd {
typedef unsigned char uint8;
typedef signed char int8;
typedef unsigned short uint16;
typedef signed short int16;
typedef unsigned int uint32;
typedef unsigned int uint;
typedef signed int int32;
typedef unsigned long long uint64;
typedef long long int64;

typedef void* (*crnd_realloc_func)(void* p, size_t size, size_t* pActual_size, bool movable, void* pUser_data);
typedef size_t (*crnd_msize_func)(void* p, void* pUser_data);
void crnd_set_memory_callbacks(crnd_realloc_func pRealloc, crnd_msize_func pMSize, void* pUser_data);

struct crn_file_info {
    crn_file_info() : m_struct_size(sizeof(crn_file_info)) {}
    uint32 m_struct_size;
    uint32 m_actual_data_size;
    uint32 m_header_size;
    uint32 m_total_palette_size;
    uint32 m_tables_size;
    uint32 m_levels;
    uint32 m_level_compressed_size[cCRNMaxLevels];
    uint32 m_color_endpoint_palette_entries;
    uint32 m_color_selector_palette_entries;
    uint32 m_alpha_endpoint_palette_entries;
    uint32 m_alpha_selector_palette_entries;
};

struct crn_texture_info {
    crn_texture_info() : m_struct_size(sizeof(crn_texture_info)) {}
    uint32 m_struct_size;
    uint32 m_width;
    uint32 m_height;
    uint32 m_levels;
    uint32 m_faces;
    uint32 m_bytes_per_block;
    uint32 m_userdata0;
    uint32 m_userdata1;
    crn_format m_format;
};

struct crn_level_info {
    crn_level_info() : m_struct_size(sizeof(crn_level_info)) {}
    uint32 m_struct_size;
    uint32 m_width;
    uint32 m_height;
    uint32 m_faces;
    uint32 m_blocks_x;
    uint32 m_blocks_y;
    uint32 m_bytes_per_block;
    crn_format m_format;
};

uint32 crnd_crn_format_to_fourcc(crn_format fmt);
crn_format crnd_get_fundamental_dxt_format(crn_format fmt);
uint32 crnd_get_crn_format_bits_per_texel(crn_format fmt);
uint32 crnd_get_bytes_per_dxt_block(crn_format fmt);

bool crnd_validate_file(const void* pData, uint32 data_size, crn_file_info* pFile_info);
bool crnd_get_texture_info(const void* pData, uint32 data_size, crn_texture_info* pTexture_info);
bool crnd_get_level_info(const void* pData, uint32 data_size, uint32 level_index, crn_level_info* pLevel_info);

typedef void* crnd_unpack_context;

/* pData must stay valid until crnd_unpack_end. */
crnd_unpack_context crnd_unpack_begin(const void* pData, uint32 data_size);
bool crnd_get_data(crnd_unpack_context pContext, const void** ppData, uint32* pData_size);
/* ppDst: one host buffer per face; row pitch a multiple of 4 (0: minimal). */
bool crnd_unpack_level(crnd_unpack_context pContext, void** ppDst, uint32 dst_size_in_bytes,
                       uint32 row_pitch_in_bytes, uint32 level_index);
bool crnd_unpack_level_segmented(crnd_unpack_context pContext, const void* pSrc, uint32 src_size_in_bytes,
                                 void** ppDst, uint32 dst_size_in_bytes, uint32 row_pitch_in_bytes,
                                 uint32 level_index);
bool crnd_unpack_end(crnd_unpack_context pContext);

const void* crnd_get_level_data(const void* pData, uint32 data_size, uint32 level_index, uint32* pSize);
uint32 crnd_get_segmented_file_size(const void* pData, uint32 data_size);
bool crnd_create_segmented_file(const void* pData, uint32 data_size, void* pBase_data, uint base_data_size);
}  // namespace crnd

#endif  // CRN2_CXX_CRN_DECOMP_H
