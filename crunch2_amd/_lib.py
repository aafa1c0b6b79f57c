"""ctypes binding of libcrn2.so (include/crn2.h).

The library is built in-tree by crunch2_amd.build; loading fails loudly when it
is missing -- there is no fallback path.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcrn2.so")

CRN2_OK = 0
CRN2_EINVAL = -1
CRN2_ENOMEM = -2
CRN2_EUNSUPPORTED = -3
CRN2_ECORRUPT = -4
CRN2_EHIP = -5
CRN2_ENODEV = -6

FMT_DXT1, FMT_DXT3, FMT_DXT5 = 0, 1, 2
FMT_DXT5_CCXY, FMT_DXT5_XGXR, FMT_DXT5_XGBR, FMT_DXT5_AGBR = 3, 4, 5, 6
FMT_DXN_XY, FMT_DXN_YX, FMT_DXT5A = 7, 8, 9


class TextureInfo(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in
                ("width", "height", "levels", "faces", "bytes_per_block", "userdata0", "userdata1", "format")]


class LevelInfo(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in
                ("width", "height", "faces", "blocks_x", "blocks_y", "bytes_per_block", "format")]


class PackInput(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in
                ("width", "height", "levels", "faces", "format", "userdata0", "userdata1")] + [
        ("color_endpoints", C.c_void_p), ("n_color_endpoints", C.c_uint32),
        ("color_selectors_linear", C.c_void_p), ("n_color_selectors", C.c_uint32),
        ("alpha_endpoints", C.c_void_p), ("n_alpha_endpoints", C.c_uint32),
        ("alpha_selectors_linear", C.c_void_p), ("n_alpha_selectors", C.c_uint32),
        ("color_endpoint_idx", C.c_void_p), ("color_selector_idx", C.c_void_p),
        ("alpha0_endpoint_idx", C.c_void_p), ("alpha0_selector_idx", C.c_void_p),
        ("alpha1_endpoint_idx", C.c_void_p), ("alpha1_selector_idx", C.c_void_p), ("order_palettes", C.c_uint32)]


class EncodeItem(C.Structure):
    _fields_ = [("rgba", C.c_void_p), ("width", C.c_uint32), ("height", C.c_uint32), ("format", C.c_uint32),
                ("blocks", C.c_void_p)]


class MipParams(C.Structure):
    _fields_ = [("filter", C.c_uint32), ("flags", C.c_uint32), ("filter_scale", C.c_float),
                ("source_gamma", C.c_float), ("max_levels", C.c_uint32), ("min_mip_size", C.c_uint32)]


class Qdxt1Params(C.Structure):
    _fields_ = [("quality_level", C.c_uint32), ("dxt_quality", C.c_int), ("flags", C.c_int),
                ("dxt1a_alpha_threshold", C.c_uint32)]


class Qdxt5Params(C.Structure):
    _fields_ = [("quality_level", C.c_uint32), ("dxt_quality", C.c_int), ("flags", C.c_int)]


# name -> (restype, argtypes); every function declared in include/crn2.h
SIGNATURES = {
    "crn2_validate_file": (C.c_int, [C.c_void_p, C.c_uint32]),
    "crn2_get_texture_info": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(TextureInfo)]),
    "crn2_get_level_info": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(LevelInfo)]),
    "crn2_get_level_data": (C.c_void_p, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32)]),
    "crn2_get_bytes_per_dxt_block": (C.c_uint32, [C.c_uint32]),
    "crn2_unpack_begin": (C.c_void_p, [C.c_void_p, C.c_uint32]),
    "crn2_unpack_begin_device": (C.c_void_p, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "crn2_get_data": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_uint32)]),
    "crn2_unpack_level": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_uint32, C.c_uint32, C.c_uint32]),
    "crn2_unpack_level_segmented": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p),
                                              C.c_uint32, C.c_uint32, C.c_uint32]),
    "crn2_unpack_level_device": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.c_uint32, C.c_uint32,
                                           C.c_uint32, C.c_void_p]),
    "crn2_unpack_levels_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32,
                                            C.c_void_p]),
    "crn2_levels_size": (C.c_uint64, [C.c_void_p, C.c_uint32, C.c_uint32]),
    "crn2_unpack_end": (C.c_int, [C.c_void_p]),
    "crn2_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint32, C.POINTER(C.c_void_p),
                                    C.POINTER(C.c_uint64), C.POINTER(C.c_void_p)]),
    "crn2_batch_run": (C.c_int, [C.c_void_p, C.c_void_p]),
    "crn2_batch_run_graph": (C.c_int, [C.c_void_p, C.c_void_p]),
    "crn2_batch_status": (C.c_int, [C.c_void_p, C.c_void_p]),
    "crn2_batch_destroy": (C.c_int, [C.c_void_p]),
    "crn2_profile_enable": (None, [C.c_int]),
    "crn2_profile_reset": (None, []),
    "crn2_profile_read": (C.c_int, [C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "crn_get_width": (C.c_uint, [C.c_void_p, C.c_uint]),
    "crn_get_height": (C.c_uint, [C.c_void_p, C.c_uint]),
    "crn_get_levels": (C.c_uint, [C.c_void_p, C.c_uint]),
    "crn_get_dxt_format": (C.c_uint, [C.c_void_p, C.c_uint]),
    "crn_get_bytes_per_block": (C.c_uint, [C.c_void_p, C.c_uint]),
    "crn_get_uncompressed_size": (C.c_uint, [C.c_void_p, C.c_uint, C.c_uint]),
    "crn_decompress": (None, [C.c_void_p, C.c_uint, C.c_void_p, C.c_uint, C.c_uint, C.c_uint]),
    "crn2_dxt5a_blocks_device": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                           C.c_void_p, C.c_void_p]),
    "crn2_encode_dxt5a_image_device": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_int,
                                                 C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "crn2_dxt1_blocks_device": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.c_int, C.c_void_p, C.c_void_p]),
    "crn2_encode_dxt1_image_device": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_void_p,
                                                C.c_uint32, C.c_void_p, C.c_void_p]),
    "crn2_encode_image_device": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_int,
                                           C.c_void_p, C.c_void_p]),
    "crn2_compress_blocks": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_void_p]),
    "crn2_encode_set_create": (C.c_int, [C.POINTER(EncodeItem), C.c_uint32, C.c_int, C.c_int,
                                         C.POINTER(C.c_void_p)]),
    "crn2_encode_set_run": (C.c_int, [C.c_void_p, C.c_void_p]),
    "crn2_encode_set_run_graph": (C.c_int, [C.c_void_p, C.c_void_p]),
    "crn2_encode_set_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "crn2_encode_set_destroy": (C.c_int, [C.c_void_p]),
    "crn2_decode_dxt_device": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                         C.c_uint32, C.c_void_p]),
    "crn2_dxt_error_device": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                        C.c_uint32, C.c_void_p, C.c_void_p]),
    "crn2_decompress_blocks": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]),
    "crn2_dxt1_clusters_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p,
                                            C.c_void_p]),
    "crn2_dxt1_clusters_noalloc_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_void_p,
                                                    C.c_void_p, C.c_void_p]),
    "crn2_hc_color_selectors_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_int,
                                                 C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                                 C.c_void_p, C.c_void_p]),
    "crn2_autotune_mode": (C.c_int, [C.c_int]),
    "crn2_autotune_choice": (C.c_int, [C.c_int, C.c_uint64, C.POINTER(C.c_int), C.POINTER(C.c_float)]),
    "crn2_mlp_features_device": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "crn2_gemm_f32_device": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_int64, C.c_int64,
                                       C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_int64, C.c_int, C.c_int,
                                       C.c_int, C.c_uint32, C.c_void_p]),
    "crn2_mlp_bias_act_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_int,
                                           C.c_void_p]),
    "crn2_mlp_gelu_back_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "crn2_mlp_loss_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p,
                                         C.c_void_p]),
    "crn2_adam_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_float, C.c_float,
                                   C.c_float, C.c_float, C.c_uint32, C.c_void_p]),
    "crn2_mlp_infer_device": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p]),
    "crn2_dxt1_refine_device": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                          C.c_void_p]),
    "crn2_dxt1_block_error_device": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p,
                                               C.c_void_p, C.c_void_p]),
    "crn2_hc_alpha_selectors_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                                 C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                                 C.c_void_p, C.c_void_p]),
    "crn2_hc_selector_refine_device": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p]),
    "crn2_tree_codebook_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                            C.c_void_p, C.POINTER(C.c_uint32), C.c_void_p, C.c_void_p]),
    "crn2_hc_color_selector_codebook_device": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                                         C.c_void_p, C.c_int, C.c_void_p, C.POINTER(C.c_uint32),
                                                         C.c_void_p, C.c_void_p, C.c_void_p]),
    "crn2_hc_alpha_selector_codebook_device": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p,
                                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                                         C.POINTER(C.c_uint32), C.c_void_p, C.c_void_p, C.c_void_p]),
    "crn2_hc_endpoint_clusters_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                                   C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32), C.c_void_p]),
    "crn2_hc_tiles_device": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32, C.c_int,
                                       C.c_uint32, C.c_void_p, C.c_int, C.c_float, C.c_float, C.c_float, C.c_void_p,
                                       C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "crn2_hc_compress_device": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32,
                                          C.c_int, C.c_uint32, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.c_void_p, C.c_void_p]),
    "crn2_hc_color_endpoints_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                                 C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                                 C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "crn2_hc_alpha_endpoints_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p,
                                                 C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                                 C.c_void_p, C.c_void_p, C.c_void_p]),
    "crn2_dxt1_blocks": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.c_int, C.c_void_p]),
    "crn2_dxt1_clusters": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int, C.c_void_p, C.c_void_p]),
    "crn2_dxt5a_blocks": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "crn2_encode_image": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_void_p]),
    "crn2_compress_crn": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.c_uint32, C.c_int, C.POINTER(C.c_uint32), C.c_void_p, C.c_uint32, C.c_uint32,
                                    C.POINTER(C.c_void_p), C.POINTER(C.c_uint32)]),
    "crn2_compress_dds": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                    C.c_uint32, C.c_int, C.c_int, C.POINTER(C.c_void_p), C.POINTER(C.c_uint32)]),
    "crn2_write_dds": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                 C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_uint32)]),
    "crn2_crn_to_dds": (C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_void_p), C.POINTER(C.c_uint32)]),
    "crn2_crc16": (C.c_uint16, [C.c_void_p, C.c_uint32, C.c_uint16]),
    "crn2_synth_rgba": (C.c_int, [C.c_uint32, C.c_uint32, C.c_uint64, C.c_int, C.c_void_p]),
    "crn2_downsample": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p]),
    "crn2_compress_rgba": (C.c_int, [C.POINTER(C.c_void_p), C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                     C.c_uint32, C.POINTER(C.c_uint32), C.POINTER(C.c_void_p),
                                     C.POINTER(C.c_uint32)]),
    "crn2_free": (None, [C.c_void_p]),
    "crn2_mip_level_count": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
    "crn2_mip_plan_create": (C.c_int, [C.c_uint32, C.c_uint32, C.POINTER(MipParams), C.POINTER(C.c_void_p)]),
    "crn2_mip_plan_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint64), C.c_void_p,
                                     C.c_void_p]),
    "crn2_mip_plan_run": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]),
    "crn2_mip_plan_destroy": (C.c_int, [C.c_void_p]),
    "crn2_mip_plan_profile": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.POINTER(C.c_uint64)]),
    "crn2_generate_mipmaps_device": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(MipParams), C.c_void_p,
                                               C.c_uint64, C.c_void_p]),
    "crn2_generate_mipmaps": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.POINTER(MipParams), C.c_void_p,
                                        C.c_uint64]),
    "crn2_mip_contributors": (C.c_int, [C.c_uint32, C.c_uint32, C.c_int, C.c_uint32, C.c_float, C.c_void_p,
                                        C.c_void_p, C.c_void_p, C.c_uint32]),
    "crn2_mip_gamma_tables": (C.c_int, [C.c_float, C.c_void_p, C.c_void_p]),
    "crn2_qdxt1_init": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_int, C.POINTER(C.c_void_p),
                                  C.c_void_p]),
    "crn2_qdxt1_pack": (C.c_int, [C.c_void_p, C.POINTER(Qdxt1Params), C.c_float, C.c_void_p, C.c_void_p]),
    "crn2_qdxt1_info": (C.c_int, [C.c_void_p] + [C.POINTER(C.c_uint32)] * 4),
    "crn2_qdxt1_training": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "crn2_qdxt1_destroy": (C.c_int, [C.c_void_p]),
    "crn2_qdxt5_init": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_int, C.c_uint32,
                                  C.POINTER(C.c_void_p), C.c_void_p]),
    "crn2_qdxt5_pack": (C.c_int, [C.c_void_p, C.POINTER(Qdxt5Params), C.c_void_p, C.c_void_p]),
    "crn2_qdxt5_info": (C.c_int, [C.c_void_p] + [C.POINTER(C.c_uint32)] * 4),
    "crn2_qdxt5_training": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    "crn2_qdxt5_destroy": (C.c_int, [C.c_void_p]),
    "crn2_compress_dds_clustered": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                              C.c_uint32, C.c_uint32, C.c_int, C.c_int, C.c_uint32,
                                              C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_uint32)]),
    "crn2_qdxt_profile": (C.c_int, [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "crn2_vq_clusters_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32,
                                          C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32),
                                          C.c_void_p]),
    "crn2_vq_threaded_clusters_device": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32,
                                                   C.c_void_p, C.c_void_p, C.POINTER(C.c_uint32), C.c_void_p]),
}

# internal (not in include/): the block packer behind crn2_compress_rgba
INTERNAL = {
    "crn2_pack": (C.c_int, [C.POINTER(PackInput), C.POINTER(C.c_void_p), C.POINTER(C.c_uint32)]),
    "crn2_prefix_lengths": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libcrn2.so is not built (run python -m crunch2_amd.build): " + LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in list(SIGNATURES.items()) + list(INTERNAL.items()):
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what="libcrn2"):
    if rc != CRN2_OK:
        raise RuntimeError("%s failed with status %d" % (what, rc))
    return rc
