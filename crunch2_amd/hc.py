"""dxt_hc on the GPU (crnlib/crn_dxt_hc.cpp): selector assignment (colour
:1306-1360, alpha :1516-1586), selector codebook refinement, colour and alpha
endpoint codebooks -- Python face of the crn2_hc_* entry points."""
import ctypes as C

import torch

from ._lib import lib, check


def _ptr(t):
    return C.c_void_p(t.data_ptr() if t is not None else 0)


def color_selectors(blocks, endpoint_index, cluster_colors, selectors, subblocks=False, perceptual=True,
                    details=None, used=None):
    """blocks: uint8 CUDA tensor [n, 16, 4]; endpoint_index: int32/uint32
    [n] or [n, 2]; cluster_colors: uint8 [k, 4, 4]; selectors: int32 [m]
    (bit patterns).  Returns (best_index int32 [n], best_error int32 [n]);
    details (int32 [m, 64]) and used (uint8 [m]) are accumulated in place
    when given."""
    n = blocks.shape[0]
    m = selectors.shape[0]
    dev = blocks.device
    best = torch.empty(n, dtype=torch.int32, device=dev)
    err = torch.empty(n, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    check(lib().crn2_hc_color_selectors_device(_ptr(blocks.contiguous()), _ptr(endpoint_index.contiguous()),
                                               _ptr(cluster_colors.contiguous()), n, int(bool(subblocks)),
                                               int(bool(perceptual)), _ptr(selectors.contiguous()), m, _ptr(best),
                                               _ptr(err), _ptr(details), _ptr(used), C.c_void_p(stream)),
          "crn2_hc_color_selectors_device")
    return best, err


def color_endpoints(cluster_px, cluster_ofs, block_cluster, blocks, block_weights, block_encodings, perceptual=True):
    """dxt_hc colour endpoint codebook (crn_dxt_hc.cpp:663-754).  cluster_px:
    uint8 CUDA tensor [t, 4]; cluster_ofs: int32 [k + 1]; block_cluster: int32
    [n] (-1: no cluster); blocks: uint8 [n, 16, 4]; block_weights: float32 [n];
    block_encodings: uint8 [n].  Returns (endpoints int32 [k], color_values
    uint8 [k, 4, 4], block_selectors int64 [n] (zero for blocks of no
    cluster), clusters not encoded)."""
    k = cluster_ofs.shape[0] - 1
    n = blocks.shape[0]
    dev = blocks.device
    ends = torch.empty(k, dtype=torch.int32, device=dev)
    cv = torch.empty((k, 4, 4), dtype=torch.uint8, device=dev)
    bsel = torch.zeros(n, dtype=torch.int64, device=dev)
    ne = torch.zeros(1, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    check(lib().crn2_hc_color_endpoints_device(_ptr(cluster_px.contiguous()), _ptr(cluster_ofs.contiguous()), k,
                                               int(cluster_px.shape[0]), _ptr(block_cluster.contiguous()), n,
                                               _ptr(blocks.contiguous()), _ptr(block_weights.contiguous()),
                                               _ptr(block_encodings.contiguous()), int(bool(perceptual)), _ptr(ends),
                                               _ptr(cv), _ptr(bsel), _ptr(ne), C.c_void_p(stream)),
          "crn2_hc_color_endpoints_device")
    return ends, cv, bsel, int(ne.item())


def alpha_endpoints(cluster_vals, cluster_ofs, block_cluster, block_vals, block_encodings):
    """dxt_hc alpha endpoint codebook for one alpha channel
    (crn_dxt_hc.cpp:1014-1130).  cluster_vals: uint8 CUDA tensor [t];
    cluster_ofs: int32 [k + 1]; block_cluster: int32 [n] (-1: none);
    block_vals: uint8 [n, 16]; block_encodings: uint8 [n].  Returns
    (endpoints int32 [k] = first | second << 8 | refined << 16, alpha_values
    uint8 [k, 8], refined_values uint8 [k, 8], block_selectors int64 [n])."""
    k = cluster_ofs.shape[0] - 1
    n = block_vals.shape[0]
    dev = block_vals.device
    ends = torch.empty(k, dtype=torch.int32, device=dev)
    av = torch.empty((k, 8), dtype=torch.uint8, device=dev)
    rv = torch.empty((k, 8), dtype=torch.uint8, device=dev)
    bsel = torch.zeros(n, dtype=torch.int64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    check(lib().crn2_hc_alpha_endpoints_device(_ptr(cluster_vals.contiguous()), _ptr(cluster_ofs.contiguous()), k,
                                               int(cluster_vals.shape[0]), _ptr(block_cluster.contiguous()), n,
                                               _ptr(block_vals.contiguous()), _ptr(block_encodings.contiguous()),
                                               _ptr(ends), _ptr(av), _ptr(rv), _ptr(bsel), C.c_void_p(stream)),
          "crn2_hc_alpha_endpoints_device")
    return ends, av, rv, bsel


def alpha_selectors(block_vals, endpoint_index, values, selectors, scatter_values=None, details=None, used=None):
    """block_vals: uint8 CUDA tensor [n, 16]; endpoint_index: int32 [n];
    values / scatter_values: uint8 [k, 8]; selectors: int64 [m] (3 bits per
    texel).  Returns (best_index int32 [n], best_error int32 [n]); details
    (int32 [m, 128]) and used (uint8 [m]) are accumulated in place when given."""
    n = block_vals.shape[0]
    m = selectors.shape[0]
    dev = block_vals.device
    best = torch.empty(n, dtype=torch.int32, device=dev)
    err = torch.empty(n, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    sv = scatter_values.contiguous() if scatter_values is not None else None
    check(lib().crn2_hc_alpha_selectors_device(_ptr(block_vals.contiguous()), _ptr(endpoint_index.contiguous()),
                                               _ptr(values.contiguous()), _ptr(sv), n,
                                               _ptr(selectors.contiguous()), m, _ptr(best), _ptr(err),
                                               _ptr(details), _ptr(used), C.c_void_p(stream)),
          "crn2_hc_alpha_selectors_device")
    return best, err


def selector_refine(details, bits):
    """details: int32 CUDA tensor of accumulated totals, [m, 64] for bits 2 or
    [m, 128] for bits 3.  Returns the refined codebook:
    int32 [m] (bits 2) or int64 [m] (bits 3)."""
    m = details.shape[0]
    out = torch.empty(m, dtype=torch.int32 if bits == 2 else torch.int64, device=details.device)
    stream = torch.cuda.current_stream().cuda_stream
    check(lib().crn2_hc_selector_refine_device(_ptr(details.contiguous()), m, int(bits), _ptr(out),
                                               C.c_void_p(stream)), "crn2_hc_selector_refine_device")
    return out


def tree_codebook(vectors, weights, max_splits, leaves=True):
    """Top-down VQ codebook (crn_tree_clusterizer.h:89-171).  vectors: float32
    CUDA tensor [n, 6] or [n, 16]; weights: int32 [n] (non-negative).  Returns
    (codebook float32 [k, dim], vector_leaf int32 [n] or None)."""
    n, dim = vectors.shape
    dev = vectors.device
    cb = torch.empty((max_splits, dim), dtype=torch.float32, device=dev)
    leaf = torch.empty(n, dtype=torch.int32, device=dev) if leaves else None
    k = C.c_uint32(0)
    stream = torch.cuda.current_stream().cuda_stream
    check(lib().crn2_tree_codebook_device(_ptr(vectors.contiguous()), _ptr(weights.contiguous()), n, dim,
                                          max_splits, _ptr(cb), C.byref(k), _ptr(leaf), C.c_void_p(stream)),
          "crn2_tree_codebook_device")
    return cb[:k.value], leaf


def color_selector_codebook(block_selectors, codebook_size, blocks, endpoint_index, cluster_colors,
                            perceptual=True):
    """dxt_hc colour selector codebook (crn_dxt_hc.cpp:1379-1504).
    block_selectors: int64 CUDA tensor [n] (selector << 32 | weight); blocks
    uint8 [n, 16, 4]; endpoint_index int32 [n]; cluster_colors uint8 [k, 4, 4].
    Returns (selectors int32 [m], selector_index int32 [n], used uint8 [m])."""
    n = blocks.shape[0]
    dev = blocks.device
    sel = torch.empty(codebook_size, dtype=torch.int32, device=dev)
    idx = torch.empty(n, dtype=torch.int32, device=dev)
    used = torch.empty(codebook_size, dtype=torch.uint8, device=dev)
    m = C.c_uint32(0)
    stream = torch.cuda.current_stream().cuda_stream
    check(lib().crn2_hc_color_selector_codebook_device(_ptr(block_selectors.contiguous()), n, codebook_size,
                                                       _ptr(blocks.contiguous()), _ptr(endpoint_index.contiguous()),
                                                       _ptr(cluster_colors.contiguous()), int(bool(perceptual)),
                                                       _ptr(sel), C.byref(m), _ptr(idx), _ptr(used),
                                                       C.c_void_p(stream)),
          "crn2_hc_color_selector_codebook_device")
    return sel[:m.value], idx, used[:m.value]


def alpha_selector_codebook(block_selectors, codebook_size, block_vals, endpoint_index, alpha_values,
                            refined_values=None):
    """dxt_hc alpha selector codebook (crn_dxt_hc.cpp:1588-1720) for one or
    two channels.  block_selectors int64 [c, n]; block_vals uint8 [c, n, 16];
    endpoint_index int32 [c, n]; alpha_values / refined_values uint8 [k, 8].
    Returns (selectors int64 [m], selector_index int32 [c, n], used uint8 [m])."""
    nc, n = block_selectors.shape
    dev = block_vals.device
    sel = torch.empty(codebook_size, dtype=torch.int64, device=dev)
    idx = torch.empty((nc, n), dtype=torch.int32, device=dev)
    used = torch.empty(codebook_size, dtype=torch.uint8, device=dev)
    m = C.c_uint32(0)
    stream = torch.cuda.current_stream().cuda_stream
    rv = refined_values.contiguous() if refined_values is not None else None
    check(lib().crn2_hc_alpha_selector_codebook_device(_ptr(block_selectors.contiguous()), n, nc, codebook_size,
                                                       _ptr(block_vals.contiguous()), _ptr(endpoint_index.contiguous()),
                                                       _ptr(alpha_values.contiguous()), _ptr(rv), _ptr(sel),
                                                       C.byref(m), _ptr(idx), _ptr(used), C.c_void_p(stream)),
          "crn2_hc_alpha_selector_codebook_device")
    return sel[:m.value], idx, used[:m.value]


def endpoint_clusters(vectors, weights, codebook_size):
    """dxt_hc endpoint clustering (crn_dxt_hc.cpp:888-972 colour, dim 6;
    1160-1239 alpha, dim 2).  vectors: float32 CUDA tensor [n, dim]; weights
    int32 [n].  Returns (cluster int32 [n], codebook float32 [k, dim])."""
    n, dim = vectors.shape
    dev = vectors.device
    cl = torch.empty(n, dtype=torch.int32, device=dev)
    cb = torch.empty((codebook_size, dim), dtype=torch.float32, device=dev)
    k = C.c_uint32(0)
    stream = torch.cuda.current_stream().cuda_stream
    check(lib().crn2_hc_endpoint_clusters_device(_ptr(vectors.contiguous()), _ptr(weights.contiguous()), n, dim,
                                                 codebook_size, _ptr(cl), _ptr(cb), C.byref(k), C.c_void_p(stream)),
          "crn2_hc_endpoint_clusters_device")
    return cl, cb[:k.value]


def tiles(blocks, levels, num_faces=1, has_color=True, alpha_comp=(), perceptual=True, color_derating=2.0,
          alpha_derating=2.0, ratio=3.0):
    """dxt_hc adaptive tile choice (crn_dxt_hc.cpp:386-561).  blocks: uint8
    CUDA tensor [n, 16, 4]; levels: list of (first_block, num_blocks,
    block_width[, weight]).  Returns (block_encodings uint8 [n], tile_indices
    int32 [n], tile_src int32 [n], tile_vec float32 [n, 10])."""
    n = blocks.shape[0]
    dev = blocks.device
    desc = (C.c_uint32 * (3 * len(levels)))(*[int(x) for lv in levels for x in lv[:3]])
    ac = (C.c_uint32 * 2)(*(list(alpha_comp) + [0, 0])[:2])
    enc = torch.empty(n, dtype=torch.uint8, device=dev)
    ti = torch.empty(n, dtype=torch.int32, device=dev)
    src = torch.empty(n, dtype=torch.int32, device=dev)
    vec = torch.empty((n, 10), dtype=torch.float32, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    check(lib().crn2_hc_tiles_device(_ptr(blocks.contiguous()), n, len(levels), desc, num_faces,
                                     int(bool(has_color)), len(alpha_comp), ac, int(bool(perceptual)),
                                     color_derating, alpha_derating, ratio, _ptr(enc), _ptr(ti), _ptr(src), _ptr(vec),
                                     C.c_void_p(stream)),
          "crn2_hc_tiles_device")
    return enc, ti, src, vec


def compress(blocks, levels, num_faces=1, has_color=True, alpha_comp=(), perceptual=True,
             codebook_sizes=(3072, 3072, 3072, 3072), tile_params=(2.0, 2.0, 3.0)):
    """dxt_hc::compress (crn_dxt_hc.cpp:97-316) on the GPU.  blocks: uint8
    CUDA tensor [n, 16, 4]; levels: (first_block, num_blocks, block_width,
    weight) per level.  Returns (endpoint_indices int16 [n, 4] = colour,
    alpha0, alpha1, reference; selector_indices int16 [n, 3]; colour
    endpoints int32; colour selectors int32; alpha endpoints int32; alpha
    selectors int64) -- the palettes that do not apply are None."""
    n = blocks.shape[0]
    dev = blocks.device
    desc = (C.c_uint32 * (3 * len(levels)))(*[int(x) for lv in levels for x in lv[:3]])
    wts = (C.c_float * len(levels))(*[float(lv[3]) for lv in levels])
    ac = (C.c_uint32 * 2)(*(list(alpha_comp) + [0, 0])[:2])
    sizes = (C.c_uint32 * 4)(*codebook_sizes)
    tp = (C.c_float * 3)(*tile_params)
    counts = (C.c_uint32 * 4)()
    ep = torch.empty((n, 4), dtype=torch.int16, device=dev)
    sel = torch.empty((n, 3), dtype=torch.int16, device=dev)
    ce = torch.empty(codebook_sizes[0], dtype=torch.int32, device=dev) if has_color else None
    cs = torch.empty(codebook_sizes[1], dtype=torch.int32, device=dev) if has_color else None
    ae = torch.empty(codebook_sizes[2], dtype=torch.int32, device=dev) if alpha_comp else None
    as_ = torch.empty(codebook_sizes[3], dtype=torch.int64, device=dev) if alpha_comp else None
    stream = torch.cuda.current_stream().cuda_stream
    check(lib().crn2_hc_compress_device(_ptr(blocks.contiguous()), n, len(levels), desc, wts, num_faces,
                                        int(bool(has_color)), len(alpha_comp), ac, int(bool(perceptual)), sizes, tp,
                                        _ptr(ep), _ptr(sel), _ptr(ce), _ptr(ae), _ptr(cs), _ptr(as_),
                                        counts, C.c_void_p(stream)),
          "crn2_hc_compress_device")
    cut = lambda t, k: None if t is None else t[:k]
    return ep, sel, cut(ce, counts[0]), cut(cs, counts[1]), cut(ae, counts[2]), cut(as_, counts[3])
