// texture_api.hip -- whole-level DXTn encode (the plain, unclustered path of
// crnlib's dxt_image) and the host-memory block API on top of it.
//
// Reference: dxt_image::init_internal's element layout per format
// (crnlib/crn_dxt_image.cpp:120-200: DXT1 = colour; DXT3 = DXT3 alpha of
// channel 3 + colour; DXT5 = DXT5 alpha of channel 3 + colour; DXT5A = DXT5
// alpha of channel 3; DXN_XY = DXT5 alpha of channels 0 and 1; DXN_YX = of
// channels 1 and 0), set_block_pixels (:1436-1540: 3-colour DXT1 blocks only
// for DXT1 itself, dxt3_block::set_alpha's (a * 15 + 128) / 255,
// crn_dxt.cpp:327-348) and the block compressor of crnlib.cpp:345-410.
// Every element is produced by the GPU searches (crn2_encode_dxt1_image_device,
// crn2_encode_dxt5a_image_device) writing into its slot of the interleaved
// output; DXT3 alpha is one small kernel here.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <vector>

#include "autotune.h"
#include "crn2_internal.h"
#include "encode_set.h"

namespace {

__global__ void __launch_bounds__(256) dxt3_alpha_kernel(const uint32_t* __restrict__ rgba, uint32_t width,
                                                         uint32_t height, uint32_t blocks_x, uint32_t nblocks,
                                                         uint64_t* __restrict__ out, uint32_t stride)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nblocks) return;
    const uint32_t bx = b % blocks_x, by = b / blocks_x;
    uint64_t v = 0;
#pragma unroll
    for (int y = 0; y < 4; y++) {
        const uint32_t sy = min(by * 4 + y, height - 1);
#pragma unroll
        for (int x = 0; x < 4; x++) {
            const uint32_t a = rgba[(size_t)sy * width + min(bx * 4 + x, width - 1)] >> 24;
            v |= (uint64_t)((a * 15u + 128u) / 255u) << (4 * (4 * y + x));
        }
    }
    out[(size_t)b * stride] = v;
}

}  // namespace

extern "C" int crn2_encode_image_device(const void* rgba, uint32_t width, uint32_t height, uint32_t format,
                                        int quality, int flags, void* blocks, void* stream)
{
    if (!rgba || !blocks || !width || !height || quality < 0 || quality > 4 || (flags & ~0xFF1F)) return CRN2_EINVAL;
    if (reinterpret_cast<uintptr_t>(rgba) & 3) return CRN2_EINVAL;
    const uint64_t bx = (width + 3) / 4, n = bx * ((height + 3) / 4);
    if (n > 0xFFFFFFFFull) return CRN2_EINVAL;
    const bool perceptual = flags & 1, both = (flags >> 1) & 1, gray = (flags >> 2) & 1;
    // DXT1 only: DXT1A for texels with alpha below the threshold in bits
    // 8..15, and transparent indices for black texels
    const bool dxt1a = (flags >> 3) & 1, black = (flags >> 4) & 1;
    const int thresh = (flags >> 8) & 255;
    hipStream_t st = static_cast<hipStream_t>(stream);
    uint64_t* out = static_cast<uint64_t*>(blocks);
    const int cq = quality;
    const int cflags = (perceptual ? 1 : 0) | (gray ? 4 : 0);
    switch (format) {
        case CRN2_FMT_DXT1:
            return crn2_encode_dxt1_image_device(rgba, width, height, cq,
                                                 cflags | (both ? 2 : 0) | (dxt1a ? (64 | (thresh << 8)) : 0) |
                                                     (black ? 32 : 0),
                                                 out, 1, nullptr, stream);
        case CRN2_FMT_DXT3: {
            const uint32_t tb = 256;
            hipLaunchKernelGGL(dxt3_alpha_kernel, dim3((uint32_t)((n + tb - 1) / tb)), dim3(tb), 0, st,
                               static_cast<const uint32_t*>(rgba), width, height, (uint32_t)bx, (uint32_t)n, out, 2);
            if (hipGetLastError() != hipSuccess) return CRN2_EHIP;
            return crn2_encode_dxt1_image_device(rgba, width, height, cq, cflags, out + 1, 2, nullptr, stream);
        }
        case CRN2_FMT_DXT5:
        case CRN2_FMT_DXT5_CCXY:
        case CRN2_FMT_DXT5_XGXR:
        case CRN2_FMT_DXT5_XGBR:
        case CRN2_FMT_DXT5_AGBR: {
            const int rc = crn2_encode_dxt5a_image_device(rgba, width, height, 3, quality, both, out, 2, nullptr,
                                                          stream);
            if (rc) return rc;
            return crn2_encode_dxt1_image_device(rgba, width, height, cq, cflags, out + 1, 2, nullptr, stream);
        }
        case CRN2_FMT_DXT5A:
            return crn2_encode_dxt5a_image_device(rgba, width, height, 3, quality, both, out, 1, nullptr, stream);
        case CRN2_FMT_DXN_XY:
        case CRN2_FMT_DXN_YX: {
            const uint32_t c0 = format == CRN2_FMT_DXN_XY ? 0 : 1;
            const int rc = crn2_encode_dxt5a_image_device(rgba, width, height, c0, quality, both, out, 2, nullptr,
                                                          stream);
            if (rc) return rc;
            return crn2_encode_dxt5a_image_device(rgba, width, height, 1 - c0, quality, both, out + 1, 2, nullptr,
                                                  stream);
        }
        default:
            return CRN2_EUNSUPPORTED;
    }
}

extern "C" int crn2_compress_blocks(const uint32_t* pixels, uint32_t nblocks, uint32_t format, int quality,
                                    int flags, void* out)
{
    if (!pixels || !out) return CRN2_EINVAL;
    if (!nblocks) return CRN2_OK;
    if (nblocks > (1u << 26)) return CRN2_EINVAL;
    if (format > CRN2_FMT_DXT5A) return CRN2_EUNSUPPORTED;
    const uint32_t bpb = (format == CRN2_FMT_DXT1 || format == CRN2_FMT_DXT5A) ? 8 : 16;
    // nblocks blocks of 16 consecutive texels = one image 4 texels wide and
    // 4 * nblocks high, whose block column is the blocks in order
    void* d_px = nullptr;
    void* d_out = nullptr;
    const size_t in = (size_t)nblocks * 64, outb = (size_t)nblocks * bpb;
    if (hipMalloc(&d_px, in) != hipSuccess) return CRN2_ENOMEM;
    if (hipMalloc(&d_out, outb) != hipSuccess) {
        (void)hipFree(d_px);
        return CRN2_ENOMEM;
    }
    int rc = CRN2_OK;
    if (hipMemcpy(d_px, pixels, in, hipMemcpyHostToDevice) != hipSuccess) rc = CRN2_EHIP;
    if (rc == CRN2_OK) rc = crn2_encode_image_device(d_px, 4, 4 * nblocks, format, quality, flags, d_out, nullptr);
    if (rc == CRN2_OK && hipMemcpy(out, d_out, outb, hipMemcpyDeviceToHost) != hipSuccess) rc = CRN2_EHIP;
    (void)hipFree(d_px);
    (void)hipFree(d_out);
    return rc;
}

namespace {
// Device staging for the host-memory entry points: frees what it allocated.
struct Staging {
    void* p[8] = {};
    int n = 0;
    ~Staging()
    {
        for (int i = 0; i < n; i++) (void)hipFree(p[i]);
    }
    void* alloc(size_t bytes)
    {
        void* q = nullptr;
        if (n >= 8 || hipMalloc(&q, bytes ? bytes : 1) != hipSuccess) return nullptr;
        p[n++] = q;
        return q;
    }
    void* in(const void* host, size_t bytes)
    {
        void* q = alloc(bytes);
        if (q && bytes && hipMemcpy(q, host, bytes, hipMemcpyHostToDevice) != hipSuccess) return nullptr;
        return q;
    }
};
inline int out(void* host, const void* dev, size_t bytes)
{
    return (!bytes || hipMemcpy(host, dev, bytes, hipMemcpyDeviceToHost) == hipSuccess) ? CRN2_OK : CRN2_EHIP;
}
}  // namespace

extern "C" int crn2_dxt1_blocks(const void* rgba, uint32_t nblocks, int quality, int flags, void* results)
{
    if (!rgba || !results) return CRN2_EINVAL;
    if (!nblocks) return CRN2_OK;
    Staging s;
    void* d_in = s.in(rgba, (size_t)nblocks * 64);
    void* d_out = s.alloc((size_t)nblocks * 24);
    if (!d_in || !d_out) return CRN2_ENOMEM;
    int rc = crn2_dxt1_blocks_device(d_in, nblocks, quality, flags, d_out, nullptr);
    if (rc == CRN2_OK) rc = out(results, d_out, (size_t)nblocks * 24);
    return rc;
}

extern "C" int crn2_dxt1_clusters(const void* rgba, const uint32_t* offsets, uint32_t nclusters, int flags,
                                  void* results, uint8_t* selectors)
{
    if (!rgba || !offsets || !results) return CRN2_EINVAL;
    if (!nclusters) return CRN2_OK;
    const uint32_t total = offsets[nclusters];
    for (uint32_t c = 0; c < nclusters; c++)
        if (offsets[c + 1] < offsets[c]) return CRN2_EINVAL;
    Staging s;
    void* d_px = s.in(rgba, (size_t)total * 4);
    void* d_ofs = s.in(offsets, ((size_t)nclusters + 1) * 4);
    void* d_out = s.alloc((size_t)nclusters * 24);
    uint8_t* d_sel = selectors ? static_cast<uint8_t*>(s.alloc(total)) : nullptr;
    if (!d_px || !d_ofs || !d_out || (selectors && !d_sel)) return CRN2_ENOMEM;
    int rc = crn2_dxt1_clusters_device(d_px, static_cast<const uint32_t*>(d_ofs), nclusters, flags, d_out, d_sel,
                                       nullptr);
    if (rc == CRN2_OK && hipDeviceSynchronize() != hipSuccess) rc = CRN2_EHIP;
    if (rc == CRN2_OK) rc = out(results, d_out, (size_t)nclusters * 24);
    if (rc == CRN2_OK && selectors) rc = out(selectors, d_sel, total);
    return rc;
}

extern "C" int crn2_dxt5a_blocks(const void* values, uint32_t nblocks, int quality, int use_both, void* selectors48,
                                 void* errors, void* meta)
{
    if (!values || !selectors48 || !errors || !meta) return CRN2_EINVAL;
    if (!nblocks) return CRN2_OK;
    Staging s;
    void* d_v = s.in(values, (size_t)nblocks * 16);
    void* d_s = s.alloc((size_t)nblocks * 8);
    void* d_e = s.alloc((size_t)nblocks * 8);
    void* d_m = s.alloc((size_t)nblocks * 4);
    if (!d_v || !d_s || !d_e || !d_m) return CRN2_ENOMEM;
    int rc = crn2_dxt5a_blocks_device(d_v, nblocks, quality, use_both, d_s, d_e, d_m, nullptr);
    if (rc == CRN2_OK) rc = out(selectors48, d_s, (size_t)nblocks * 8);
    if (rc == CRN2_OK) rc = out(errors, d_e, (size_t)nblocks * 8);
    if (rc == CRN2_OK) rc = out(meta, d_m, (size_t)nblocks * 4);
    return rc;
}

extern "C" int crn2_encode_image(const void* rgba, uint32_t width, uint32_t height, uint32_t format, int quality,
                                 int flags, void* blocks)
{
    if (!rgba || !blocks || !width || !height) return CRN2_EINVAL;
    if (format > CRN2_FMT_DXT5A) return CRN2_EUNSUPPORTED;
    const uint32_t bpb = (format == CRN2_FMT_DXT1 || format == CRN2_FMT_DXT5A) ? 8 : 16;
    const size_t n = (size_t)((width + 3) / 4) * ((height + 3) / 4);
    Staging s;
    void* d_in = s.in(rgba, (size_t)width * height * 4);
    void* d_out = s.alloc(n * bpb);
    if (!d_in || !d_out) return CRN2_ENOMEM;
    int rc = crn2_encode_image_device(d_in, width, height, format, quality, flags, d_out, nullptr);
    if (rc == CRN2_OK) rc = out(blocks, d_out, n * bpb);
    return rc;
}

extern "C" int crn2_compress_crn(const uint8_t* const* levels, uint32_t width, uint32_t height, uint32_t nlevels,
                                 uint32_t faces, uint32_t format, int perceptual, const uint32_t* codebook_sizes,
                                 const float* tile_params, uint32_t userdata0, uint32_t userdata1, uint8_t** out_file,
                                 uint32_t* out_size)
{
    if (!levels || !codebook_sizes || !out_file || !out_size) return CRN2_EINVAL;
    *out_file = nullptr;
    *out_size = 0;
    if (!width || !height || width > 4096 || height > 4096 || !nlevels || nlevels > 16 || (faces != 1 && faces != 6))
        return CRN2_EINVAL;
    int has_color = 0;
    uint32_t num_alpha = 0, alpha_comp[2] = {3, 0};
    switch (format) {
        case CRN2_FMT_DXT1: has_color = 1; break;
        case CRN2_FMT_DXT5: has_color = 1; num_alpha = 1; break;
        case CRN2_FMT_DXT5A: num_alpha = 1; break;
        case CRN2_FMT_DXN_XY: num_alpha = 2; alpha_comp[0] = 0; alpha_comp[1] = 1; break;
        case CRN2_FMT_DXN_YX: num_alpha = 2; alpha_comp[0] = 1; alpha_comp[1] = 0; break;
        default: return CRN2_EUNSUPPORTED;
    }
    for (uint32_t i = 0; i < faces * nlevels; i++)
        if (!levels[i]) return CRN2_EINVAL;
    // crn_comp::quantize_images (crnlib/crn_comp.cpp:706-738): per level the
    // faces' blocks over a grid padded to a multiple of 8 texels, edge clamp,
    // level weights min(12, 1.3^level)
    std::vector<uint32_t> desc(3 * nlevels);
    std::vector<float> weight(nlevels);
    size_t total = 0;
    for (uint32_t l = 0; l < nlevels; l++) {
        const uint32_t w = std::max<uint32_t>(1, width >> l), h = std::max<uint32_t>(1, height >> l);
        const uint32_t bw = ((w + 7) & ~7u) >> 2, bh = ((h + 7) & ~7u) >> 2;
        desc[3 * l] = (uint32_t)total;
        desc[3 * l + 1] = bw * bh * faces;
        desc[3 * l + 2] = bw;
        weight[l] = std::min(12.0f, powf(1.3f, (float)l));
        total += (size_t)bw * bh * faces;
    }
    std::vector<uint32_t> blocks(total * 16);
    size_t b = 0;
    for (uint32_t l = 0; l < nlevels; l++) {
        const uint32_t w = std::max<uint32_t>(1, width >> l), h = std::max<uint32_t>(1, height >> l);
        const uint32_t bw = desc[3 * l + 2], bh = desc[3 * l + 1] / (bw * faces);
        for (uint32_t f = 0; f < faces; f++) {
            const uint32_t* img = reinterpret_cast<const uint32_t*>(levels[f * nlevels + l]);
            for (uint32_t by = 0; by < bh; by++)
                for (uint32_t bx = 0; bx < bw; bx++, b++)
                    for (uint32_t t = 0; t < 16; t++) {
                        const uint32_t x = std::min(bx * 4 + (t & 3), w - 1), y = std::min(by * 4 + (t >> 2), h - 1);
                        blocks[b * 16 + t] = img[(size_t)y * w + x];
                    }
        }
    }
    Staging s;
    void* d_blocks = s.in(blocks.data(), total * 64);
    uint16_t* d_ep = static_cast<uint16_t*>(s.alloc(total * 8));
    uint16_t* d_sel = static_cast<uint16_t*>(s.alloc(total * 6));
    uint32_t* d_ce = static_cast<uint32_t*>(s.alloc((size_t)std::max(codebook_sizes[0], 1u) * 4));
    uint32_t* d_ae = static_cast<uint32_t*>(s.alloc((size_t)std::max(codebook_sizes[2], 1u) * 4));
    uint32_t* d_cs = static_cast<uint32_t*>(s.alloc((size_t)std::max(codebook_sizes[1], 1u) * 4));
    uint64_t* d_as = static_cast<uint64_t*>(s.alloc((size_t)std::max(codebook_sizes[3], 1u) * 8));
    if (!d_blocks || !d_ep || !d_sel || !d_ce || !d_ae || !d_cs || !d_as) return CRN2_ENOMEM;
    uint32_t counts[4] = {0, 0, 0, 0};
    int rc = crn2_hc_compress_device(d_blocks, (uint32_t)total, nlevels, desc.data(), weight.data(), faces, has_color,
                                     num_alpha, alpha_comp, perceptual, codebook_sizes, tile_params, d_ep, d_sel,
                                     d_ce, d_ae, d_cs, d_as, counts, nullptr);
    if (rc != CRN2_OK) return rc;
    if (hipDeviceSynchronize() != hipSuccess) return CRN2_EHIP;
    std::vector<uint16_t> ep(total * 4), sel(total * 3);
    std::vector<uint32_t> ce(counts[0] + 1), cs(counts[1] + 1), ae(counts[2] + 1);
    std::vector<uint64_t> as(counts[3] + 1);
    if (out(ep.data(), d_ep, total * 8) || out(sel.data(), d_sel, total * 6) ||
        out(ce.data(), d_ce, (size_t)counts[0] * 4) || out(cs.data(), d_cs, (size_t)counts[1] * 4) ||
        out(ae.data(), d_ae, (size_t)counts[2] * 4) || out(as.data(), d_as, (size_t)counts[3] * 8))
        return CRN2_EHIP;
    std::vector<uint32_t> idx[6];
    for (int k = 0; k < 6; k++) idx[k].resize(total);
    for (size_t i = 0; i < total; i++) {
        idx[0][i] = ep[4 * i];        // colour endpoint
        idx[1][i] = sel[3 * i];       // colour selector
        idx[2][i] = ep[4 * i + 1];    // alpha 0 endpoint
        idx[3][i] = sel[3 * i + 1];   // alpha 0 selector
        idx[4][i] = ep[4 * i + 2];    // alpha 1 endpoint
        idx[5][i] = sel[3 * i + 2];   // alpha 1 selector
    }
    crn2_pack_input in;
    memset(&in, 0, sizeof(in));
    in.width = width;
    in.height = height;
    in.levels = nlevels;
    in.faces = faces;
    in.format = format;
    in.userdata0 = userdata0;
    in.userdata1 = userdata1;
    in.order_palettes = 1;
    if (has_color) {
        in.color_endpoints = ce.data();
        in.n_color_endpoints = counts[0];
        in.color_selectors_linear = cs.data();
        in.n_color_selectors = counts[1];
        in.color_endpoint_idx = idx[0].data();
        in.color_selector_idx = idx[1].data();
    }
    if (num_alpha) {
        in.alpha_endpoints = ae.data();
        in.n_alpha_endpoints = counts[2];
        in.alpha_selectors_linear = as.data();
        in.n_alpha_selectors = counts[3];
        in.alpha0_endpoint_idx = idx[2].data();
        in.alpha0_selector_idx = idx[3].data();
        if (num_alpha > 1) {
            in.alpha1_endpoint_idx = idx[4].data();
            in.alpha1_selector_idx = idx[5].data();
        }
    }
    return crn2_pack(&in, out_file, out_size);
}

// ------------------------------------------------------ texture-set encode
// A whole set of level images (any sizes, DXT1 / DXT5 / DXT5A / DXN formats)
// encoded in two launches: every colour element of the set in one flat order
// (dxt1_set_kernel), then every alpha element (dxt5a_set_kernel), as the
// reference's task pool hands out blocks of one image (crn_dxt_image.cpp:
// 480-485) but across images.  Launch shapes (threads per workgroup x
// workgroups, identical results) come from the autotuner, timed on a sample
// of the set's own elements the first time a plan runs with tuning on.
struct crn2_encode_set {
    uint32_t magic;
    crn2::SetJob* d_jobs[2];   // colour, alpha
    uint32_t njobs[2];
    uint64_t total[2];
    int quality;
    int shape[2];              // index into kSetShapes; -1 = not chosen yet
    int cus;                   // compute units of the device the plan was made on
    hipGraphExec_t graph;
};

namespace {
constexpr uint32_t kSetMagic = 0x5E7E7C0Du;
struct SetShape { uint32_t tb, wg_per_cu; };   // wg_per_cu 0: one thread per element
constexpr SetShape kSetShapes[] = {{256, 0}, {128, 0}, {64, 0}, {256, 8}, {128, 16}, {64, 32}};
constexpr int kNumSetShapes = sizeof(kSetShapes) / sizeof(kSetShapes[0]);
constexpr uint64_t kTuneSample = 1u << 17;   // elements each candidate shape is timed on

int set_launch(const crn2_encode_set* s, int t, int shape, uint64_t begin, uint64_t end, hipStream_t st)
{
    const int cus = s->cus;
    const SetShape sh = kSetShapes[shape];
    // colour elements take one wavefront each, alpha elements one lane
    const uint64_t per = t == 0 ? sh.tb / 64 : sh.tb;
    const uint64_t n = end - begin, full = (n + per - 1) / per;
    const uint64_t cap = sh.wg_per_cu ? (uint64_t)cus * sh.wg_per_cu : full;
    const uint32_t grid = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(full, std::min<uint64_t>(cap, 1u << 30)));
    return t == 0 ? crn2_dxt1_set_launch(s->d_jobs[0], s->njobs[0], begin, end, s->quality, grid, sh.tb, st)
                  : crn2_dxt5a_set_launch(s->d_jobs[1], s->njobs[1], begin, end, s->quality, grid, sh.tb, st);
}

int set_choose_shapes(crn2_encode_set* s, hipStream_t st)
{
    if (s->total[0]) {
        const int rc = crn2_dxt1_set_prepare();
        if (rc) return rc;
    }
    for (int t = 0; t < 2; t++) {
        if (s->shape[t] >= 0 || !s->total[t]) continue;
        const uint64_t end = std::min<uint64_t>(s->total[t], kTuneSample);
        // the tuner keys its choice by kernel and the sample's size class
        s->shape[t] = crn2::tuner().pick(t == 0 ? crn2::TUNE_DXT1_SET : crn2::TUNE_DXT5A_SET, end, kNumSetShapes, 0,
                                         st, [&](int sh) {
                                             return set_launch(s, t, sh, 0, end, st) == CRN2_OK ? hipSuccess
                                                                                                  : hipErrorUnknown;
                                         });
    }
    return CRN2_OK;
}

int set_run(crn2_encode_set* s, hipStream_t st)
{
    for (int t = 0; t < 2; t++)
        if (s->total[t]) {
            const int rc = set_launch(s, t, s->shape[t] < 0 ? 0 : s->shape[t], 0, s->total[t], st);
            if (rc) return rc;
        }
    return CRN2_OK;
}
}  // namespace

extern "C" int crn2_encode_set_create(const crn2_encode_item* items, uint32_t n, int quality, int flags,
                                      crn2_encode_set** out)
{
    if (!out) return CRN2_EINVAL;
    *out = nullptr;
    if (!items || !n || quality < 0 || quality > 4 || (flags & ~0xFF1F)) return CRN2_EINVAL;
    const bool perceptual = flags & 1, both = (flags >> 1) & 1, gray = (flags >> 2) & 1;
    const uint32_t dxt1_extra = (((flags >> 3) & 1) ? (64u | (((uint32_t)flags >> 8 & 255u) << 8)) : 0u) |
                                (((flags >> 4) & 1) ? 32u : 0u);
    std::vector<crn2::SetJob> jobs[2];
    uint64_t total[2] = {0, 0};
    auto add = [&](int t, const crn2_encode_item& it, uint64_t* o, uint32_t stride, uint32_t fl, uint32_t comp) {
        crn2::SetJob j;
        j.rgba = static_cast<const uint32_t*>(it.rgba);
        j.out = o;
        j.first = total[t];
        j.width = it.width;
        j.height = it.height;
        j.blocks_x = (it.width + 3) / 4;
        j.stride = stride;
        j.flags = fl;
        j.comp = comp;
        jobs[t].push_back(j);
        total[t] += (uint64_t)j.blocks_x * ((it.height + 3) / 4);
    };
    const uint32_t cflags = (perceptual ? 1u : 0u) | (gray ? 4u : 0u);
    for (uint32_t i = 0; i < n; i++) {
        const crn2_encode_item& it = items[i];
        if (!it.rgba || !it.blocks || !it.width || !it.height || (reinterpret_cast<uintptr_t>(it.rgba) & 3) ||
            (reinterpret_cast<uintptr_t>(it.blocks) & 7))
            return CRN2_EINVAL;
        uint64_t* o = static_cast<uint64_t*>(it.blocks);
        switch (it.format) {
            case CRN2_FMT_DXT1: add(0, it, o, 1, cflags | (both ? 2u : 0u) | dxt1_extra, 0); break;
            case CRN2_FMT_DXT5:
                add(1, it, o, 2, both, 3);
                add(0, it, o + 1, 2, cflags, 0);
                break;
            case CRN2_FMT_DXT5A: add(1, it, o, 1, both, 3); break;
            case CRN2_FMT_DXN_XY:
            case CRN2_FMT_DXN_YX: {
                const uint32_t c0 = it.format == CRN2_FMT_DXN_XY ? 0 : 1;
                add(1, it, o, 2, both, c0);
                add(1, it, o + 1, 2, both, 1 - c0);
                break;
            }
            default: return CRN2_EUNSUPPORTED;
        }
    }
    crn2_encode_set* s = new (std::nothrow) crn2_encode_set;
    if (!s) return CRN2_ENOMEM;
    s->magic = kSetMagic;
    s->quality = quality;
    s->graph = nullptr;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&s->cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || s->cus <= 0)
        s->cus = 256;
    for (int t = 0; t < 2; t++) {
        s->d_jobs[t] = nullptr;
        s->njobs[t] = (uint32_t)jobs[t].size();
        s->total[t] = total[t];
        s->shape[t] = -1;
        if (!jobs[t].empty()) {
            const size_t bytes = jobs[t].size() * sizeof(crn2::SetJob);
            if (hipMalloc(reinterpret_cast<void**>(&s->d_jobs[t]), bytes) != hipSuccess ||
                hipMemcpy(s->d_jobs[t], jobs[t].data(), bytes, hipMemcpyHostToDevice) != hipSuccess) {
                crn2_encode_set_destroy(s);
                return CRN2_ENOMEM;
            }
        }
    }
    *out = s;
    return CRN2_OK;
}

extern "C" int crn2_encode_set_run(crn2_encode_set* s, void* stream)
{
    if (!s || s->magic != kSetMagic) return CRN2_EINVAL;
    hipStream_t st = static_cast<hipStream_t>(stream);
    const int rc = set_choose_shapes(s, st);
    return rc ? rc : set_run(s, st);
}

extern "C" int crn2_encode_set_run_graph(crn2_encode_set* s, void* stream)
{
    if (!s || s->magic != kSetMagic) return CRN2_EINVAL;
    hipStream_t st = static_cast<hipStream_t>(stream);
    if (!s->graph) {
        int rc = set_choose_shapes(s, st);
        if (rc) return rc;
        if (hipStreamSynchronize(st) != hipSuccess) return CRN2_EHIP;
        hipStream_t cs = nullptr;
        if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return CRN2_EHIP;
        hipGraph_t g = nullptr;
        if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = CRN2_EHIP;
        if (!rc) rc = set_run(s, cs);
        if (hipStreamEndCapture(cs, &g) != hipSuccess && !rc) rc = CRN2_EHIP;
        if (!rc && hipGraphInstantiate(&s->graph, g, nullptr, nullptr, 0) != hipSuccess) {
            s->graph = nullptr;
            rc = CRN2_EHIP;
        }
        if (g) (void)hipGraphDestroy(g);
        (void)hipStreamDestroy(cs);
        if (rc) return rc;
    }
    return hipGraphLaunch(s->graph, st) == hipSuccess ? CRN2_OK : CRN2_EHIP;
}

extern "C" int crn2_encode_set_info(const crn2_encode_set* s, uint64_t* colour_elements, uint64_t* alpha_elements,
                                    int* colour_shape, int* alpha_shape)
{
    if (!s || s->magic != kSetMagic) return CRN2_EINVAL;
    if (colour_elements) *colour_elements = s->total[0];
    if (alpha_elements) *alpha_elements = s->total[1];
    if (colour_shape) *colour_shape = s->shape[0];
    if (alpha_shape) *alpha_shape = s->shape[1];
    return CRN2_OK;
}

extern "C" int crn2_encode_set_destroy(crn2_encode_set* s)
{
    if (!s || s->magic != kSetMagic) return CRN2_EINVAL;
    if (s->graph) (void)hipGraphExecDestroy(s->graph);
    for (int t = 0; t < 2; t++)
        if (s->d_jobs[t]) (void)hipFree(s->d_jobs[t]);
    s->magic = 0;
    delete s;
    return CRN2_OK;
}
