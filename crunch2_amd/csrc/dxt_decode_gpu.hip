// dxt_decode_gpu.hip -- DXTn blocks back to RGBA8 texels on gfx950, and the
// per-channel squared error of a DXTn level against its source image.
//
// Reference: crn_decompress_block (crnlib/crnlib.cpp:420-543) with the block
// helpers it calls: dxt1_block::get_block_colors (crn_dxt.cpp:291-300; four
// colours (2a + b) / 3 when c0 > c1, else three colours (a + b) >> 1 and
// transparent black, crn_dxt.cpp:234-260), unpack_color's 565 expansion
// (crn_dxt.cpp:167-181), dxt3_block::get_alpha scaled (crn_dxt.cpp:350-366),
// dxt5_block::get_block_values (crn_dxt.cpp:407-455; eight values when
// l > h, else six plus 0 and 255) and the selector layouts of crn_dxt.h
// (2 bits per texel in row bytes for colour, a 48-bit little-endian field of
// 3-bit selectors for DXT5 alpha).  DXN_XY puts the first alpha block in red
// and the second in green, DXN_YX the reverse, both with blue and alpha 255;
// DXT5A puts its values in alpha over white.  The swizzled DXT5 formats decode
// as plain DXT5, as the reference does.  One deliberate difference: the
// reference's DXT5 case has no `break` and falls through into the DXN code
// (crnlib.cpp:485-503), which overwrites every texel it just wrote; this
// decode ends the DXT5 case where its alpha is written.
//
// The error pass (the "on-GPU PSNR" of the encoders' benchmarks) follows
// error_metrics::compute (crnlib/crn_image_utils.cpp:1048-1120): squared
// differences per channel summed over the texels inside the image; the caller
// forms MSE and PSNR from the four sums.
//
// Layout: one thread per 4x4 block, consecutive threads on consecutive blocks
// of a block row, so each of a block's four texel-row writes (16 bytes) is
// contiguous across the threads of a wave.  The pass is HBM-bound (8 or 16
// bytes read, 64 written per block).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>

#include "crn2_internal.h"

namespace {

__device__ inline uint32_t expand5(uint32_t v) { return (v << 3) | (v >> 2); }
__device__ inline uint32_t expand6(uint32_t v) { return (v << 2) | (v >> 4); }

// The four colours of a DXT1 colour block as RGBA8 words (r in the low byte).
__device__ inline void dxt1_colors(uint32_t c0, uint32_t c1, uint32_t* pal)
{
    const uint32_t r0 = expand5(c0 >> 11), g0 = expand6((c0 >> 5) & 63), b0 = expand5(c0 & 31);
    const uint32_t r1 = expand5(c1 >> 11), g1 = expand6((c1 >> 5) & 63), b1 = expand5(c1 & 31);
    pal[0] = r0 | (g0 << 8) | (b0 << 16) | 0xFF000000u;
    pal[1] = r1 | (g1 << 8) | (b1 << 16) | 0xFF000000u;
    if (c0 > c1) {
        pal[2] = ((r0 * 2 + r1) / 3) | (((g0 * 2 + g1) / 3) << 8) | (((b0 * 2 + b1) / 3) << 16) | 0xFF000000u;
        pal[3] = ((r1 * 2 + r0) / 3) | (((g1 * 2 + g0) / 3) << 8) | (((b1 * 2 + b0) / 3) << 16) | 0xFF000000u;
    } else {
        pal[2] = ((r0 + r1) >> 1) | (((g0 + g1) >> 1) << 8) | (((b0 + b1) >> 1) << 16) | 0xFF000000u;
        pal[3] = 0;
    }
}

// The eight values of a DXT5 alpha block.
__device__ inline void dxt5_values(uint32_t l, uint32_t h, uint32_t* v)
{
    v[0] = l;
    v[1] = h;
    if (l > h) {
        v[2] = (l * 6 + h) / 7;
        v[3] = (l * 5 + h * 2) / 7;
        v[4] = (l * 4 + h * 3) / 7;
        v[5] = (l * 3 + h * 4) / 7;
        v[6] = (l * 2 + h * 5) / 7;
        v[7] = (l + h * 6) / 7;
    } else {
        v[2] = (l * 4 + h) / 5;
        v[3] = (l * 3 + h * 2) / 5;
        v[4] = (l * 2 + h * 3) / 5;
        v[5] = (l + h * 4) / 5;
        v[6] = 0;
        v[7] = 255;
    }
}

// kind: 0 DXT1, 1 DXT3, 2 DXT5 (and its swizzles), 3 DXN_XY, 4 DXN_YX, 5 DXT5A
__device__ inline void decode_block(int kind, const uint32_t* w, uint32_t* t)
{
    if (kind <= 2) {
        const uint32_t* cw = kind == 0 ? w : w + 2;
        uint32_t pal[4];
        dxt1_colors(cw[0] & 0xFFFFu, cw[0] >> 16, pal);
#pragma unroll
        for (int i = 0; i < 16; i++) t[i] = pal[(cw[1] >> (2 * i)) & 3u];
        if (kind == 1) {
#pragma unroll
            for (int i = 0; i < 16; i++) {
                const uint32_t a = (w[i >> 3] >> (4 * (i & 7))) & 15u;
                t[i] = (t[i] & 0x00FFFFFFu) | (((a << 4) | a) << 24);
            }
        } else if (kind == 2) {
            uint32_t v[8];
            dxt5_values(w[0] & 0xFF, (w[0] >> 8) & 0xFF, v);
            const uint64_t s = ((uint64_t)w[1] << 16) | (w[0] >> 16);
#pragma unroll
            for (int i = 0; i < 16; i++) t[i] = (t[i] & 0x00FFFFFFu) | (v[(s >> (3 * i)) & 7u] << 24);
        }
    } else if (kind <= 4) {
        uint32_t v0[8], v1[8];
        dxt5_values(w[0] & 0xFF, (w[0] >> 8) & 0xFF, v0);
        dxt5_values(w[2] & 0xFF, (w[2] >> 8) & 0xFF, v1);
        const uint64_t s0 = ((uint64_t)w[1] << 16) | (w[0] >> 16);
        const uint64_t s1 = ((uint64_t)w[3] << 16) | (w[2] >> 16);
#pragma unroll
        for (int i = 0; i < 16; i++) {
            const uint32_t a = v0[(s0 >> (3 * i)) & 7u], b = v1[(s1 >> (3 * i)) & 7u];
            t[i] = (kind == 3 ? (a | (b << 8)) : (b | (a << 8))) | 0xFFFF0000u;
        }
    } else {
        uint32_t v[8];
        dxt5_values(w[0] & 0xFF, (w[0] >> 8) & 0xFF, v);
        const uint64_t s = ((uint64_t)w[1] << 16) | (w[0] >> 16);
#pragma unroll
        for (int i = 0; i < 16; i++) t[i] = 0x00FFFFFFu | (v[(s >> (3 * i)) & 7u] << 24);
    }
}

struct DecodeArgs {
    const uint8_t* blocks;
    uint64_t block_pitch;      // bytes from one block row to the next
    uint32_t bx, by, width, height, kind, bpb;
    uint8_t* rgba;             // decoded texels (may be null)
    uint64_t rgba_pitch;
    const uint8_t* src;        // source texels for the error sums (may be null)
    uint64_t src_pitch;
    unsigned long long* sse;   // [4], accumulated
};

__global__ void __launch_bounds__(256) dxt_decode_kernel(DecodeArgs a)
{
    const uint32_t b = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n = a.bx * a.by;
    uint64_t e[4] = {0, 0, 0, 0};
    if (b < n) {
        const uint32_t xb = b % a.bx, yb = b / a.bx;
        const uint8_t* p = a.blocks + (uint64_t)yb * a.block_pitch + (uint64_t)xb * a.bpb;
        uint32_t w[4];
        if (a.bpb == 16) {
            const uint4 v = *reinterpret_cast<const uint4*>(p);
            w[0] = v.x;
            w[1] = v.y;
            w[2] = v.z;
            w[3] = v.w;
        } else {
            const uint2 v = *reinterpret_cast<const uint2*>(p);
            w[0] = v.x;
            w[1] = v.y;
            w[2] = w[3] = 0;
        }
        uint32_t t[16];
        decode_block((int)a.kind, w, t);
        const uint32_t x0 = xb * 4, y0 = yb * 4;
        const bool full = x0 + 4 <= a.width;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const uint32_t y = y0 + r;
            if (y >= a.height) break;
            if (a.rgba) {
                uint32_t* d = reinterpret_cast<uint32_t*>(a.rgba + (uint64_t)y * a.rgba_pitch) + x0;
                if (full && !(reinterpret_cast<uintptr_t>(d) & 15)) {
                    *reinterpret_cast<uint4*>(d) = make_uint4(t[4 * r], t[4 * r + 1], t[4 * r + 2], t[4 * r + 3]);
                } else {
                    for (int c = 0; c < 4; c++)
                        if (x0 + c < a.width) d[c] = t[4 * r + c];
                }
            }
            if (a.src) {
                const uint32_t* s = reinterpret_cast<const uint32_t*>(a.src + (uint64_t)y * a.src_pitch) + x0;
                for (int c = 0; c < 4; c++) {
                    if (x0 + c >= a.width) break;
                    const uint32_t u = s[c], v = t[4 * r + c];
#pragma unroll
                    for (int k = 0; k < 4; k++) {
                        const int d = (int)((u >> (8 * k)) & 0xFF) - (int)((v >> (8 * k)) & 0xFF);
                        e[k] += (uint64_t)(d * d);
                    }
                }
            }
        }
    }
    if (a.sse) {
        // wave sums, then one atomic per wave and channel
#pragma unroll
        for (int k = 0; k < 4; k++) {
            uint64_t v = e[k];
            for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
            if ((threadIdx.x & 63) == 0 && v) atomicAdd(&a.sse[k], (unsigned long long)v);
        }
    }
}

int kind_of(uint32_t format, uint32_t* bpb)
{
    switch (format) {
        case CRN2_FMT_DXT1: *bpb = 8; return 0;
        case CRN2_FMT_DXT3: *bpb = 16; return 1;
        case CRN2_FMT_DXT5: case CRN2_FMT_DXT5_CCXY: case CRN2_FMT_DXT5_XGXR: case CRN2_FMT_DXT5_XGBR:
        case CRN2_FMT_DXT5_AGBR: *bpb = 16; return 2;
        case CRN2_FMT_DXN_XY: *bpb = 16; return 3;
        case CRN2_FMT_DXN_YX: *bpb = 16; return 4;
        case CRN2_FMT_DXT5A: *bpb = 8; return 5;
        default: return -1;
    }
}

int launch(const void* blocks, uint32_t format, uint32_t width, uint32_t height, uint32_t block_row_pitch,
           void* rgba, uint32_t rgba_row_pitch, const void* src, uint32_t src_row_pitch, uint64_t* sse,
           hipStream_t st)
{
    uint32_t bpb = 0;
    const int kind = kind_of(format, &bpb);
    if (kind < 0) return CRN2_EUNSUPPORTED;
    if (!blocks || !width || !height) return CRN2_EINVAL;
    if (reinterpret_cast<uintptr_t>(blocks) & (bpb - 1)) return CRN2_EINVAL;
    DecodeArgs a;
    a.bx = (width + 3) / 4;
    a.by = (height + 3) / 4;
    a.block_pitch = block_row_pitch ? block_row_pitch : (uint64_t)a.bx * bpb;
    if (a.block_pitch < (uint64_t)a.bx * bpb || (a.block_pitch & (bpb - 1))) return CRN2_EINVAL;
    a.width = width;
    a.height = height;
    a.kind = (uint32_t)kind;
    a.bpb = bpb;
    a.blocks = static_cast<const uint8_t*>(blocks);
    a.rgba = static_cast<uint8_t*>(rgba);
    a.rgba_pitch = rgba_row_pitch ? rgba_row_pitch : (uint64_t)width * 4;
    a.src = static_cast<const uint8_t*>(src);
    a.src_pitch = src_row_pitch ? src_row_pitch : (uint64_t)width * 4;
    a.sse = reinterpret_cast<unsigned long long*>(sse);
    if ((a.rgba && (a.rgba_pitch < (uint64_t)width * 4 || (a.rgba_pitch & 3) ||
                    (reinterpret_cast<uintptr_t>(rgba) & 3))) ||
        (a.src && (a.src_pitch < (uint64_t)width * 4 || (a.src_pitch & 3) || (reinterpret_cast<uintptr_t>(src) & 3))))
        return CRN2_EINVAL;
    const uint64_t n = (uint64_t)a.bx * a.by;
    const uint32_t tb = 256;
    hipLaunchKernelGGL(dxt_decode_kernel, dim3((uint32_t)((n + tb - 1) / tb)), dim3(tb), 0, st, a);
    return hipGetLastError() == hipSuccess ? CRN2_OK : CRN2_EHIP;
}

}  // namespace

extern "C" int crn2_decode_dxt_device(const void* blocks, uint32_t format, uint32_t width, uint32_t height,
                                      uint32_t block_row_pitch, void* rgba, uint32_t rgba_row_pitch, void* stream)
{
    if (!rgba) return CRN2_EINVAL;
    return launch(blocks, format, width, height, block_row_pitch, rgba, rgba_row_pitch, nullptr, 0, nullptr,
                  static_cast<hipStream_t>(stream));
}

extern "C" int crn2_dxt_error_device(const void* blocks, uint32_t format, uint32_t width, uint32_t height,
                                     uint32_t block_row_pitch, const void* src_rgba, uint32_t src_row_pitch,
                                     uint64_t* sse, void* stream)
{
    if (!src_rgba || !sse) return CRN2_EINVAL;
    return launch(blocks, format, width, height, block_row_pitch, nullptr, 0, src_rgba, src_row_pitch, sse,
                  static_cast<hipStream_t>(stream));
}

extern "C" int crn2_decompress_blocks(const void* blocks, uint32_t nblocks, uint32_t format, uint32_t* rgba)
{
    uint32_t bpb = 0;
    if (kind_of(format, &bpb) < 0) return CRN2_EUNSUPPORTED;
    if (!blocks || !rgba) return CRN2_EINVAL;
    if (!nblocks) return CRN2_OK;
    if (nblocks > (1u << 26)) return CRN2_EINVAL;
    // blocks laid out as one column of nblocks blocks: a 4 x 4n image, 16
    // consecutive texels per block
    void* d_blocks = nullptr;
    void* d_rgba = nullptr;
    const size_t in = (size_t)nblocks * bpb, out = (size_t)nblocks * 64;
    if (hipMalloc(&d_blocks, in) != hipSuccess) return CRN2_ENOMEM;
    if (hipMalloc(&d_rgba, out) != hipSuccess) {
        (void)hipFree(d_blocks);
        return CRN2_ENOMEM;
    }
    int rc = CRN2_OK;
    if (hipMemcpy(d_blocks, blocks, in, hipMemcpyHostToDevice) != hipSuccess) rc = CRN2_EHIP;
    if (rc == CRN2_OK) rc = launch(d_blocks, format, 4, 4 * nblocks, bpb, d_rgba, 16, nullptr, 0, nullptr, nullptr);
    if (rc == CRN2_OK && hipMemcpy(rgba, d_rgba, out, hipMemcpyDeviceToHost) != hipSuccess) rc = CRN2_EHIP;
    (void)hipFree(d_blocks);
    (void)hipFree(d_rgba);
    return rc;
}
