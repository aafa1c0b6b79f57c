// encode_set.h -- one element job of a batched texture-set encode, shared by
// the colour and alpha set kernels (dxt1_gpu.hip, dxt5a_gpu.hip) and the plan
// that builds their job tables (texture_api.hip).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace crn2 {

// One level image of one texture, for one element type (DXT1 colour or one
// DXT5 alpha channel).  Elements of all jobs form one flat order; job j owns
// elements [first, first + blocks_x * blocks_y) and writes element b of its
// image at out[b * stride].
struct SetJob {
    const uint32_t* rgba;   // RGBA8 texels, width x height, rows of width texels
    uint64_t* out;          // 8-byte element slot of the job's first block
    uint64_t first;
    uint32_t width, height, blocks_x, stride;
    uint32_t flags;         // colour: crn2_dxt1 flags; alpha: use both block types
    uint32_t comp;          // alpha: channel 0..3
};

// The job of flat element i: the last job whose first element is <= i.
__device__ inline uint32_t set_job_of(const SetJob* __restrict__ jobs, uint32_t njobs, uint64_t i)
{
    uint32_t lo = 0, hi = njobs - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (jobs[mid].first <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

}  // namespace crn2

extern "C" {
// Per-device set-up the colour set kernel needs (the single-colour tables);
// done before any stream capture, since it copies to device memory.
int crn2_dxt1_set_prepare(void);
// Queue the set kernels over flat elements [begin, end): grid workgroups of tb
// threads, each thread taking elements begin + tid, + grid * tb, ... (internal).
int crn2_dxt1_set_launch(const crn2::SetJob* jobs, uint32_t njobs, uint64_t begin, uint64_t end, int quality,
                         uint32_t grid, uint32_t tb, hipStream_t stream);
int crn2_dxt5a_set_launch(const crn2::SetJob* jobs, uint32_t njobs, uint64_t begin, uint64_t end, int quality,
                          uint32_t grid, uint32_t tb, hipStream_t stream);
}
