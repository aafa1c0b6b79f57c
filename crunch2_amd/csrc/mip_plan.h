// mip_plan.h -- host-side set-up of mip generation shared by mip_host.cpp
// (contributor lists, gamma tables, level sizes and resampling order) and
// mip_gpu.hip (device tables and the passes).
#pragma once
#include <stdint.h>

#include <vector>

#include "../../include/crn2.h"

namespace crn2mip {

// The contributors of every destination sample along one axis, as crnlib's
// Resampler::make_clist builds them: pix/wt of destination i are entries
// offs[i] .. offs[i + 1] - 1, in the order the sums visit them.
struct List {
    std::vector<uint32_t> offs;
    std::vector<uint32_t> pix;
    std::vector<float> wt;
    uint32_t ops() const { return (uint32_t)pix.size(); }
};

struct Level {
    uint32_t width, height;
    bool xfirst;  // X pass over the source rows, then Y; else Y then X
    List lx, ly;
};

float filter_support(uint32_t filter);
float filter_value(uint32_t filter, float t);
bool make_list(uint32_t src, uint32_t dst, bool wrap, uint32_t filter, float filter_scale, List& out);
void gamma_tables(float gamma, float* s2l /*256*/, uint8_t* l2s /*8192*/);
uint32_t level_count(uint32_t width, uint32_t height, uint32_t min_size, uint32_t max_levels);
// levels 1 .. n-1 (level 0 is the source); false for invalid parameters
bool plan_levels(uint32_t width, uint32_t height, const crn2_mip_params& p, std::vector<Level>& out);

}  // namespace crn2mip
