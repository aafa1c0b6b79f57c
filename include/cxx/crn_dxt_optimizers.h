/*
 * crn_dxt_optimizers.h (libcrn2) -- crnlib's exported endpoint optimizer
 * classes with the reference's C++ signatures: crnlib::dxt1_endpoint_optimizer
 * (crnlib/crn_dxt1.h:163-352) and crnlib::dxt5_endpoint_optimizer
 * (crnlib/crn_dxt5a.h:31-88).  params and results keep the reference's member
 * names, order and defaults; compute() has the reference's mangled name and
 * runs the GPU search.  The objects hold no search state (the reference's
 * members are its CPU working storage), so any number may share a GPU.
 *
 * compute() covers:
 *   dxt1: 16 texels at any quality (the per-block kernel: both block
 *         types, perceptual or uniform, grayscale sampling, forced 3-colour
 *         blocks, DXT1A transparency with m_pixels_have_alpha and
 *         m_dxt1a_alpha_threshold, m_use_transparent_indices_for_black), and
 *         any number of texels on the path dxt_hc uses (uber, no 3-colour
 *         blocks, no alpha, no grayscale: the cluster kernel).  Other
 *         combinations return false.
 *   dxt5: 16 values, any quality.
 * m_endpoint_caching is ignored: results equal the reference's with caching
 * off.  compute_batch runs n independent calls in one launch.
 */
#ifndef CRN2_CXX_CRN_DXT_OPTIMIZERS_H
#define CRN2_CXX_CRN_DXT_OPTIMIZERS_H

#include "crnlib.h"

namespace crnlib {
typedef unsigned char uint8;
typedef unsigned short uint16;
typedef unsigned int uint32;
typedef uint32 uint;
typedef unsigned long long uint64;

template <typename component_type, typename parameter_type>
class color_quad;
typedef color_quad<uint8, int> color_quad_u8;   // 4 bytes, r first

class dxt1_endpoint_optimizer {
public:
    dxt1_endpoint_optimizer();

    struct params {
        params()
            : m_block_index(0), m_pPixels(nullptr), m_num_pixels(0), m_dxt1a_alpha_threshold(128U),
              m_quality(cCRNDXTQualityUber), m_pixels_have_alpha(false), m_use_alpha_blocks(true),
              m_perceptual(true), m_grayscale_sampling(false), m_endpoint_caching(true),
              m_use_transparent_indices_for_black(false), m_force_alpha_blocks(false)
        {
        }
        uint m_block_index;
        const color_quad_u8* m_pPixels;
        uint m_num_pixels;
        uint m_dxt1a_alpha_threshold;
        crn_dxt_quality m_quality;
        bool m_pixels_have_alpha;
        bool m_use_alpha_blocks;
        bool m_perceptual;
        bool m_grayscale_sampling;
        bool m_endpoint_caching;
        bool m_use_transparent_indices_for_black;
        bool m_force_alpha_blocks;
    };

    struct results {
        results() : m_pSelectors(nullptr) {}
        uint64 m_error;
        uint16 m_low_color;
        uint16 m_high_color;
        uint8* m_pSelectors;
        bool m_alpha_block;
        bool m_reordered;
        bool m_alternate_rounding;
        bool m_enforce_selector;
        uint8 m_enforced_selector;
    };

    bool compute(const params& p, results& r);
    /* n calls in one launch; every p[i] must take the same path (16 texels
     * with equal flags and quality). */
    static bool compute_batch(const params* p, results* r, uint n);

private:
    void* m_reserved;
};

class dxt5_endpoint_optimizer {
public:
    dxt5_endpoint_optimizer();

    struct params {
        params()
            : m_block_index(0), m_pPixels(nullptr), m_num_pixels(0), m_comp_index(3), m_quality(cCRNDXTQualityUber),
              m_use_both_block_types(true)
        {
        }
        uint m_block_index;
        const color_quad_u8* m_pPixels;
        uint m_num_pixels;
        uint m_comp_index;
        crn_dxt_quality m_quality;
        bool m_use_both_block_types;
    };

    struct results {
        uint8* m_pSelectors;
        uint64 m_error;
        uint8 m_first_endpoint;
        uint8 m_second_endpoint;
        uint8 m_block_type;
        bool m_reordered;
    };

    bool compute(const params& p, results& r);
    static bool compute_batch(const params* p, results* r, uint n);

private:
    void* m_reserved;
};
}  // namespace crnlib

#endif  // CRN2_CXX_CRN_DXT_OPTIMIZERS_H
