// autotune.h -- launch-shape autotuner shared by the encoder kernels.
//
// A kernel family offers a few launch shapes (workgroup size, waves per
// workgroup, ...) that compute identical results.  With tuning on, the first
// call for a (kernel, size class) times every shape on the caller's own
// inputs with HIP events and keeps the fastest; later calls in the process
// reuse the choice.  With tuning off the default shape runs.  The candidate
// runs must be idempotent (they overwrite their outputs).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <utility>

namespace crn2 {

enum TuneKernel : int {
    TUNE_HC_SEL = 0, TUNE_DXT1 = 1, TUNE_DXT5A = 2, TUNE_MLP_INFER = 3, TUNE_DXT1_SET = 4, TUNE_DXT5A_SET = 5,
    TUNE_COUNT = 6
};

struct Autotuner {
    std::mutex mu;
    int mode = 0;                                       // 0 off, 1 tune on first use
    std::map<std::pair<int, int>, int> best;            // (kernel, size class) -> shape index
    std::map<std::pair<int, int>, float> best_ms;

    static int size_class(uint64_t n)
    {
        int c = 0;
        while (n > 1) {
            n >>= 1;
            c++;
        }
        return c;
    }

    // run(shape) queues one run of the given shape on `s` and returns a HIP status.
    template <class Run>
    int pick(int kernel, uint64_t n, int nshapes, int default_shape, hipStream_t s, Run&& run)
    {
        const auto key = std::make_pair(kernel, size_class(n));
        {
            std::lock_guard<std::mutex> lk(mu);
            if (!mode) return default_shape;
            auto it = best.find(key);
            if (it != best.end()) return it->second;
        }
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess) return default_shape;
        if (hipEventCreate(&b) != hipSuccess) {
            (void)hipEventDestroy(a);
            return default_shape;
        }
        int choice = default_shape;
        float choice_ms = 0;
        for (int sh = 0; sh < nshapes; sh++) {
            if (run(sh) != hipSuccess) continue;                // warm-up
            if (hipEventRecord(a, s) != hipSuccess) break;
            bool ok = true;
            for (int r = 0; r < 3 && ok; r++) ok = run(sh) == hipSuccess;
            if (!ok || hipEventRecord(b, s) != hipSuccess || hipEventSynchronize(b) != hipSuccess) continue;
            float ms = 0;
            if (hipEventElapsedTime(&ms, a, b) != hipSuccess) continue;
            if (sh == 0 || ms < choice_ms || choice_ms == 0) {
                choice = sh;
                choice_ms = ms;
            }
        }
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        std::lock_guard<std::mutex> lk(mu);
        best[key] = choice;
        best_ms[key] = choice_ms / 3;
        return choice;
    }
};

Autotuner& tuner();

}  // namespace crn2
