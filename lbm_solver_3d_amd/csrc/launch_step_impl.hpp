// launch_step_impl.hpp -- definition of launch_step, included by the two
// kernel translation units that instantiate it.
#pragma once
#include "launch.hpp"

namespace lbmhip {

template <class S, bool ODD, int MODE>
static hipError_t launch_step_v(const StepArgs<S>& a, const FastCoef& k, hipStream_t s) {
    const int G = a.gx * a.gy * a.gz;
    hipLaunchKernelGGL((aa_step_kernel<S, ODD, MODE>), dim3(G), dim3(BLOCK), 0, s, a, k);
    return hipGetLastError();
}

template <class S, bool ODD>
static hipError_t launch_step_p(const StepArgs<S>& a, const FastCoef& k, int mode, hipStream_t s) {
    switch (mode & 7) {
        case 0: return launch_step_v<S, ODD, 0>(a, k, s);
        case 1: return launch_step_v<S, ODD, 1>(a, k, s);
        case 2: return launch_step_v<S, ODD, 2>(a, k, s);
        case 3: return launch_step_v<S, ODD, 3>(a, k, s);
        case 4: return launch_step_v<S, ODD, 4>(a, k, s);
        case 5: return launch_step_v<S, ODD, 5>(a, k, s);
        case 6: return launch_step_v<S, ODD, 6>(a, k, s);
        default: return launch_step_v<S, ODD, 7>(a, k, s);
    }
}

template <class S>
hipError_t launch_step(const StepArgs<S>& a, const FastCoef& k, bool odd, int mode, hipStream_t s) {
    return odd ? launch_step_p<S, true>(a, k, mode, s) : launch_step_p<S, false>(a, k, mode, s);
}

}  // namespace lbmhip
