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
    switch (mode & (STEP_MACRO | STEP_MASS)) {
        case 0: return launch_step_v<S, ODD, 0>(a, k, s);
        case STEP_MACRO: return launch_step_v<S, ODD, STEP_MACRO>(a, k, s);
        case STEP_MASS: return launch_step_v<S, ODD, STEP_MASS>(a, k, s);
        default: return launch_step_v<S, ODD, STEP_MACRO | STEP_MASS>(a, k, s);
    }
}

template <class S>
hipError_t launch_step(const StepArgs<S>& a, const FastCoef& k, bool odd, int mode, hipStream_t s) {
    return odd ? launch_step_p<S, true>(a, k, mode, s) : launch_step_p<S, false>(a, k, mode, s);
}

template <class S, int MODE>
static hipError_t launch_boundary_v(const StepArgs<S>& a, const FastCoef& k, const BoundaryArgs& b, int blocks,
                                    hipStream_t s) {
    if (blocks <= 0) return hipSuccess;
    hipLaunchKernelGGL((aa_boundary_kernel<S, MODE>), dim3(blocks), dim3(BLOCK), 0, s, a, k, b);
    return hipGetLastError();
}

template <class S>
hipError_t launch_boundary(const StepArgs<S>& a, const FastCoef& k, const BoundaryArgs& b, int blocks, int mode,
                           hipStream_t s) {
    switch (mode & (STEP_MACRO | STEP_MASS)) {
        case 0: return launch_boundary_v<S, 0>(a, k, b, blocks, s);
        case STEP_MACRO: return launch_boundary_v<S, STEP_MACRO>(a, k, b, blocks, s);
        case STEP_MASS: return launch_boundary_v<S, STEP_MASS>(a, k, b, blocks, s);
        default: return launch_boundary_v<S, STEP_MACRO | STEP_MASS>(a, k, b, blocks, s);
    }
}

template <class S, bool ODD, int MODE>
static hipError_t launch_batch_step_v(const BatchMember<S>* m, const uint32_t* map, int Gtot, long long t, int xcd,
                                      hipStream_t s) {
    hipLaunchKernelGGL((batch_step_kernel<S, ODD, MODE>), dim3(Gtot), dim3(BLOCK), 0, s, m, map, Gtot, t, xcd);
    return hipGetLastError();
}

template <class S, bool ODD>
static hipError_t launch_batch_step_p(const BatchMember<S>* m, const uint32_t* map, int Gtot, long long t, int xcd,
                                      int mode, hipStream_t s) {
    switch (mode & (STEP_MACRO | STEP_MASS)) {
        case 0: return launch_batch_step_v<S, ODD, 0>(m, map, Gtot, t, xcd, s);
        case STEP_MACRO: return launch_batch_step_v<S, ODD, STEP_MACRO>(m, map, Gtot, t, xcd, s);
        case STEP_MASS: return launch_batch_step_v<S, ODD, STEP_MASS>(m, map, Gtot, t, xcd, s);
        default: return launch_batch_step_v<S, ODD, STEP_MACRO | STEP_MASS>(m, map, Gtot, t, xcd, s);
    }
}

template <class S>
hipError_t launch_batch_step(const BatchMember<S>* m, const uint32_t* map, int Gtot, long long t, int xcd, bool odd,
                             int mode, hipStream_t s) {
    if (Gtot <= 0) return hipSuccess;
    return odd ? launch_batch_step_p<S, true>(m, map, Gtot, t, xcd, mode, s)
               : launch_batch_step_p<S, false>(m, map, Gtot, t, xcd, mode, s);
}

template <class S, int MODE>
static hipError_t launch_batch_boundary_v(const BatchMember<S>* m, const uint32_t* map, int blocks, long long t,
                                          hipStream_t s) {
    hipLaunchKernelGGL((batch_boundary_kernel<S, MODE>), dim3(blocks), dim3(BLOCK), 0, s, m, map, t);
    return hipGetLastError();
}

template <class S>
hipError_t launch_batch_boundary(const BatchMember<S>* m, const uint32_t* map, int blocks, long long t, int mode,
                                 hipStream_t s) {
    if (blocks <= 0) return hipSuccess;
    switch (mode & (STEP_MACRO | STEP_MASS)) {
        case 0: return launch_batch_boundary_v<S, 0>(m, map, blocks, t, s);
        case STEP_MACRO: return launch_batch_boundary_v<S, STEP_MACRO>(m, map, blocks, t, s);
        case STEP_MASS: return launch_batch_boundary_v<S, STEP_MASS>(m, map, blocks, t, s);
        default: return launch_batch_boundary_v<S, STEP_MACRO | STEP_MASS>(m, map, blocks, t, s);
    }
}

}  // namespace lbmhip
