// launch_step_impl.hpp -- definition of launch_step, included by the two
// kernel translation units that instantiate it.
#pragma once
#include <type_traits>

#include "launch.hpp"

namespace lbmhip {

// Calls fn(std::integral_constant<int, M>) with M the kernel-template mode of
// a step: the caller's MACRO and MASS bits, plus STEP_BODY_FORCE where the
// FP32 collision needs the forcing terms (the FP64 one always has them).
template <class S, class Fn>
static hipError_t with_kernel_mode(int mode, Fn&& fn) {
    using std::integral_constant;
    constexpr int F = STEP_BODY_FORCE;
    if constexpr (!std::is_same<S, double>::value) {
        if (mode & F) {
            switch (mode & (STEP_MACRO | STEP_MASS)) {
                case 0: return fn(integral_constant<int, F>{});
                case STEP_MACRO: return fn(integral_constant<int, STEP_MACRO | F>{});
                case STEP_MASS: return fn(integral_constant<int, STEP_MASS | F>{});
                default: return fn(integral_constant<int, STEP_MACRO | STEP_MASS | F>{});
            }
        }
    }
    switch (mode & (STEP_MACRO | STEP_MASS)) {
        case 0: return fn(integral_constant<int, 0>{});
        case STEP_MACRO: return fn(integral_constant<int, STEP_MACRO>{});
        case STEP_MASS: return fn(integral_constant<int, STEP_MASS>{});
        default: return fn(integral_constant<int, STEP_MACRO | STEP_MASS>{});
    }
}

template <class S, bool ODD, int MODE>
static hipError_t launch_step_v(const StepArgs<S>& a, const FastCoef& k, unsigned lds, hipStream_t s) {
    if (ODD && a.odd_rows) {  // row blocks; the host checked the rows (PX a multiple of 64, at most 1024)
        const int rows = a.NY * a.NZ, rpb = a.rows_per_block;
        const dim3 grid((rows + rpb - 1) / rpb), block(64 * a.row_waves * rpb);
        if (rpb == 1 && (uint32_t)a.NX == a.PX)  // one whole row per block
            hipLaunchKernelGGL((aa_odd_row_kernel<S, MODE, false>), grid, block, lds, s, a, k);
        else
            hipLaunchKernelGGL((aa_odd_row_kernel<S, MODE, true>), grid, block, lds, s, a, k);
        return hipGetLastError();
    }
    const int G = a.gx * a.gy * a.gz;
    hipLaunchKernelGGL((aa_step_kernel<S, ODD, MODE>), dim3(G), dim3(BLOCK), lds, s, a, k);
    return hipGetLastError();
}

template <class S, bool ODD>
static hipError_t launch_step_p(const StepArgs<S>& a, const FastCoef& k, int mode, unsigned lds, hipStream_t s) {
    return with_kernel_mode<S>(mode, [&](auto M) { return launch_step_v<S, ODD, decltype(M)::value>(a, k, lds, s); });
}

template <class S>
hipError_t launch_step(const StepArgs<S>& a, const FastCoef& k, bool odd, int mode, unsigned lds, hipStream_t s) {
    return odd ? launch_step_p<S, true>(a, k, mode, lds, s) : launch_step_p<S, false>(a, k, mode, lds, s);
}

template <class S>
hipError_t step_blocks_per_cu(bool odd, unsigned lds, int* blocks) {
    return odd ? hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, aa_step_kernel<S, true, 0>, BLOCK, lds)
               : hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks, aa_step_kernel<S, false, 0>, BLOCK, lds);
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
    return with_kernel_mode<S>(mode, [&](auto M) { return launch_boundary_v<S, decltype(M)::value>(a, k, b, blocks, s); });
}

template <class S, bool ODD, int MODE>
static hipError_t launch_batch_step_v(const BatchMember<S>* m, const uint32_t* map, int Gtot, long long t, int xcd,
                                      unsigned lds, hipStream_t s) {
    hipLaunchKernelGGL((batch_step_kernel<S, ODD, MODE>), dim3(Gtot), dim3(BLOCK), lds, s, m, map, Gtot, t, xcd);
    return hipGetLastError();
}

template <class S, bool ODD>
static hipError_t launch_batch_step_p(const BatchMember<S>* m, const uint32_t* map, int Gtot, long long t, int xcd,
                                      int mode, unsigned lds, hipStream_t s) {
    return with_kernel_mode<S>(
        mode, [&](auto M) { return launch_batch_step_v<S, ODD, decltype(M)::value>(m, map, Gtot, t, xcd, lds, s); });
}

template <class S>
hipError_t launch_batch_step(const BatchMember<S>* m, const uint32_t* map, int Gtot, long long t, int xcd, bool odd,
                             int mode, unsigned lds, hipStream_t s) {
    if (Gtot <= 0) return hipSuccess;
    return odd ? launch_batch_step_p<S, true>(m, map, Gtot, t, xcd, mode, lds, s)
               : launch_batch_step_p<S, false>(m, map, Gtot, t, xcd, mode, lds, s);
}

template <class S, int MODE>
static hipError_t launch_batch_odd_rows_v(const BatchMember<S>* m, const uint32_t* map, int blocks, int threads,
                                          long long t, unsigned lds, hipStream_t s) {
    hipLaunchKernelGGL((batch_odd_row_kernel<S, MODE>), dim3(blocks), dim3(threads), lds, s, m, map, t);
    return hipGetLastError();
}

template <class S>
hipError_t launch_batch_odd_rows(const BatchMember<S>* m, const uint32_t* map, int blocks, int threads, long long t,
                                 int mode, unsigned lds, hipStream_t s) {
    if (blocks <= 0) return hipSuccess;
    return with_kernel_mode<S>(
        mode, [&](auto M) { return launch_batch_odd_rows_v<S, decltype(M)::value>(m, map, blocks, threads, t, lds, s); });
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
    return with_kernel_mode<S>(
        mode, [&](auto M) { return launch_batch_boundary_v<S, decltype(M)::value>(m, map, blocks, t, s); });
}

}  // namespace lbmhip
