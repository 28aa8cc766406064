// kernels_fast.hip -- FP32-arithmetic step kernels for FP32 / FP16 / BF16
// population storage (compiled without contraction; the collision's
// multiply-adds are explicit fmaf).
#include "launch_step_impl.hpp"

namespace lbmhip {
template hipError_t launch_step<float>(const StepArgs<float>&, const FastCoef&, bool, int, unsigned, hipStream_t);
template hipError_t step_blocks_per_cu<float>(bool, unsigned, int*);
template hipError_t launch_step<_Float16>(const StepArgs<_Float16>&, const FastCoef&, bool, int, unsigned, hipStream_t);
template hipError_t step_blocks_per_cu<_Float16>(bool, unsigned, int*);
template hipError_t launch_step<bf16_t>(const StepArgs<bf16_t>&, const FastCoef&, bool, int, unsigned, hipStream_t);
template hipError_t step_blocks_per_cu<bf16_t>(bool, unsigned, int*);
template hipError_t launch_boundary<float>(const StepArgs<float>&, const FastCoef&, const BoundaryArgs&, int, int,
                                            hipStream_t);
template hipError_t launch_boundary<_Float16>(const StepArgs<_Float16>&, const FastCoef&, const BoundaryArgs&, int, int,
                                            hipStream_t);
template hipError_t launch_boundary<bf16_t>(const StepArgs<bf16_t>&, const FastCoef&, const BoundaryArgs&, int, int,
                                            hipStream_t);
#define LBMHIP_INSTANTIATE_BATCH(S)                                                                              \
    template hipError_t launch_batch_step<S>(const BatchMember<S>*, const uint32_t*, int, long long, int, bool, \
                                             int, unsigned, hipStream_t);                                       \
    template hipError_t launch_batch_boundary<S>(const BatchMember<S>*, const uint32_t*, int, long long, int,    \
                                                 hipStream_t);                                                  \
    template hipError_t launch_batch_odd_rows<S>(const BatchMember<S>*, const uint32_t*, int, int, long long, int, \
                                                 unsigned, hipStream_t);
LBMHIP_INSTANTIATE_BATCH(float)
LBMHIP_INSTANTIATE_BATCH(_Float16)
LBMHIP_INSTANTIATE_BATCH(bf16_t)
#undef LBMHIP_INSTANTIATE_BATCH
}  // namespace lbmhip
