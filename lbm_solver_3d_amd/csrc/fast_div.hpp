// fast_div.hpp -- division by a run-time constant with a multiply-high, an
// add and a shift, shared by host and device code.
//
// The kernels split a block index into tile coordinates with divisions by
// the tile-grid sizes, which are the same for every block of a launch.  A
// 32-bit integer division compiles to some thirty scalar and vector
// instructions per wavefront; with the divisor's magic number computed once
// on the host it is three scalar instructions.
//
// Method: Granlund and Montgomery, "Division by invariant integers using
// multiplication" (PLDI 1994), Figure 4.1 with N = 32: for d >= 1, let
// l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1 (which fits 32
// bits); then floor(n / d) = floor((mulhi(m, n) + n) / 2^l) for every
// 0 <= n < 2^32.  Here the sum mulhi(m, n) + n is formed in 32 bits, so the
// numerator must stay below 2^31 (tile, row and chunk indices do).
#pragma once
#include <cstdint>

#include "lattice.hpp"

namespace lbmhip {

struct FastDiv {
    uint32_t d, m, l;
};

inline FastDiv make_fast_div(uint32_t d) {
    FastDiv f{d, 0u, 0u};
    if (d == 0) return f;  // not a divisor; quot() is never called with it
    while ((uint64_t(1) << f.l) < d) ++f.l;
    f.m = (uint32_t)(((uint64_t(1) << 32) * ((uint64_t(1) << f.l) - d)) / d + 1);
    return f;
}

// floor(n / d) for 0 <= n < 2^31.
LBM_HD inline uint32_t quot(uint32_t n, const FastDiv& f) {
#if defined(__HIP_DEVICE_COMPILE__)
    const uint32_t hi = __umulhi(n, f.m);
#else
    const uint32_t hi = (uint32_t)(((uint64_t)n * f.m) >> 32);
#endif
    return (hi + n) >> f.l;
}

}  // namespace lbmhip
