// storage.hpp -- population storage formats.
//
//   double : f_i itself, FP64 arithmetic in the reference's operation order.
//   float  : g_i = f_i - w_i, FP32 arithmetic.
//   half   : g_i * 2^15 as IEEE binary16, FP32 arithmetic.
//   bf16   : g_i as bfloat16 (round to nearest even), FP32 arithmetic.
//
// Storing the deviation from the rest equilibrium w_i instead of f_i keeps
// the significant bits of the small non-equilibrium part (the "DDF shifting"
// of M. Lehmann et al., Phys. Rev. E 106, 015308 (2022)).  The 2^15 scale of
// the half format moves |g| ~ 1e-8..1e-1 into the normal range of binary16
// (6.1e-5 .. 65504), so |f_i - w_i| < 2 is representable.
#pragma once
#include <cstdint>

#include "lattice.hpp"

namespace lbmhip {

struct bf16_t {
    uint16_t bits;
};

template <class S>
struct Store;

template <>
struct Store<double> {
    using C = double;
    static constexpr bool shifted = false;
    static constexpr double scale = 1.0;
    LBM_HD static inline C load(double v) { return v; }
    LBM_HD static inline C load_add(double v, C c, float) { return v + c; }
    LBM_HD static inline double store(C v) { return v; }
};

template <>
struct Store<float> {
    using C = float;
    static constexpr bool shifted = true;
    static constexpr double scale = 1.0;
    LBM_HD static inline C load(float v) { return v; }
    LBM_HD static inline C load_add(float v, C c, float) { return v + c; }
    LBM_HD static inline float store(C v) { return v; }
};

template <>
struct Store<_Float16> {
    using C = float;
    static constexpr bool shifted = true;
    static constexpr double scale = 32768.0;
    LBM_HD static inline C load(_Float16 v) { return (float)v; }
    // (float)v + c, rounded once.  `one` is 1.0f held in a register: with a
    // literal 1 the compiler folds the multiply away and emits a conversion
    // and an add; this way it emits one mixed-precision fused multiply-add.
    LBM_HD static inline C load_add(_Float16 v, C c, float one) { return __builtin_fmaf((float)v, one, c); }
    LBM_HD static inline _Float16 store(C v) { return (_Float16)v; }
};

template <>
struct Store<bf16_t> {
    using C = float;
    static constexpr bool shifted = true;
    static constexpr double scale = 1.0;
    LBM_HD static inline C load(bf16_t v) {
        const uint32_t u = (uint32_t)v.bits << 16;
        return __builtin_bit_cast(float, u);
    }
    LBM_HD static inline C load_add(bf16_t v, C c, float) { return load(v) + c; }
    LBM_HD static inline bf16_t store(C v) {
        uint32_t u = __builtin_bit_cast(uint32_t, v);
        u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest, ties to even (finite inputs)
        return bf16_t{(uint16_t)(u >> 16)};
    }
};

// Real population value <-> stored value (host side, for uploads/downloads
// and the remap's double-precision paths).
template <class S>
LBM_HD inline double to_real(S v, int i) {
    using St = Store<S>;
    const double g = (double)St::load(v);
    if constexpr (St::shifted) return W[i] + g * (1.0 / St::scale);
    else return g;
}
template <class S>
LBM_HD inline S from_real(double f, int i) {
    using St = Store<S>;
    if constexpr (St::shifted) return St::store((typename St::C)((f - W[i]) * St::scale));
    else return St::store(f);
}

}  // namespace lbmhip
