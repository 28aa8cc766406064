// fixed_point.hpp -- exact 128-bit fixed-point form of a double, shared by
// host and device code (the excess-mass partial sums).
//
// A value v is held as the two's-complement integer V = trunc(v * 2^64),
// rounded toward zero, in 128 bits: hi = V >> 64 (signed), lo = V mod 2^64,
// so v ~ hi + lo * 2^-64.  The conversion takes the double's sign, biased
// exponent and 53-bit significand apart with bit operations and shifts the
// significand into place; nothing is rounded except the bits below 2^-64,
// so |v - V * 2^-64| < 2^-64 (5.4e-20) for every |v| < 2^63.  NaN and
// infinities convert to 0 and |v| >= 2^63 saturates; neither occurs for the
// excess mass of a node of a finite lattice.  Integer addition is
// associative, so a sum of converted values does not depend on the order or
// grouping of its terms.
#pragma once
#include <cstdint>

#include "lattice.hpp"

namespace lbmhip {

using u128 = unsigned __int128;

LBM_HD inline u128 fixed_from_double(double v) {
    const uint64_t bits = __builtin_bit_cast(uint64_t, v);
    const int ex = (int)((bits >> 52) & 0x7FFu);
    if (ex == 0x7FF || ex == 0) return 0;  // NaN / infinity; zero and subnormals (|v| < 2^-1022)
    const uint64_t m = (bits & ((1ull << 52) - 1)) | (1ull << 52);
    // |v| = m * 2^(ex - 1075), hence |v| * 2^64 = m * 2^(ex - 1011)
    const int sh = ex - 1011;
    u128 a;
    if (sh >= 0) a = sh > 74 ? (((u128)1 << 127) - 1) : ((u128)m << sh);
    else a = sh <= -64 ? (u128)0 : (u128)(m >> (unsigned)(-sh));
    return (bits >> 63) ? (u128)0 - a : a;
}

LBM_HD inline u128 fixed_join(uint64_t hi, uint64_t lo) { return ((u128)hi << 64) | lo; }
LBM_HD inline int64_t fixed_hi(u128 v) { return (int64_t)(uint64_t)(v >> 64); }
LBM_HD inline uint64_t fixed_lo(u128 v) { return (uint64_t)v; }

// hi + lo * 2^-64 as a double (|hi| < 2^53: one rounding per part).
LBM_HD inline double fixed_to_double(u128 v) {
    return (double)fixed_hi(v) + (double)fixed_lo(v) * 0x1p-64;
}

}  // namespace lbmhip
