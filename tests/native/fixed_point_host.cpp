// Host build of the fixed-point conversion used by the excess-mass sums
// (lbm_solver_3d_amd/csrc/fixed_point.hpp), for the CPU tests: reads one
// double per argument (decimal or C99 hex-float text) and prints
// "hi lo" of its 128-bit fixed-point form, then the sum of all of them and
// that sum converted back to a double.
#include <cstdio>
#include <cstdlib>

#include "fixed_point.hpp"

int main(int argc, char** argv) {
    using namespace lbmhip;
    u128 total = 0;
    for (int k = 1; k < argc; ++k) {
        const double v = std::strtod(argv[k], nullptr);
        const u128 q = fixed_from_double(v);
        total += q;
        std::printf("%lld %llu\n", (long long)fixed_hi(q), (unsigned long long)fixed_lo(q));
    }
    std::printf("%lld %llu %a\n", (long long)fixed_hi(total), (unsigned long long)fixed_lo(total),
                fixed_to_double(total));
    return 0;
}
