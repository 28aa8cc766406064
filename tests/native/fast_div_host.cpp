// Host build of csrc/fast_div.hpp for tests/test_fast_div_cpu.py.
//   fast_div_host sweep D N   -> number of (n, d), 1 <= d <= D, 0 <= n < N, with quot != n / d
//   fast_div_host pairs       -> reads "n d" lines on stdin, prints quot(n, make_fast_div(d)) per line
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "fast_div.hpp"

int main(int argc, char** argv) {
    using namespace lbmhip;
    if (argc == 4 && std::strcmp(argv[1], "sweep") == 0) {
        const uint32_t D = (uint32_t)std::strtoul(argv[2], nullptr, 10);
        const uint32_t N = (uint32_t)std::strtoul(argv[3], nullptr, 10);
        unsigned long long bad = 0;
        for (uint32_t d = 1; d <= D; ++d) {
            const FastDiv f = make_fast_div(d);
            for (uint32_t n = 0; n < N; ++n) bad += quot(n, f) != n / d;
        }
        std::printf("%llu\n", bad);
        return 0;
    }
    if (argc == 2 && std::strcmp(argv[1], "pairs") == 0) {
        unsigned long n, d;
        while (std::scanf("%lu %lu", &n, &d) == 2) std::printf("%u\n", quot((uint32_t)n, make_fast_div((uint32_t)d)));
        return 0;
    }
    std::fprintf(stderr, "usage: fast_div_host sweep D N | pairs\n");
    return 2;
}
