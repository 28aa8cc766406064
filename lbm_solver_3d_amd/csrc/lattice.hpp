// lattice.hpp -- D3Q19 velocity set, weights, node types and flag-byte bits,
// shared by host and device code.
//
// The velocity order, opposite pairs and weights are the reference's
// (src/utils/lattice.hpp:18-29): opposite directions are adjacent, so
// opp(i) = i + 1 for odd i and i - 1 for even i > 0.  The population index i
// therefore means the same direction in every dump and every kernel.
#pragma once
#include <cstdint>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define LBM_HD __host__ __device__
#else
#define LBM_HD
#endif

namespace lbmhip {

constexpr int Q = 19;

// lattice.hpp:20-23 (same values; stored as constexpr tables usable in
// device code through template parameters and constant folding).
constexpr int CX[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
constexpr int CY[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
constexpr int CZ[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
constexpr int OPP[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};

// lattice.hpp:24-27 -- written as the same quotients so the doubles agree.
constexpr double W[Q] = {1.0 / 3.0,
                         1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18,
                         1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36,
                         1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};

// lattice.hpp:36 -- the byte values are also the on-disk VTK values.
enum NodeType : uint8_t { FLUID = 0, WALL = 1, PARTICLE = 2 };

// Device flag byte = NodeType in bits 0-1 plus bookkeeping bits that are
// masked off whenever flags leave the device.
constexpr uint8_t TYPE_MASK = 0x03;
constexpr uint8_t FLIPPED = 0x40;   // node changed type in the remap being applied
constexpr uint8_t BOUNDARY = 0x80;  // some neighbour n + c_i (i = 1..18) is not FLUID

constexpr double kPi = 3.14159265358979323846;

}  // namespace lbmhip
