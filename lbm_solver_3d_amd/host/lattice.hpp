// lattice.hpp -- drop-in for the reference's src/utils/lattice.hpp:
// namespace lbm::d3q19 (Q, cx, cy, cz, opp, w, cs2), NodeType, equilibrium(),
// cross().  Same names, values and population order, so host code written
// against the reference compiles unchanged.
#pragma once
#include <cstdint>

namespace lbm {
namespace d3q19 {
inline constexpr int Q = 19;
inline constexpr int cx[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
inline constexpr int cy[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
inline constexpr int cz[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
inline constexpr int opp[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
inline constexpr double w[Q] = {1.0 / 3.0,
                                1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18,
                                1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36,
                                1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};
inline constexpr double cs2 = 1.0 / 3.0;
}  // namespace d3q19

enum NodeType : uint8_t { FLUID = 0, WALL = 1, PARTICLE = 2 };

// Second-order equilibrium w_i rho (1 + 3 c.u + 4.5 (c.u)^2 - 1.5 u.u).
inline double equilibrium(int i, double rho, double ux, double uy, double uz) {
    using namespace d3q19;
    const double cu = cx[i] * ux + cy[i] * uy + cz[i] * uz;
    const double usq = ux * ux + uy * uy + uz * uz;
    return w[i] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq);
}

inline void cross(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

}  // namespace lbm
