// analytic.hpp -- drop-in for the reference's src/utils/analytic.hpp: exact
// Poiseuille solutions and the L2 error of u_x on the cross-section x = 0.
#pragma once
#include <cstdint>
#include <string>

#include "lbm.hpp"

namespace lbm {

// Rectangular duct -a < y < a, -b < z < b driven by force density G along x.
double ductVelocity(double y, double z, double a, double b, double G, double mu);
// Plane channel -a < y < a:  u = G (a^2 - y^2) / (2 mu).
double planeVelocity(double y, double a, double G, double mu);
// Relative L2 error vs the exact solution; writes <base>_profile_{y,z}.dat.
// Calls L.syncHost(false) first.  Negative if there are no walls.
double poiseuilleError(const LBM& L, const std::string& base);
// The same on host arrays of NX * NY * NZ entries (reference layout):
// flags and u_x of the cross-section x = 0, force density G along x.
double poiseuilleErrorOf(int NX, int NY, int NZ, bool wallsY, bool wallsZ, double nu, double G, const uint8_t* flag,
                         const double* ux, const std::string& base);

}  // namespace lbm
