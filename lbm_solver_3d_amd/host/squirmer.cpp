// squirmer.cpp -- host-side Squirmer methods with the reference's
// arithmetic (src/utils/squirmer.cpp:20-77); the device runs the same
// update inside LBM::advanceSquirmer.
#include "squirmer.hpp"

#include <cmath>

#include "lattice.hpp"
#include "params.hpp"

namespace lbm {

namespace {
constexpr double kPi = 3.14159265358979323846;
}

void Squirmer::init(const Params& p, double nu, double xc, double yc, double zc) {
    R = p.R; B1 = p.B1; beta = p.beta; planar = p.planar; planeXZ = p.planeXZ;
    M = p.rhoP * 4.0 / 3.0 * kPi * R * R * R;
    I = 0.4 * M * R * R;
    X[0] = xc; X[1] = yc; X[2] = zc;
    const double a = p.angle * kPi / 180.0;
    e[0] = std::cos(a);
    e[1] = planeXZ ? 0.0 : std::sin(a);
    e[2] = planeXZ ? std::sin(a) : 0.0;
    eps = p.eps; hc = p.hc; hmin = p.hmin;
    FS = 6.0 * kPi * nu * R * U0();
}

void Squirmer::surfaceVelocity(const double r[3], double ub[3], bool withSlip) const {
    double wr[3];
    cross(W, r, wr);
    for (int d = 0; d < 3; ++d) ub[d] = U[d] + wr[d];
    if (!withSlip) return;
    const double rn = std::sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    if (rn < 1e-12) return;
    const double rh[3] = {r[0] / rn, r[1] / rn, r[2] / rn};
    const double ct = e[0] * rh[0] + e[1] * rh[1] + e[2] * rh[2];
    const double amp = B1 * (1.0 + beta * ct);
    for (int d = 0; d < 3; ++d) ub[d] += amp * (ct * rh[d] - e[d]);
}

void Squirmer::integrate() {
    for (int d = 0; d < 3; ++d) {
        U[d] += (Fh[d] + Fc[d] + Fw[d]) / M;
        W[d] += (Th[d] + Tc[d]) / I;
        Fc[d] = Tc[d] = 0.0;
    }
    if (planar) {
        if (planeXZ) { U[1] = 0.0; W[0] = 0.0; W[2] = 0.0; }
        else { U[2] = 0.0; W[0] = 0.0; W[1] = 0.0; }
    }
    for (int d = 0; d < 3; ++d) X[d] += U[d];
    double we[3];
    cross(W, e, we);
    for (int d = 0; d < 3; ++d) e[d] += we[d];
    if (planar) {
        if (planeXZ) e[1] = 0.0;
        else e[2] = 0.0;
    }
    const double n = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    for (int d = 0; d < 3; ++d) e[d] /= n;
}

double softWallForce(double h, double eps, double FS, double hc) {
    if (h >= hc) return 0.0;
    const double s = 1.0 - (h < 0.0 ? 0.0 : h) / hc;
    return eps * FS * s * s;
}

}  // namespace lbm
