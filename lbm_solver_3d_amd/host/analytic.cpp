// analytic.cpp -- exact fully developed channel and duct flows and the
// cross-section error of u_x on host arrays (drop-in for the reference's
// src/utils/analytic.cpp:11-64; same functions, file formats and values;
// the class LBM overload is in lbm_io.cpp).
//
// Duct -a < y < a, -b < z < b driven by a body-force density G along x
// (F. M. White, Viscous Fluid Flow, 3rd ed., eq. 3-48):
//
//   u(y, z) = 16 a^2 G / (mu pi^3)
//             * sum_{j >= 0} (-1)^j [1 - cosh(k_j z) / cosh(k_j b)] cos(k_j y) / n_j^3,
//   n_j = 2j + 1,  k_j = n_j pi / (2a).
//
// The series is cut after 200 odd harmonics (n <= 399), the truncation the
// reference's printed profiles and errors were made with.
#include "analytic.hpp"

#include <cmath>
#include <cstdio>
#include <vector>

namespace lbm {

namespace {

constexpr double kPi = 3.14159265358979323846;
constexpr int kHarmonics = 200;

// The duct series with its per-harmonic constants evaluated once, for
// evaluation at many points of one cross-section.
class DuctSeries {
public:
    DuctSeries(double a, double b, double G, double mu) : b_(b) {
        amp_ = 16.0 * a * a * G / (mu * kPi * kPi * kPi);
        for (int j = 0; j < kHarmonics; ++j) {
            const double n = 2.0 * j + 1.0;
            k_[j] = n * kPi / (2.0 * a);
            w_[j] = ((j & 1) ? -1.0 : 1.0) / (n * n * n);
            // cosh(k b) = e^{k b} (1 + e^{-2 k b}) / 2: keep the bracket
            den_[j] = 1.0 + std::exp(-2.0 * k_[j] * b);
        }
    }

    double operator()(double y, double z) const {
        const double az = std::fabs(z);
        // highest harmonics (smallest terms) first
        double sum = 0.0;
        for (int j = kHarmonics - 1; j >= 0; --j) {
            // cosh(k z) / cosh(k b) = e^{-k (b - |z|)} (1 + e^{-2 k |z|}) / (1 + e^{-2 k b}),
            // which stays finite where cosh(k b) alone would overflow
            const double ch = std::exp(-k_[j] * (b_ - az)) * (1.0 + std::exp(-2.0 * k_[j] * az)) / den_[j];
            sum += w_[j] * (1.0 - ch) * std::cos(k_[j] * y);
        }
        return amp_ * sum;
    }

private:
    double b_, amp_;
    double k_[kHarmonics], w_[kHarmonics], den_[kHarmonics];
};

void write_profile(const std::string& path, const char* axis, int fixed, int count, double centre,
                   const std::vector<double>& lbm, const std::vector<double>& exact, const std::vector<char>& fluid) {
    std::FILE* fp = std::fopen(path.c_str(), "w");
    if (!fp) return;
    std::fprintf(fp, "# %s-%sc  ux_lbm  ux_exact   (%s = %d)\n", axis, axis, axis[0] == 'y' ? "z" : "y", fixed);
    for (int k = 0; k < count; ++k)
        if (fluid[(size_t)k]) std::fprintf(fp, "%9.3f %.10e %.10e\n", k - centre, lbm[(size_t)k], exact[(size_t)k]);
    std::fclose(fp);
}

}  // namespace

double ductVelocity(double y, double z, double a, double b, double G, double mu) {
    return DuctSeries(a, b, G, mu)(y, z);
}

double planeVelocity(double y, double a, double G, double mu) { return 0.5 * G / mu * (a - y) * (a + y); }

double poiseuilleErrorOf(int NX, int NY, int NZ, bool wallsY, bool wallsZ, double nu, double G, const uint8_t* flag,
                         const double* ux, const std::string& base) {
    if (!wallsY && !wallsZ) return -1.0;
    const double mu = nu;
    auto node = [&](int y, int z) { return (size_t)NX * ((size_t)y + (size_t)NY * z); };  // x = 0
    // walls halfway between the solid layer and the first fluid node
    const double yc = 0.5 * (NY - 1), zc = 0.5 * (NZ - 1);
    const double a = 0.5 * (NY - 2), b = 0.5 * (NZ - 2);

    // exact u_x over the cross-section x = 0
    std::vector<double> ue((size_t)NY * NZ);
    if (wallsY && wallsZ) {
        const DuctSeries duct(a, b, G, mu);
        for (int z = 0; z < NZ; ++z)
            for (int y = 0; y < NY; ++y) ue[(size_t)z * NY + y] = duct(y - yc, z - zc);
    } else {
        for (int z = 0; z < NZ; ++z)
            for (int y = 0; y < NY; ++y)
                ue[(size_t)z * NY + y] = wallsY ? planeVelocity(y - yc, a, G, mu) : planeVelocity(z - zc, b, G, mu);
    }

    // relative L2 norm of u_lbm - u_exact over the fluid nodes
    double err2 = 0.0, ref2 = 0.0;
    for (int z = 0; z < NZ; ++z)
        for (int y = 0; y < NY; ++y) {
            const size_t n = node(y, z);
            if (flag[n] != FLUID) continue;
            const double e = ue[(size_t)z * NY + y], d = ux[n] - e;
            err2 += d * d;
            ref2 += e * e;
        }

    // centre-line profiles through the middle row and column
    const int zm = NZ / 2, ym = NY / 2;
    std::vector<double> lb, ex;
    std::vector<char> fl;
    for (int y = 0; y < NY; ++y) {
        const size_t n = node(y, zm);
        lb.push_back(ux[n]);
        ex.push_back(ue[(size_t)zm * NY + y]);
        fl.push_back(flag[n] == FLUID);
    }
    write_profile(base + "_profile_y.dat", "y", zm, NY, yc, lb, ex, fl);
    lb.clear(); ex.clear(); fl.clear();
    for (int z = 0; z < NZ; ++z) {
        const size_t n = node(ym, z);
        lb.push_back(ux[n]);
        ex.push_back(ue[(size_t)z * NY + ym]);
        fl.push_back(flag[n] == FLUID);
    }
    write_profile(base + "_profile_z.dat", "z", ym, NZ, zc, lb, ex, fl);

    return ref2 > 0.0 ? std::sqrt(err2 / ref2) : -1.0;
}

}  // namespace lbm
