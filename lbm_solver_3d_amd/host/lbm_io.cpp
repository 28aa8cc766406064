// lbm_io.cpp -- the writers and the Poiseuille error taking a class LBM
// (reference io.cpp:44-97, analytic.cpp:25-64 signatures): refresh the host
// mirrors from the device, then call the array forms of io.cpp and
// analytic.cpp.
#include "analytic.hpp"
#include "io.hpp"

namespace lbm {

void writeVTK(const LBM& L, const std::string& path) {
    L.syncHost(false);
    writeVTKFields(path, L.NX, L.NY, L.NZ, L.flag.data(), L.rho.data(), L.ux.data(), L.uy.data(), L.uz.data());
}

void writeSquirmerState(const LBM& L, int t, const std::string& path) { writeSquirmerStateOf(L.sq, L.nu, t, path); }

double poiseuilleError(const LBM& L, const std::string& base) {
    if (!L.wallsY && !L.wallsZ) return -1.0;
    L.syncHost(false);
    return poiseuilleErrorOf(L.NX, L.NY, L.NZ, L.wallsY, L.wallsZ, L.nu, L.F[0], L.flag.data(), L.ux.data(), base);
}

}  // namespace lbm
