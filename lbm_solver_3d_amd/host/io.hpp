// io.hpp -- drop-in for the reference's src/utils/io.hpp: legacy-VTK field
// output, trajectory rows and the final squirmer state, in the reference's
// byte and text formats (io.cpp:44-97), so its fixtures and readers apply.
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>

#include "lbm.hpp"

namespace lbm {

// STRUCTURED_POINTS, binary big-endian: flag (uint8), density (float64),
// velocity (3 x float64).  Calls L.syncHost(false) first.
void writeVTK(const LBM& L, const std::string& path);
// The same file from host arrays of NX * NY * NZ entries (reference layout).
void writeVTKFields(const std::string& path, int NX, int NY, int NZ, const uint8_t* flag, const double* rho,
                    const double* ux, const double* uy, const double* uz);

// One row every traj_every steps:  t X Y Z ex ey ez Ux Uy Uz Wx Wy Wz Fwall_y Fwall_z
class TrajectoryWriter {
public:
    explicit TrajectoryWriter(const std::string& path);
    ~TrajectoryWriter();
    void write(int t, const Squirmer& s);
    void flush();

private:
    std::FILE* fp_ = nullptr;
};

void writeSquirmerState(const LBM& L, int t, const std::string& path);
// The same file from a squirmer state and the fluid viscosity.
void writeSquirmerStateOf(const Squirmer& s, double nu, int t, const std::string& path);
void ensureDirectory(const std::string& dir);

}  // namespace lbm
