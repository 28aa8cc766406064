// params.hpp -- drop-in for the reference's src/utils/params.hpp: struct
// Params with the same fields, defaults and methods (set, loadFile,
// parseCommandLine, validate, print), so the reference input decks run
// unchanged.  Extra keys of this build:
//   precision = fp64 | fp32 | fp16 | bf16   population storage (default fp64)
//   plane     = xy | xz                      orientation / planar plane (default xy)
//   device    = <int>                        GPU ordinal
#pragma once
#include <cstdio>
#include <string>

namespace lbm {

struct Params {
    int NX = 64, NY = 34, NZ = 32;
    double tau = 1.0;
    double Lambda = 3.0 / 16.0;
    double Fx = 0.0, Fy = 0.0, Fz = 0.0;
    bool wallsY = true;
    bool wallsZ = false;
    int steps = 10000;
    int printEvery = 1000;
    int vtkEvery = 0;
    int trajEvery = 50;
    double tol = 0.0;
    int massCorrectionEvery = 10;
    std::string outputDir = "outputs";
    std::string prefix = "run";
    bool squirmer = false;
    double R = 4.0;
    double B1 = 0.015;
    double beta = 0.0;
    double rhoP = 1.0;
    double x0 = -1, y0 = -1, z0 = -1;
    double angle = 30.0;
    bool planar = true;
    double eps = 0.2;
    double hc = 2.0;
    double hmin = 0.5;
    // ---- this build ----
    int precision = 0;      // 0 fp64, 1 fp32, 2 fp16, 3 bf16
    bool planeXZ = false;   // orientation angle and planar motion in x-z
    int device = 0;

    void set(const std::string& key, const std::string& value);
    void loadFile(const std::string& path);
    void parseCommandLine(int argc, char** argv);
    void validate() const;
    void print(std::FILE* fp) const;
};

}  // namespace lbm
