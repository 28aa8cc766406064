// squirmer.hpp -- drop-in for the reference's src/utils/squirmer.hpp: struct
// Squirmer (model, state, forces) and softWallForce.  In class LBM the
// authoritative squirmer state lives on the GPU; LBM::sq is its host mirror.
//
//  Surface slip:  u_s = B1 (1 + beta cos th) (cos th r^ - e),   cos th = e . r^
//  Swimming speed in free space: U0 = 2 B1 / 3.
#pragma once

namespace lbm {

struct Params;

struct Squirmer {
    double R = 4.0, B1 = 0.015, beta = 0.0;
    double M = 0.0, I = 0.0;
    bool planar = true;
    bool planeXZ = false;  // this build: planar motion in x-z instead of x-y

    double X[3] = {0, 0, 0};
    double e[3] = {1, 0, 0};
    double U[3] = {0, 0, 0};
    double W[3] = {0, 0, 0};

    double Fh[3] = {0, 0, 0}, Th[3] = {0, 0, 0};
    double Fc[3] = {0, 0, 0}, Tc[3] = {0, 0, 0};
    double Fw[3] = {0, 0, 0};

    double eps = 0.2, hc = 2.0, hmin = 0.5;
    double FS = 0.0;
    long contacts = 0;

    void init(const Params& p, double nu, double xc, double yc, double zc);
    double U0() const { return 2.0 * B1 / 3.0; }
    void surfaceVelocity(const double r[3], double ub[3], bool withSlip) const;
    void integrate();
};

double softWallForce(double h, double eps, double FS, double hc);

}  // namespace lbm
