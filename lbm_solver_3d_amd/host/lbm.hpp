// lbm.hpp -- drop-in for the reference's src/utils/lbm.hpp: class LBM with
// the same constructor, methods and public data members, computing on one
// GPU through the C ABI of liblbm_hip.so (include/lbm_hip.h).
//
// Layout of the public arrays is the reference's: f[i * N + n] with
// n = x + NX * (y + NY * z).  The arrays are host mirrors of device state:
// syncHost() refreshes them (populations in reference layout, pre-collision
// moments of the last step, node flags).  wallForce and sq are refreshed
// after every step / advanceSquirmer unless lazy mode is on.
#pragma once
#include <cstdint>
#include <vector>

#include "lattice.hpp"
#include "params.hpp"
#include "squirmer.hpp"

struct lbm_hip_ctx;

namespace lbm {

class LBM {
public:
    explicit LBM(const Params& p);
    ~LBM();
    LBM(const LBM&) = delete;
    LBM& operator=(const LBM&) = delete;

    // One LBM time step: TRT collision (+ Guo forcing) fused with streaming
    // and halfway bounce-back on WALL and PARTICLE links, plus the
    // momentum-exchange force on the walls and the squirmer.
    void step();

    // Squirmer update: soft wall force, Newton-Euler integration, hard gap,
    // and re-mapping of the sphere onto the lattice.
    void advanceSquirmer();

    // Restore the mean fluid density to 1.
    void correctMass();

    // ---- diagnostics (moments of the last step) ----
    long countFluid() const;
    double meanDensity() const;
    double maxSpeed() const;
    double sumUx() const;

    inline size_t idx(int x, int y, int z) const { return x + (size_t)NX * (y + (size_t)NY * z); }
    inline size_t fi(int i, size_t n) const { return (size_t)i * N + n; }

    // Minimum-image offset of a point from the squirmer centre.
    void offsetFromSquirmer(double px, double py, double pz, double r[3]) const;

    // ---- data (public for I/O and tests; host mirrors) ----
    int NX, NY, NZ;
    size_t N;
    double tau, omPlus, omMinus, nu;
    double F[3];
    bool wallsY, wallsZ;

    // mutable: they cache device data, refreshed by syncHost() also through
    // a const LBM& (writeVTK and poiseuilleError take one, as in the reference)
    mutable std::vector<double> f, fnew;
    mutable std::vector<double> rho, ux, uy, uz;
    mutable std::vector<uint8_t> flag;

    double wallForce[3] = {0, 0, 0};

    bool hasSquirmer = false;
    Squirmer sq;

    // ---- this build ----
    // FLUID nodes of the last step's moments with a non-finite or
    // out-of-range density or velocity (lbm_hip_diag::unstable_nodes).
    long unstableNodes() const;
    // Refresh the host mirrors f, rho, ux, uy, uz, flag (fnew mirrors f).
    void syncHost(bool populations = true) const;
    // Upload host-side changes of sq to the device.
    void pushSquirmer();
    // Upload f (reference layout) to the device.
    void pushPopulations();
    // Lazy mode: step() writes only what the coupling needs (no rho/u, no
    // wall force) unless requestOutputs() was called for the next step, and
    // sq / wallForce are refreshed only by syncState().
    void setLazy(bool lazy) { lazy_ = lazy; }
    void requestOutputs() { want_outputs_ = true; }
    void syncState();
    // The driver loop of main.cpp:60-64 on the device (graph-replayed):
    // steps x (step, advanceSquirmer, correctMass every massEvery).
    void run(int steps, int massEvery, bool outputsOnLastStep = true);
    lbm_hip_ctx* context() const { return ctx_; }

private:
    void check(int rc) const;
    lbm_hip_ctx* ctx_ = nullptr;
    bool lazy_ = false;
    bool want_outputs_ = false;
    long long stepsDone_ = 0;
};

}  // namespace lbm
