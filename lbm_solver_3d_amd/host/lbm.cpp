// lbm.cpp -- class LBM over the C ABI.  All lattice work runs on the GPU;
// this file only sequences calls and moves data between host mirrors and
// the device.
#include "lbm.hpp"

#include <cmath>
#include <stdexcept>
#include <string>

#include "lbm_hip.h"

namespace lbm {

void LBM::check(int rc) const {
    if (rc != LBM_HIP_OK) throw std::runtime_error(std::string("liblbm_hip: ") + lbm_hip_last_error());
}

static void to_squirmer(const lbm_squirmer_state& s, Squirmer& q) {
    q.R = s.R; q.B1 = s.B1; q.beta = s.beta; q.M = s.M; q.I = s.I; q.FS = s.FS;
    for (int d = 0; d < 3; ++d) {
        q.X[d] = s.X[d]; q.e[d] = s.e[d]; q.U[d] = s.U[d]; q.W[d] = s.W[d];
        q.Fh[d] = s.Fh[d]; q.Th[d] = s.Th[d]; q.Fc[d] = s.Fc[d]; q.Tc[d] = s.Tc[d]; q.Fw[d] = s.Fw[d];
    }
    q.eps = s.eps; q.hc = s.hc; q.hmin = s.hmin;
    q.contacts = (long)s.contacts;
    q.planar = s.planar != 0; q.planeXZ = s.plane_xz != 0;
}

static void from_squirmer(const Squirmer& q, lbm_squirmer_state& s) {
    s.R = q.R; s.B1 = q.B1; s.beta = q.beta; s.M = q.M; s.I = q.I; s.FS = q.FS;
    for (int d = 0; d < 3; ++d) {
        s.X[d] = q.X[d]; s.e[d] = q.e[d]; s.U[d] = q.U[d]; s.W[d] = q.W[d];
        s.Fh[d] = q.Fh[d]; s.Th[d] = q.Th[d]; s.Fc[d] = q.Fc[d]; s.Tc[d] = q.Tc[d]; s.Fw[d] = q.Fw[d];
    }
    s.eps = q.eps; s.hc = q.hc; s.hmin = q.hmin;
    s.contacts = q.contacts;
    s.planar = q.planar; s.plane_xz = q.planeXZ;
}

LBM::LBM(const Params& p)
    : NX(p.NX), NY(p.NY), NZ(p.NZ), N((size_t)p.NX * p.NY * p.NZ), tau(p.tau), F{p.Fx, p.Fy, p.Fz},
      wallsY(p.wallsY), wallsZ(p.wallsZ) {
    lbm_hip_params hp;
    lbm_hip_default_params(&hp);
    hp.nx = p.NX; hp.ny = p.NY; hp.nz = p.NZ;
    hp.tau = p.tau; hp.lambda = p.Lambda; hp.fx = p.Fx; hp.fy = p.Fy; hp.fz = p.Fz;
    hp.walls_y = p.wallsY; hp.walls_z = p.wallsZ;
    hp.squirmer = p.squirmer;
    hp.radius = p.R; hp.b1 = p.B1; hp.beta = p.beta; hp.rho_p = p.rhoP;
    hp.x0 = p.x0; hp.y0 = p.y0; hp.z0 = p.z0; hp.angle = p.angle; hp.planar = p.planar;
    hp.eps = p.eps; hp.hc = p.hc; hp.hmin = p.hmin;
    hp.plane_xz = p.planeXZ; hp.precision = p.precision; hp.device = p.device;
    check(lbm_hip_create(&hp, &ctx_));
    lbm_hip_info info;
    check(lbm_hip_get_info(ctx_, &info));
    omPlus = info.om_plus;
    omMinus = info.om_minus;
    nu = info.nu;
    hasSquirmer = info.has_squirmer != 0;
    syncState();
}

LBM::~LBM() { lbm_hip_destroy(ctx_); }

void LBM::step() {
    int flags = LBM_STEP_ALL;
    if (lazy_) flags = want_outputs_ ? LBM_STEP_ALL : LBM_STEP_MASS;
    want_outputs_ = false;
    check(lbm_hip_step(ctx_, flags));
    ++stepsDone_;
    if (!lazy_) {
        if (hasSquirmer) syncState();  // Fh, Th and the wall force of this step
        else check(lbm_hip_get_wall_force(ctx_, wallForce));
    }
}

void LBM::advanceSquirmer() {
    if (!hasSquirmer) return;
    check(lbm_hip_advance_squirmer(ctx_));
    if (!lazy_) {  // the squirmer moved; the wall force is the step's, unchanged
        lbm_squirmer_state s;
        check(lbm_hip_get_squirmer(ctx_, &s));
        to_squirmer(s, sq);
    }
}

void LBM::correctMass() { check(lbm_hip_correct_mass(ctx_)); }

void LBM::run(int steps, int massEvery, bool outputsOnLastStep) {
    check(lbm_hip_run(ctx_, steps, massEvery, outputsOnLastStep ? LBM_STEP_ALL : 0));
    stepsDone_ += steps;
    if (!lazy_) syncState();
}

void LBM::syncState() {
    lbm_squirmer_state s;
    check(lbm_hip_get_squirmer(ctx_, &s));
    to_squirmer(s, sq);
    if (lbm_hip_get_wall_force(ctx_, wallForce) != LBM_HIP_OK) {
        // the last step did not accumulate the wall force; keep the old value
    }
}

void LBM::pushSquirmer() {
    lbm_squirmer_state s;
    from_squirmer(sq, s);
    check(lbm_hip_set_squirmer(ctx_, &s));
}

void LBM::syncHost(bool populations) const {
    rho.resize(N); ux.resize(N); uy.resize(N); uz.resize(N); flag.resize(N);
    if (populations) f.resize((size_t)d3q19::Q * N);
    check(lbm_hip_download(ctx_, populations ? f.data() : nullptr, rho.data(), ux.data(), uy.data(), uz.data(),
                           flag.data()));
    if (populations) fnew = f;
}

void LBM::pushPopulations() {
    if (f.size() != (size_t)d3q19::Q * N) throw std::runtime_error("pushPopulations: f has the wrong size");
    check(lbm_hip_upload_populations(ctx_, f.data()));
}

void LBM::offsetFromSquirmer(double px, double py, double pz, double r[3]) const {
    r[0] = px - sq.X[0]; r[0] -= NX * std::round(r[0] / NX);
    r[1] = py - sq.X[1]; if (!wallsY) r[1] -= NY * std::round(r[1] / NY);
    r[2] = pz - sq.X[2]; if (!wallsZ) r[2] -= NZ * std::round(r[2] / NZ);
}

long LBM::countFluid() const {
    lbm_hip_info info;
    check(lbm_hip_get_info(ctx_, &info));
    return (long)info.count_fluid;
}
double LBM::meanDensity() const {
    lbm_hip_diag d;
    check(lbm_hip_diagnostics(ctx_, &d));
    return d.mean_density;
}
double LBM::maxSpeed() const {
    lbm_hip_diag d;
    check(lbm_hip_diagnostics(ctx_, &d));
    return d.max_speed;
}
double LBM::sumUx() const {
    lbm_hip_diag d;
    check(lbm_hip_diagnostics(ctx_, &d));
    return d.sum_ux;
}
long LBM::unstableNodes() const {
    lbm_hip_diag d;
    check(lbm_hip_diagnostics(ctx_, &d));
    return (long)d.unstable_nodes;
}

}  // namespace lbm
