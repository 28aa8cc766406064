// ref_harness.cpp -- a C entry layer around the reference solver, compiled
// together with the reference's own sources (src/utils/lbm.cpp, squirmer.cpp,
// params.cpp) where they lie, into oracle/_ref/.  TEST INFRASTRUCTURE ONLY:
// it lets the tests read the reference's populations after N steps (the CLI
// only writes rho/u/flag) and time the reference path as the CPU baseline.
//
// Every call maps to one reference entry point:
//   ref_create        -> Params::set (params.cpp:37-75), Params::validate
//                        (params.cpp:120-140), LBM::LBM (lbm.cpp:60-94)
//   ref_step          -> LBM::step            (lbm.cpp:128-224)
//   ref_advance       -> LBM::advanceSquirmer (lbm.cpp:257-263)
//   ref_correct_mass  -> LBM::correctMass     (lbm.cpp:324-339)
//   ref_get_*         -> the public data members (lbm.hpp:65-78)
#include <cstdint>
#include <cstring>
#include <exception>
#include <stdexcept>
#include <string>

#include "lbm.hpp"
#include "params.hpp"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {
thread_local std::string g_err;
}

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

// `kv` is a whitespace-separated list of key=value pairs, applied in order to
// the reference's default Params.
void* ref_create(const char* kv) {
    try {
        lbm::Params p;
        std::string s(kv ? kv : "");
        size_t pos = 0;
        while (pos < s.size()) {
            while (pos < s.size() && (s[pos] == ' ' || s[pos] == '\t' || s[pos] == '\n')) ++pos;
            if (pos >= s.size()) break;
            size_t end = pos;
            while (end < s.size() && s[end] != ' ' && s[end] != '\t' && s[end] != '\n') ++end;
            const std::string tok = s.substr(pos, end - pos);
            const size_t eq = tok.find('=');
            if (eq == std::string::npos) throw std::runtime_error("expected key=value: " + tok);
            p.set(tok.substr(0, eq), tok.substr(eq + 1));
            pos = end;
        }
        p.validate();
        return new lbm::LBM(p);
    } catch (const std::exception& e) {
        g_err = e.what();
        return nullptr;
    }
}

void ref_destroy(void* h) { delete static_cast<lbm::LBM*>(h); }
void ref_step(void* h) { static_cast<lbm::LBM*>(h)->step(); }
void ref_advance(void* h) { static_cast<lbm::LBM*>(h)->advanceSquirmer(); }
void ref_correct_mass(void* h) { static_cast<lbm::LBM*>(h)->correctMass(); }

// The driver loop of main.cpp:60-64 without output.
void ref_run(void* h, int steps, int mass_every) {
    auto* L = static_cast<lbm::LBM*>(h);
    for (int t = 1; t <= steps; ++t) {
        L->step();
        if (L->hasSquirmer) {
            L->advanceSquirmer();
            if (mass_every > 0 && t % mass_every == 0) L->correctMass();
        }
    }
}

void ref_dims(void* h, int* out) {
    auto* L = static_cast<lbm::LBM*>(h);
    out[0] = L->NX; out[1] = L->NY; out[2] = L->NZ; out[3] = L->hasSquirmer ? 1 : 0;
}

void ref_get_fields(void* h, double* f, double* rho, double* ux, double* uy, double* uz, uint8_t* flag) {
    auto* L = static_cast<lbm::LBM*>(h);
    const size_t N = L->N;
    if (f) std::memcpy(f, L->f.data(), sizeof(double) * 19 * N);
    if (rho) std::memcpy(rho, L->rho.data(), sizeof(double) * N);
    if (ux) std::memcpy(ux, L->ux.data(), sizeof(double) * N);
    if (uy) std::memcpy(uy, L->uy.data(), sizeof(double) * N);
    if (uz) std::memcpy(uz, L->uz.data(), sizeof(double) * N);
    if (flag) std::memcpy(flag, L->flag.data(), N);
}

// out[0..26] = X e U W Fh Th Fc Tc Fw, out[27..29] = wallForce, out[30] = contacts,
// out[31..33] = omPlus omMinus nu, out[34..37] = M I FS U0
void ref_get_state(void* h, double* out) {
    auto* L = static_cast<lbm::LBM*>(h);
    const lbm::Squirmer& s = L->sq;
    for (int d = 0; d < 3; ++d) {
        out[0 + d] = s.X[d]; out[3 + d] = s.e[d]; out[6 + d] = s.U[d]; out[9 + d] = s.W[d];
        out[12 + d] = s.Fh[d]; out[15 + d] = s.Th[d]; out[18 + d] = s.Fc[d]; out[21 + d] = s.Tc[d];
        out[24 + d] = s.Fw[d]; out[27 + d] = L->wallForce[d];
    }
    out[30] = (double)s.contacts;
    out[31] = L->omPlus; out[32] = L->omMinus; out[33] = L->nu;
    out[34] = s.M; out[35] = s.I; out[36] = s.FS; out[37] = s.U0();
}

void ref_diagnostics(void* h, double* out) {
    auto* L = static_cast<lbm::LBM*>(h);
    out[0] = (double)L->countFluid();
    out[1] = L->meanDensity();
    out[2] = L->maxSpeed();
    out[3] = L->sumUx();
}

void ref_set_threads(int n) {
#ifdef _OPENMP
    omp_set_num_threads(n);
#else
    (void)n;
#endif
}

}  // extern "C"
