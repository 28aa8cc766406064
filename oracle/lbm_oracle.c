/*
 * lbm_oracle.c -- CPU FP64 restatement of the reference D3Q19 TRT lattice
 * Boltzmann step and squirmer coupling.  TEST INFRASTRUCTURE ONLY (the
 * checker): see lbm_oracle.h.
 *
 * Every floating-point expression keeps the evaluation order of the
 * reference source line it cites, and the file is compiled with
 * -ffp-contract=off so no multiply-add is fused.  Built that way, a
 * single-threaded run is bit-identical to the reference compiled from its
 * own sources with the same flag.
 */
#include "lbm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define Q 19
static const double PI_ = 3.14159265358979323846;

/* lattice.hpp:20-27 -- opposite pairs stored adjacently. */
static const int cx[Q] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int cy[Q] = {0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int cz[Q] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1};
static const int opp[Q] = {0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17};
static const double w[Q] = {1.0 / 3.0,
                            1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18, 1.0 / 18,
                            1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36,
                            1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36, 1.0 / 36};

enum { FLUID = 0, WALL = 1, PARTICLE = 2 }; /* lattice.hpp:36 */

int orc_cx(int i) { return cx[i]; }
int orc_cy(int i) { return cy[i]; }
int orc_cz(int i) { return cz[i]; }
int orc_opp(int i) { return opp[i]; }
double orc_w(int i) { return w[i]; }

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static inline size_t idx(const orc_lbm* L, int x, int y, int z) {
    return (size_t)x + (size_t)L->NX * ((size_t)y + (size_t)L->NY * (size_t)z);
}

/* lattice.hpp:52-56 */
static void cross3(const double a[3], const double b[3], double c[3]) {
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

/* lattice.hpp:42-47 */
double orc_equilibrium(int i, double rho, double ux, double uy, double uz) {
    const double cu = cx[i] * ux + cy[i] * uy + cz[i] * uz;
    const double usq = ux * ux + uy * uy + uz * uz;
    return w[i] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq);
}

/* params.hpp:16-56 defaults */
void orc_default_params(orc_params* p) {
    memset(p, 0, sizeof *p);
    p->nx = 64; p->ny = 34; p->nz = 32;
    p->tau = 1.0; p->lambda = 3.0 / 16.0;
    p->fx = p->fy = p->fz = 0.0;
    p->walls_y = 1; p->walls_z = 0;
    p->squirmer = 0;
    p->radius = 4.0; p->b1 = 0.015; p->beta = 0.0; p->rho_p = 1.0;
    p->x0 = -1; p->y0 = -1; p->z0 = -1;
    p->angle = 30.0; p->planar = 1;
    p->eps = 0.2; p->hc = 2.0; p->hmin = 0.5;
    p->plane_xz = 0;
}

/* squirmer.cpp:35-46 */
void orc_surface_velocity(const orc_squirmer* s, const double r[3], double ub[3], int with_slip) {
    double wr[3];
    cross3(s->W, r, wr);
    for (int d = 0; d < 3; ++d) ub[d] = s->U[d] + wr[d];
    if (!with_slip) return;
    const double rn = sqrt(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
    if (rn < 1e-12) return;
    const double rh[3] = {r[0] / rn, r[1] / rn, r[2] / rn};
    const double ct = s->e[0] * rh[0] + s->e[1] * rh[1] + s->e[2] * rh[2];
    const double amp = s->B1 * (1.0 + s->beta * ct);
    for (int d = 0; d < 3; ++d) ub[d] += amp * (ct * rh[d] - s->e[d]);
}

/* squirmer.cpp:52-68 (plus the x-z plane extension for `planar`) */
void orc_integrate(orc_squirmer* s) {
    for (int d = 0; d < 3; ++d) {
        s->U[d] += (s->Fh[d] + s->Fc[d] + s->Fw[d]) / s->M;
        s->W[d] += (s->Th[d] + s->Tc[d]) / s->I;
        s->Fc[d] = s->Tc[d] = 0.0;
    }
    if (s->planar) {
        if (s->plane_xz) { s->U[1] = 0.0; s->W[0] = 0.0; s->W[2] = 0.0; }
        else { s->U[2] = 0.0; s->W[0] = 0.0; s->W[1] = 0.0; }
    }
    for (int d = 0; d < 3; ++d) s->X[d] += s->U[d];
    double we[3];
    cross3(s->W, s->e, we);
    for (int d = 0; d < 3; ++d) s->e[d] += we[d];
    if (s->planar) {
        if (s->plane_xz) s->e[1] = 0.0;
        else s->e[2] = 0.0;
    }
    const double n = sqrt(s->e[0] * s->e[0] + s->e[1] * s->e[1] + s->e[2] * s->e[2]);
    for (int d = 0; d < 3; ++d) s->e[d] /= n;
}

/* squirmer.cpp:73-77 */
double orc_soft_wall_force(double h, double eps, double FS, double hc) {
    if (h >= hc) return 0.0;
    const double hp = (h < 0.0) ? 0.0 : h; /* std::max(h, 0.0) */
    const double s = 1.0 - hp / hc;
    return eps * FS * s * s;
}

/* squirmer.cpp:20-30 */
static void squirmer_init(orc_squirmer* s, const orc_params* p, double nu, double xc, double yc, double zc) {
    s->R = p->radius; s->B1 = p->b1; s->beta = p->beta; s->planar = p->planar;
    s->plane_xz = p->plane_xz;
    s->M = p->rho_p * 4.0 / 3.0 * PI_ * s->R * s->R * s->R;
    s->I = 0.4 * s->M * s->R * s->R;
    s->X[0] = xc; s->X[1] = yc; s->X[2] = zc;
    const double a = p->angle * PI_ / 180.0;
    if (p->plane_xz) { s->e[0] = cos(a); s->e[1] = 0.0; s->e[2] = sin(a); }
    else { s->e[0] = cos(a); s->e[1] = sin(a); s->e[2] = 0.0; }
    s->eps = p->eps; s->hc = p->hc; s->hmin = p->hmin;
    s->FS = 6.0 * PI_ * nu * s->R * (2.0 * s->B1 / 3.0);
}

/* lbm.cpp:100-104 */
void orc_offset_from_squirmer(const orc_lbm* L, double px, double py, double pz, double r[3]) {
    r[0] = px - L->sq.X[0]; r[0] -= L->NX * round(r[0] / L->NX);
    r[1] = py - L->sq.X[1]; if (!L->wallsY) r[1] -= L->NY * round(r[1] / L->NY);
    r[2] = pz - L->sq.X[2]; if (!L->wallsZ) r[2] -= L->NZ * round(r[2] / L->NZ);
}

static void update_particle_flags(orc_lbm* L, int initial);

/* lbm.cpp:60-94 */
orc_lbm* orc_create(const orc_params* p) {
    orc_lbm* L = (orc_lbm*)calloc(1, sizeof(orc_lbm));
    if (!L) return NULL;
    L->NX = p->nx; L->NY = p->ny; L->NZ = p->nz;
    L->N = (size_t)p->nx * p->ny * p->nz;
    L->tau = p->tau;
    L->F[0] = p->fx; L->F[1] = p->fy; L->F[2] = p->fz;
    L->wallsY = p->walls_y; L->wallsZ = p->walls_z;
    const size_t N = L->N;
    L->f = (double*)malloc(sizeof(double) * Q * N);
    L->fnew = (double*)malloc(sizeof(double) * Q * N);
    L->rho = (double*)malloc(sizeof(double) * N);
    L->ux = (double*)calloc(N, sizeof(double));
    L->uy = (double*)calloc(N, sizeof(double));
    L->uz = (double*)calloc(N, sizeof(double));
    L->flag = (uint8_t*)malloc(N);
    if (!L->f || !L->fnew || !L->rho || !L->ux || !L->uy || !L->uz || !L->flag) {
        orc_destroy(L);
        return NULL;
    }
    for (size_t n = 0; n < N; ++n) L->rho[n] = 1.0;

    L->omPlus = 1.0 / L->tau;
    L->omMinus = 1.0 / (0.5 + p->lambda / (L->tau - 0.5));
    L->nu = (L->tau - 0.5) / 3.0;

    for (int z = 0; z < L->NZ; ++z)
        for (int y = 0; y < L->NY; ++y)
            for (int x = 0; x < L->NX; ++x) {
                const int wall = (L->wallsY && (y == 0 || y == L->NY - 1)) ||
                                 (L->wallsZ && (z == 0 || z == L->NZ - 1));
                L->flag[idx(L, x, y, z)] = wall ? WALL : FLUID;
            }
    for (size_t n = 0; n < N; ++n)
        for (int i = 0; i < Q; ++i) L->f[(size_t)i * N + n] = w[i];
    memcpy(L->fnew, L->f, sizeof(double) * Q * N);

    if (p->squirmer) {
        L->hasSquirmer = 1;
        squirmer_init(&L->sq, p, L->nu,
                      p->x0 >= 0 ? p->x0 : 0.5 * (L->NX - 1),
                      p->y0 >= 0 ? p->y0 : 0.5 * (L->NY - 1),
                      p->z0 >= 0 ? p->z0 : 0.5 * (L->NZ - 1));
        update_particle_flags(L, 1);
    } else {
        /* Squirmer defaults (squirmer.hpp:16-37) for an unused member. */
        L->sq.R = 4.0; L->sq.B1 = 0.015; L->sq.planar = 1;
        L->sq.e[0] = 1.0; L->sq.eps = 0.2; L->sq.hc = 2.0; L->sq.hmin = 0.5;
    }
    return L;
}

void orc_destroy(orc_lbm* L) {
    if (!L) return;
    free(L->f); free(L->fnew); free(L->rho);
    free(L->ux); free(L->uy); free(L->uz); free(L->flag);
    free(L);
}

/* lbm.cpp:128-224: moments, TRT + Guo collision, push streaming, halfway
 * and moving-boundary bounce-back, momentum-exchange force and torque. */
void orc_step(orc_lbm* L) {
    const int NX = L->NX, NY = L->NY, NZ = L->NZ;
    const size_t N = L->N;
    const double omPlus = L->omPlus, omMinus = L->omMinus;
    const double gP = 1.0 - 0.5 * omPlus, gM = 1.0 - 0.5 * omMinus;
    const double Fx = L->F[0], Fy = L->F[1], Fz = L->F[2];
    const double* f = L->f;
    double* fnew = L->fnew;
    const uint8_t* flag = L->flag;
    double FWx = 0, FWy = 0, FWz = 0;
    double FPx = 0, FPy = 0, FPz = 0, TPx = 0, TPy = 0, TPz = 0;

#pragma omp parallel for collapse(2) schedule(static) \
    reduction(+ : FWx, FWy, FWz, FPx, FPy, FPz, TPx, TPy, TPz)
    for (int z = 0; z < NZ; ++z)
        for (int y = 0; y < NY; ++y)
            for (int x = 0; x < NX; ++x) {
                const size_t n = idx(L, x, y, z);
                if (flag[n] != FLUID) continue;

                double fl[Q];
                double r = 0, mx = 0, my = 0, mz = 0;
                for (int i = 0; i < Q; ++i) {
                    fl[i] = f[(size_t)i * N + n];
                    r += fl[i]; mx += cx[i] * fl[i]; my += cy[i] * fl[i]; mz += cz[i] * fl[i];
                }
                const double u = (mx + 0.5 * Fx) / r, v = (my + 0.5 * Fy) / r, wz = (mz + 0.5 * Fz) / r;
                L->rho[n] = r; L->ux[n] = u; L->uy[n] = v; L->uz[n] = wz;
                const double usq = u * u + v * v + wz * wz, uF = u * Fx + v * Fy + wz * Fz;

                for (int i = 0; i < Q; ++i) {
                    const int io = opp[i];
                    const double cu = cx[i] * u + cy[i] * v + cz[i] * wz;
                    const double cF = cx[i] * Fx + cy[i] * Fy + cz[i] * Fz;
                    const double feqP = w[i] * r * (1.0 + 4.5 * cu * cu - 1.5 * usq);
                    const double feqM = w[i] * r * 3.0 * cu;
                    const double fP = 0.5 * (fl[i] + fl[io]), fM = 0.5 * (fl[i] - fl[io]);
                    const double SP = w[i] * (9.0 * cu * cF - 3.0 * uF), SM = w[i] * 3.0 * cF;
                    const double fpost = fl[i] - omPlus * (fP - feqP) - omMinus * (fM - feqM)
                                         + gP * SP + gM * SM;

                    int xn = x + cx[i], yn = y + cy[i], zn = z + cz[i];
                    if (xn < 0) xn += NX; else if (xn >= NX) xn -= NX;
                    if (yn < 0) yn += NY; else if (yn >= NY) yn -= NY;
                    if (zn < 0) zn += NZ; else if (zn >= NZ) zn -= NZ;
                    const size_t nb = idx(L, xn, yn, zn);
                    const uint8_t t = flag[nb];
                    if (t == FLUID) { fnew[(size_t)i * N + nb] = fpost; continue; }

                    double ub[3] = {0, 0, 0}, rb[3] = {0, 0, 0};
                    if (t == PARTICLE) {
                        orc_offset_from_squirmer(L, x + 0.5 * cx[i], y + 0.5 * cy[i], z + 0.5 * cz[i], rb);
                        orc_surface_velocity(&L->sq, rb, ub, 1);
                    }
                    const double cub = cx[i] * ub[0] + cy[i] * ub[1] + cz[i] * ub[2];
                    const double fb = fpost - 6.0 * w[i] * r * cub;
                    fnew[(size_t)io * N + n] = fb;

                    const double s = fpost + fb, dl = fpost - fb;
                    const double dpx = s * cx[i] - dl * ub[0];
                    const double dpy = s * cy[i] - dl * ub[1];
                    const double dpz = s * cz[i] - dl * ub[2];
                    if (t == WALL) { FWx += dpx; FWy += dpy; FWz += dpz; }
                    else {
                        FPx += dpx; FPy += dpy; FPz += dpz;
                        TPx += rb[1] * dpz - rb[2] * dpy;
                        TPy += rb[2] * dpx - rb[0] * dpz;
                        TPz += rb[0] * dpy - rb[1] * dpx;
                    }
                }
            }
    double* tmp = L->f; L->f = L->fnew; L->fnew = tmp;
    L->wallForce[0] = FWx; L->wallForce[1] = FWy; L->wallForce[2] = FWz;
    if (L->hasSquirmer) {
        L->sq.Fh[0] = FPx; L->sq.Fh[1] = FPy; L->sq.Fh[2] = FPz;
        L->sq.Th[0] = TPx; L->sq.Th[1] = TPy; L->sq.Th[2] = TPz;
    }
}

/* lbm.cpp:231-241 */
static void apply_wall_potential(orc_lbm* L) {
    orc_squirmer* s = &L->sq;
    s->Fw[0] = s->Fw[1] = s->Fw[2] = 0.0;
    if (L->wallsY) {
        s->Fw[1] += orc_soft_wall_force(s->X[1] - s->R - 0.5, s->eps, s->FS, s->hc);
        s->Fw[1] -= orc_soft_wall_force((L->NY - 1.5) - s->X[1] - s->R, s->eps, s->FS, s->hc);
    }
    if (L->wallsZ) {
        s->Fw[2] += orc_soft_wall_force(s->X[2] - s->R - 0.5, s->eps, s->FS, s->hc);
        s->Fw[2] -= orc_soft_wall_force((L->NZ - 1.5) - s->X[2] - s->R, s->eps, s->FS, s->hc);
    }
}

/* lbm.cpp:245-253 */
static void clamp_axis(orc_squirmer* s, int d, int Nd) {
    const double lo = 0.5 + s->R + s->hmin, hi = (Nd - 1.5) - s->R - s->hmin;
    if (s->X[d] < lo) { s->X[d] = lo; if (s->U[d] < 0) s->U[d] = 0; ++s->contacts; }
    if (s->X[d] > hi) { s->X[d] = hi; if (s->U[d] > 0) s->U[d] = 0; ++s->contacts; }
}
static void enforce_minimum_gap(orc_lbm* L) {
    if (L->wallsY) clamp_axis(&L->sq, 1, L->NY);
    if (L->wallsZ) clamp_axis(&L->sq, 2, L->NZ);
}

/* lbm.cpp:271-320: sequential, in-place remap of the sphere. */
static void update_particle_flags(orc_lbm* L, int initial) {
    orc_squirmer* s = &L->sq;
    const int NX = L->NX, NY = L->NY, NZ = L->NZ;
    const size_t N = L->N;
    const double R = s->R;
    const int x0 = (int)floor(s->X[0] - R) - 2, x1 = (int)ceil(s->X[0] + R) + 2;
    int y0 = (int)floor(s->X[1] - R) - 2, y1 = (int)ceil(s->X[1] + R) + 2;
    const int z0 = (int)floor(s->X[2] - R) - 2, z1 = (int)ceil(s->X[2] + R) + 2;
    if (L->wallsY) { if (y0 < 0) y0 = 0; if (y1 > NY - 1) y1 = NY - 1; }

    for (int zz = z0; zz <= z1; ++zz)
        for (int yy = y0; yy <= y1; ++yy)
            for (int xx = x0; xx <= x1; ++xx) {
                const int x = ((xx % NX) + NX) % NX;
                const int y = ((yy % NY) + NY) % NY;
                const int z = ((zz % NZ) + NZ) % NZ;
                const size_t n = idx(L, x, y, z);
                if (L->flag[n] == WALL) continue;
                double r[3];
                orc_offset_from_squirmer(L, x, y, z, r);
                const int inside = r[0] * r[0] + r[1] * r[1] + r[2] * r[2] < R * R;
                const int was = L->flag[n] == PARTICLE;
                if (inside == was) continue;
                if (initial) { L->flag[n] = inside ? PARTICLE : FLUID; continue; }

                if (inside) {
                    double p[3] = {0, 0, 0}, t[3];
                    for (int i = 0; i < Q; ++i) {
                        const double fv = L->f[(size_t)i * N + n];
                        p[0] += fv * cx[i]; p[1] += fv * cy[i]; p[2] += fv * cz[i];
                    }
                    cross3(r, p, t);
                    for (int d = 0; d < 3; ++d) { s->Fc[d] += p[d]; s->Tc[d] += t[d]; }
                    L->flag[n] = PARTICLE;
                } else {
                    double rs = 0; int cnt = 0;
                    for (int i = 1; i < Q; ++i) {
                        const int xn = (x + cx[i] + NX) % NX, yn = (y + cy[i] + NY) % NY,
                                  zn = (z + cz[i] + NZ) % NZ;
                        const size_t nb = idx(L, xn, yn, zn);
                        if (L->flag[nb] == FLUID) { rs += L->rho[nb]; ++cnt; }
                    }
                    const double rr = cnt ? rs / cnt : 1.0;
                    double ub[3];
                    orc_surface_velocity(s, r, ub, 0);
                    for (int i = 0; i < Q; ++i)
                        L->f[(size_t)i * N + n] = orc_equilibrium(i, rr, ub[0], ub[1], ub[2]);
                    L->rho[n] = rr; L->ux[n] = ub[0]; L->uy[n] = ub[1]; L->uz[n] = ub[2];
                    double p[3] = {rr * ub[0], rr * ub[1], rr * ub[2]}, t[3];
                    cross3(r, p, t);
                    for (int d = 0; d < 3; ++d) { s->Fc[d] -= p[d]; s->Tc[d] -= t[d]; }
                    L->flag[n] = FLUID;
                }
            }
}

/* lbm.cpp:257-263 */
void orc_advance_squirmer(orc_lbm* L) {
    if (!L->hasSquirmer) return;
    apply_wall_potential(L);
    orc_integrate(&L->sq);
    enforce_minimum_gap(L);
    update_particle_flags(L, 0);
}

/* Summation of the mass correction: 0 = the reference's (lbm.cpp:325-331,
 * m += f into a double, per OpenMP thread); 1 = exact.  The reference's
 * partial sums grow to ~N / threads, whose ulp exceeds the per-node
 * deviations from rho = 1, and the far-field populations are all w_i, so
 * every addition drops the same low bits: at 16.7M nodes the shift d is off
 * by ~1e-10 per correction, by an amount that depends on the thread count.
 * The exact mode sums each node's excess (sum_i f_i) - 1 as a 128-bit
 * fixed-point integer (trunc(e * 2^64), as fixed_point.hpp of the product
 * does) and sets d = -excess / n_fluid: the same algorithm without that
 * rounding. */
static int g_mass_sum_exact = 0;
void orc_set_mass_sum_exact(int exact) { g_mass_sum_exact = exact; }

static __int128 fixed_of(double v) {
    if (v == 0.0 || !isfinite(v)) return 0;
    int ex;
    const double m = frexp(fabs(v), &ex); /* |v| = m 2^ex, m in [0.5, 1) */
    const uint64_t mant = (uint64_t)ldexp(m, 53); /* exact: 53-bit significand */
    const int sh = ex - 53 + 64;                  /* |v| 2^64 = mant 2^sh */
    unsigned __int128 a;
    if (sh >= 0) a = sh > 74 ? (((unsigned __int128)1 << 127) - 1) : ((unsigned __int128)mant << sh);
    else a = sh <= -64 ? 0 : (unsigned __int128)(mant >> (-sh));
    return v < 0 ? -(__int128)a : (__int128)a;
}

/* lbm.cpp:324-339 */
void orc_correct_mass(orc_lbm* L) {
    const size_t N = L->N;
    double m = 0; long nf = 0;
    double d;
    if (g_mass_sum_exact) {
        __int128 ex = 0;
        for (long n = 0; n < (long)N; ++n) {
            if (L->flag[n] != FLUID) continue;
            double s = 0.0;
            for (int i = 0; i < Q; ++i) s += L->f[(size_t)i * N + n];
            ex += fixed_of(s - 1.0);
            ++nf;
        }
        if (nf == 0) return;
        const double hi = (double)(int64_t)(ex >> 64), lo = (double)(uint64_t)ex;
        d = -(hi + lo * 0x1p-64) / (double)nf;
    } else {
#pragma omp parallel for reduction(+ : m, nf)
        for (long n = 0; n < (long)N; ++n) {
            if (L->flag[n] != FLUID) continue;
            for (int i = 0; i < Q; ++i) m += L->f[(size_t)i * N + n];
            ++nf;
        }
        if (nf == 0) return;
        d = 1.0 - m / nf;
    }
#pragma omp parallel for
    for (long n = 0; n < (long)N; ++n) {
        if (L->flag[n] != FLUID) continue;
        for (int i = 0; i < Q; ++i) L->f[(size_t)i * N + n] += w[i] * d;
    }
}

/* lbm.cpp:342-364 */
long orc_count_fluid(const orc_lbm* L) {
    long c = 0;
    for (size_t n = 0; n < L->N; ++n) c += L->flag[n] == FLUID;
    return c;
}
double orc_mean_density(const orc_lbm* L) {
    double m = 0; long nf = 0;
    for (size_t n = 0; n < L->N; ++n) if (L->flag[n] == FLUID) { m += L->rho[n]; ++nf; }
    return nf ? m / nf : 0.0;
}
double orc_max_speed(const orc_lbm* L) {
    double m = 0;
    for (size_t n = 0; n < L->N; ++n)
        if (L->flag[n] == FLUID) {
            const double v = sqrt(L->ux[n] * L->ux[n] + L->uy[n] * L->uy[n] + L->uz[n] * L->uz[n]);
            m = (m < v) ? v : m; /* std::max(m, v) */
        }
    return m;
}
double orc_sum_ux(const orc_lbm* L) {
    double s = 0;
    for (size_t n = 0; n < L->N; ++n) if (L->flag[n] == FLUID) s += L->ux[n];
    return s;
}

/* main.cpp:60-64: the step / advanceSquirmer / correctMass cadence. */
void orc_run(orc_lbm* L, int steps, int mass_every) {
    for (int t = 1; t <= steps; ++t) {
        orc_step(L);
        if (L->hasSquirmer) {
            orc_advance_squirmer(L);
            if (mass_every > 0 && t % mass_every == 0) orc_correct_mass(L);
        }
    }
}
