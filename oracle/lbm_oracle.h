/*
 * lbm_oracle.h -- CPU FP64 restatement of the reference D3Q19 TRT step and
 * its squirmer coupling.  TEST INFRASTRUCTURE ONLY: this library is the
 * checker that the HIP path is compared against (tests, smoke check and the
 * bench's CPU-baseline leg).  It is never linked into, loaded by or called
 * from the product library.
 *
 * It follows, function by function and in the same floating-point operation
 * order, the reference files
 *   src/utils/lattice.hpp  (velocity set, weights, node types, equilibrium)
 *   src/utils/lbm.hpp/.cpp (layout, step, bounce-back, remap, mass correction)
 *   src/utils/squirmer.hpp/.cpp (slip velocity, Newton-Euler, wall potential)
 *   src/utils/params.hpp   (parameter defaults)
 * with the push scheme and two population arrays the reference uses.
 *
 * One extension beyond the reference: `plane_xz` initialises the orientation
 * in the x-z plane and confines planar motion to x-z (the reference only has
 * the x-y plane, params.hpp:48-50, squirmer.cpp:26-27).  With plane_xz = 0
 * every function is the reference's.
 */
#ifndef LBM_ORACLE_H
#define LBM_ORACLE_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* struct Params (reference params.hpp:14-56), the fields the LBM ctor reads. */
typedef struct orc_params {
    int nx, ny, nz;
    double tau, lambda, fx, fy, fz;
    int walls_y, walls_z;
    int squirmer;
    double radius, b1, beta, rho_p, x0, y0, z0, angle;
    int planar;
    double eps, hc, hmin;
    int plane_xz; /* extension: orientation/planar plane x-z instead of x-y */
} orc_params;

/* struct Squirmer (reference squirmer.hpp:14-37). */
typedef struct orc_squirmer {
    double R, B1, beta, M, I;
    int planar, plane_xz;
    double X[3], e[3], U[3], W[3];
    double Fh[3], Th[3], Fc[3], Tc[3], Fw[3];
    double eps, hc, hmin, FS;
    long contacts;
} orc_squirmer;

/* class LBM (reference lbm.hpp:19-78): public data. */
typedef struct orc_lbm {
    int NX, NY, NZ;
    size_t N;
    double tau, omPlus, omMinus, nu;
    double F[3];
    int wallsY, wallsZ;
    double *f, *fnew;             /* f[i*N + n], n = x + NX*(y + NY*z) */
    double *rho, *ux, *uy, *uz;
    uint8_t *flag;
    double wallForce[3];
    int hasSquirmer;
    orc_squirmer sq;
} orc_lbm;

void orc_default_params(orc_params* p);
orc_lbm* orc_create(const orc_params* p);
void orc_destroy(orc_lbm* L);

void orc_step(orc_lbm* L);
void orc_advance_squirmer(orc_lbm* L);
void orc_correct_mass(orc_lbm* L);
/* 0 (default): the mass correction sums as the reference does; 1: exactly
 * (see lbm_oracle.c).  Process-wide. */
void orc_set_mass_sum_exact(int exact);

long orc_count_fluid(const orc_lbm* L);
double orc_mean_density(const orc_lbm* L);
double orc_max_speed(const orc_lbm* L);
double orc_sum_ux(const orc_lbm* L);

/* Squirmer helpers, exposed so their known answers can be checked alone. */
void orc_surface_velocity(const orc_squirmer* s, const double r[3], double ub[3], int with_slip);
void orc_integrate(orc_squirmer* s);
double orc_soft_wall_force(double h, double eps, double FS, double hc);
double orc_equilibrium(int i, double rho, double ux, double uy, double uz);
void orc_offset_from_squirmer(const orc_lbm* L, double px, double py, double pz, double r[3]);

/* Run `steps` iterations of the reference driver loop (main.cpp:60-66):
 * step, then with a squirmer advanceSquirmer and correctMass every
 * `mass_every` steps (0 = never).  Used for timing and long runs. */
void orc_run(orc_lbm* L, int steps, int mass_every);

/* Access to the velocity set (lattice.hpp:20-27). */
int orc_cx(int i);
int orc_cy(int i);
int orc_cz(int i);
int orc_opp(int i);
double orc_w(int i);

int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
