/*
 * Harness of the oracle (oracle/lbm_oracle.c) for the CPU tests' sanitizer
 * build (AddressSanitizer + UndefinedBehaviorSanitizer): runs small cases
 * through every function of the oracle -- duct with walls in y and z and a
 * body force, a channel squirmer touching a wall (wall potential, hard gap,
 * remaps, mass correction), a squirmer in the x-z plane, and a fully
 * periodic box with the sphere over the corner (every box wraps) -- and
 * prints the mass and squirmer state so the test can check them.
 */
#include <math.h>
#include <stdio.h>

#include "lbm_oracle.h"

static double total_mass(const orc_lbm* L) {
    double m = 0.0;
    for (size_t n = 0; n < L->N; ++n)
        if (L->flag[n] == 0)
            for (int i = 0; i < 19; ++i) m += L->f[(size_t)i * L->N + n];
    return m;
}

static int run_case(const char* name, orc_params* p, int steps, int mass_every) {
    orc_lbm* L = orc_create(p);
    if (!L) return 1;
    orc_run(L, steps, mass_every);
    const double m = total_mass(L), nf = (double)orc_count_fluid(L);
    double r[3];
    orc_offset_from_squirmer(L, 0.25, 1.5, 2.75, r);
    printf("%s fluid %.0f mass_excess %.3e mean_rho %.15f umax %.6e X %.9f %.9f %.9f contacts %ld r %.3f\n", name, nf,
           m - nf, orc_mean_density(L), orc_max_speed(L), L->sq.X[0], L->sq.X[1], L->sq.X[2], L->sq.contacts, r[0]);
    const int bad = !isfinite(m) || (L->hasSquirmer && mass_every > 0 && fabs(m - nf) > 1e-9 * nf);
    orc_destroy(L);
    return bad;
}

int main(void) {
    int bad = 0;
    orc_params p;

    orc_default_params(&p);
    p.nx = 4; p.ny = 10; p.nz = 12; p.tau = 0.8; p.fx = 5e-5; p.fy = 1e-6; p.walls_y = 1; p.walls_z = 1;
    bad |= run_case("duct", &p, 60, 0);

    orc_default_params(&p);
    p.nx = 40; p.ny = 22; p.nz = 16; p.walls_y = 1; p.walls_z = 0; p.squirmer = 1; p.radius = 4.0; p.b1 = 0.03;
    p.beta = -1.0; p.angle = -40.0; p.y0 = 6.2; p.hmin = 0.5;
    bad |= run_case("wall_contact", &p, 150, 10);

    orc_default_params(&p);
    p.nx = 36; p.ny = 20; p.nz = 22; p.walls_y = 0; p.walls_z = 1; p.squirmer = 1; p.radius = 4.5; p.b1 = 0.02;
    p.beta = -3.0; p.angle = -25.0; p.plane_xz = 1; p.z0 = 9.0;
    bad |= run_case("plane_xz", &p, 80, 10);

    orc_default_params(&p);
    p.nx = 30; p.ny = 26; p.nz = 28; p.walls_y = 0; p.walls_z = 0; p.squirmer = 1; p.radius = 4.0; p.b1 = 0.02;
    p.beta = 2.0; p.angle = 30.0; p.planar = 0; p.x0 = 0.4; p.y0 = 25.7; p.z0 = 0.6;
    bad |= run_case("corner", &p, 60, 10);

    /* the helpers on their own */
    orc_squirmer s = {0};
    s.R = 4.0; s.B1 = 0.015; s.beta = -3.0; s.M = 268.08; s.I = 0.4 * 268.08 * 16.0; s.planar = 1;
    s.e[0] = 1.0;
    const double rr[3] = {3.0, 2.0, -1.0};
    double ub[3];
    orc_surface_velocity(&s, rr, ub, 1);
    s.Fh[0] = 1e-3;
    orc_integrate(&s);
    printf("helpers feq %.17g wall %.17g ub %.6e %.6e %.6e X %.9f\n", orc_equilibrium(7, 1.01, 0.02, -0.01, 0.005),
           orc_soft_wall_force(0.5, 0.2, 0.01, 2.0), ub[0], ub[1], ub[2], s.X[0]);
    return bad;
}
