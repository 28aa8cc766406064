/*
 * lbm_hip.h -- C ABI of liblbm_hip.so, the MI355X (gfx950) implementation of
 * the D3Q19 TRT lattice-Boltzmann step and its squirmer coupling.
 *
 * Plain C: pointers, sizes and PODs only.  Every call returns 0 on success
 * and a negative LBM_HIP_E* code on failure; lbm_hip_last_error() then holds
 * a message (thread-local).  One context owns one device allocation set and
 * one HIP stream on one GPU; calls enqueue work on that stream and return
 * without waiting, except the getters and transfers, which synchronise.
 * A context is not thread-safe (like the reference's class LBM).
 *
 * Each entry point names the reference interface it replaces
 * (byjesh1998/LBM_solver_3D, file:line).  The C++ facade in
 * lbm_solver_3d_amd/host/ re-exposes them under the reference's own class
 * and method names (lbm::LBM, lbm::Squirmer).
 */
#ifndef LBM_HIP_H
#define LBM_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ---- */
#define LBM_HIP_OK 0
#define LBM_HIP_EINVAL (-1)    /* invalid argument / parameters (validate() rules) */
#define LBM_HIP_ENOMEM (-2)    /* device or host allocation failed */
#define LBM_HIP_EDEVICE (-3)   /* HIP runtime error, or no usable GPU */
#define LBM_HIP_ESTATE (-4)    /* requested data was not produced by the last step */
#define LBM_HIP_EOVERFLOW (-5) /* remap changed more nodes than the device list holds */

/* ---- population storage precision ---- */
#define LBM_HIP_FP64 0  /* f_i as double, reference arithmetic order (parity mode) */
#define LBM_HIP_FP32 1  /* f_i - w_i as float, FP32 collision arithmetic */
#define LBM_HIP_FP16 2  /* (f_i - w_i) * 2^15 as IEEE half, FP32 collision arithmetic */
#define LBM_HIP_BF16 3  /* f_i - w_i as bfloat16, FP32 collision arithmetic */

/* ---- lbm_hip_step flags ---- */
#define LBM_STEP_MACRO 0x1      /* write rho, ux, uy, uz of every fluid node (lbm.cpp:152) */
#define LBM_STEP_WALL_FORCE 0x2 /* accumulate the wall momentum exchange (lbm.cpp:209, 219) */
#define LBM_STEP_MASS 0x4       /* emit the fluid-mass partial sums used by correct_mass */
#define LBM_STEP_ALL (LBM_STEP_MACRO | LBM_STEP_WALL_FORCE | LBM_STEP_MASS)

/*
 * Parameters: the fields of the reference struct Params that its LBM
 * constructor reads (src/utils/params.hpp:14-56), with the same meaning and
 * defaults (lbm_hip_default_params), plus:
 *   plane_xz  -- orientation angle and planar confinement in the x-z plane
 *                instead of x-y (the reference hard-codes x-y,
 *                squirmer.cpp:26-27, 58, 65); 0 = reference behaviour.
 *   precision -- LBM_HIP_FP64 / FP32 / FP16 / BF16 population storage.
 *   device    -- HIP device ordinal.
 */
typedef struct lbm_hip_params {
    int nx, ny, nz;
    double tau, lambda, fx, fy, fz;
    int walls_y, walls_z;
    int squirmer;
    double radius, b1, beta, rho_p, x0, y0, z0, angle;
    int planar;
    double eps, hc, hmin;
    int plane_xz;
    int precision;
    int device;
} lbm_hip_params;

/* struct Squirmer (src/utils/squirmer.hpp:14-37): model, state, forces. */
typedef struct lbm_squirmer_state {
    double R, B1, beta, M, I, FS;
    double X[3], e[3], U[3], W[3];
    double Fh[3], Th[3], Fc[3], Tc[3], Fw[3];
    double eps, hc, hmin;
    int64_t contacts;
    int32_t planar, plane_xz;
} lbm_squirmer_state;

/* LBM diagnostics (src/utils/lbm.cpp:342-364), plus a stability check the
 * reference does not make: unstable_nodes counts the FLUID nodes whose
 * density or velocity is not finite, whose density lies outside
 * (0, LBM_HIP_RHO_MAX) or whose speed reaches the lattice sound speed
 * 1/sqrt(3) (a run past those bounds has lost its accuracy or diverged). */
#define LBM_HIP_RHO_MAX 2.0
typedef struct lbm_hip_diag {
    int64_t count_fluid;
    double mean_density, max_speed, sum_ux;
    int64_t unstable_nodes;
} lbm_hip_diag;

/* Sizes and derived constants of a context. */
typedef struct lbm_hip_info {
    int nx, ny, nz, precision;
    uint64_t n;             /* nodes */
    int64_t steps_done;     /* time steps taken so far */
    double tau, om_plus, om_minus, nu; /* lbm.cpp:67-69 */
    double bytes_per_lup;   /* algorithmic HBM bytes per lattice update: 2*19*sizeof(storage) */
    uint64_t device_bytes;  /* device memory held by the context */
    int has_squirmer, walls_y, walls_z;
    int64_t count_fluid;    /* FLUID nodes now (LBM::countFluid, lbm.cpp:342) */
    int row_pitch;          /* nodes per stored population row (nx, or nx rounded up to 64) */
    int tile_width;         /* collide-stream tile width in x (64 / 128 / 256; 0: row runs) */
    int xcd_remap;          /* tile order over the XCDs (lbm_hip_set_tiling) */
    int occupancy_even;     /* cap on collide-stream blocks per CU, even / odd steps (lbm_hip_set_occupancy) */
    int occupancy_odd;
    int resident_even;      /* blocks per CU the tile launches allow (HIP's occupancy calculator) */
    int resident_odd;
    int odd_row_waves;      /* odd steps in row blocks: wavefronts per row (0: odd steps in tiles; the
                               occupancy_odd cap then applies to the row blocks' wavefronts) */
} lbm_hip_info;

typedef struct lbm_hip_ctx lbm_hip_ctx;

/* Reference defaults of struct Params (params.hpp:16-56); precision FP64, device 0. */
void lbm_hip_default_params(lbm_hip_params* p);

/* Params::validate (params.cpp:120-140): 0 if valid, else LBM_HIP_EINVAL with
 * the reference's message in lbm_hip_last_error().  Beyond the reference's
 * rules: nx * ny * nz < 2^32 - 65536 (32-bit node indices; population
 * offsets are 64-bit), and a known precision. */
int lbm_hip_validate(const lbm_hip_params* p);

/* LBM::LBM(const Params&) (lbm.cpp:60-94): allocate, set wall flags, f = w_i,
 * initialise the squirmer and its initial flags.  Validates first. */
int lbm_hip_create(const lbm_hip_params* p, lbm_hip_ctx** out);

/* ~LBM: release every device resource of the context. */
void lbm_hip_destroy(lbm_hip_ctx* ctx);

/* LBM::step (lbm.cpp:128-224): one fused collide-stream update with TRT + Guo
 * forcing, halfway bounce-back on WALL links and moving-boundary bounce-back
 * on PARTICLE links, and the momentum-exchange force and torque on the
 * squirmer.  `flags` (LBM_STEP_*) selects which by-products are written:
 * the reference writes all of them every step (LBM_STEP_ALL reproduces it).
 * With a squirmer, the PARTICLE-link pass (bounce-back correction, force,
 * torque) is left pending and runs in the launch of the next
 * lbm_hip_advance_squirmer, or first in any call that reads its results
 * (lbm_hip_step, correct_mass, get/set_squirmer, download, upload,
 * synchronize, particle_links, autotune): the results are the same. */
int lbm_hip_step(lbm_hip_ctx* ctx, int flags);

/* LBM::advanceSquirmer (lbm.cpp:257-263): wall potential, Newton-Euler
 * update, hard gap and the sequential remap of the sphere, on the device in
 * FP64, in one launch with the pending link pass of the last step.  No-op
 * without a squirmer. */
int lbm_hip_advance_squirmer(lbm_hip_ctx* ctx);

/* LBM::correctMass (lbm.cpp:324-339): restore the mean fluid density to 1.
 * The shift f_i += w_i d is applied as the next step loads f. */
int lbm_hip_correct_mass(lbm_hip_ctx* ctx);

/* The driver loop of main.cpp:60-64 for `steps` steps: step, then with a
 * squirmer advanceSquirmer and correctMass every `mass_every` steps
 * (0 = never).  `last_flags` is passed to the final step's lbm_hip_step; the
 * other steps write only what the coupling needs.  Captured once into a HIP
 * graph per context and replayed. */
int lbm_hip_run(lbm_hip_ctx* ctx, int steps, int mass_every, int last_flags);

/* Read / write the public struct Squirmer member LBM::sq (lbm.hpp:78). */
int lbm_hip_get_squirmer(lbm_hip_ctx* ctx, lbm_squirmer_state* out);
/* The radius may not exceed the one the context was created with, and the
 * centre may move by less than one node per axis (the flags stay those of
 * the old centre until the next remap, as in the reference's loop). */
int lbm_hip_set_squirmer(lbm_hip_ctx* ctx, const lbm_squirmer_state* in);

/* LBM::wallForce (lbm.hpp:75) after the last step; LBM_HIP_ESTATE if that
 * step ran without LBM_STEP_WALL_FORCE. */
int lbm_hip_get_wall_force(lbm_hip_ctx* ctx, double out[3]);

/* countFluid / meanDensity / maxSpeed / sumUx (lbm.cpp:342-364) and the
 * count of unstable nodes (lbm_hip_diag), one pass of a device reduction.
 * They read the moments of the last step: LBM_HIP_ESTATE if it ran without
 * LBM_STEP_MACRO. */
int lbm_hip_diagnostics(lbm_hip_ctx* ctx, lbm_hip_diag* out);

/* Copy the public arrays of class LBM (lbm.hpp:71-73) to host buffers in the
 * reference layout: f[i*N + n] with n = x + NX*(y + NY*z) (19*N doubles),
 * rho/ux/uy/uz (N doubles each), flag (N bytes, NodeType values).  Any pointer
 * may be NULL.  Populations of non-fluid nodes are not defined (the reference
 * never reads them).  rho/u: LBM_HIP_ESTATE unless the last step wrote them. */
int lbm_hip_download(lbm_hip_ctx* ctx, double* f, double* rho, double* ux, double* uy, double* uz,
                     uint8_t* flag);

/* Overwrite the populations of every node from f[i*N + n] (reference layout). */
int lbm_hip_upload_populations(lbm_hip_ctx* ctx, const double* f);

/* Wait for all work of the context. */
int lbm_hip_synchronize(lbm_hip_ctx* ctx);

int lbm_hip_get_info(lbm_hip_ctx* ctx, lbm_hip_info* out);

/* The PARTICLE links of the current flags, as the device's link list holds
 * them: (node n, direction i) with n FLUID and n + c_i PARTICLE -- the links
 * LBM::step visits in its sweep (lbm.cpp:179-199), which the reference finds
 * from the flags every step.  Writes at most `cap` of them, in list order,
 * into node[] / dir[] (either may be NULL when cap is 0) and their number
 * into *count.  No counterpart in the reference: it exposes the link list
 * for inspection and tests. */
int lbm_hip_particle_links(lbm_hip_ctx* ctx, uint32_t* node, uint8_t* dir, int64_t cap, int64_t* count);

/* Bench support: time `steps` driver-loop steps (as lbm_hip_run, graph
 * replayed) with HIP events on the context's stream -> *total_ms; then run up
 * to 64 further steps eagerly with an event pair around the collide-stream
 * update of each step (the bulk kernel, and on even steps the boundary
 * kernel after it) and return their mean duration times `steps`
 * -> *step_kernel_ms. */
int lbm_hip_time_run(lbm_hip_ctx* ctx, int steps, int mass_every, double* total_ms,
                     double* step_kernel_ms);

/* ---- tile-shape autotuning ----
 * No counterpart in the reference (its OpenMP loop has no launch shape).
 * Candidates: collide-stream tiles of 64 x 4, 128 x 2 and 256 x 1 nodes
 * (tile_width 64 / 128 / 256), each in three orders of the tiles over the
 * 8 XCDs (xcd_remap): 0 the hardware's round robin of consecutive tiles,
 * 1 a contiguous range of tiles per XCD, 2 whole tile rows per XCD (the
 * tiles that share lines at their x edges in one L2); and row runs
 * (tile_width 0: four consecutive 64-node chunks of the rows in row-major
 * order, so rows narrower than a tile leave no lanes idle) in orders 1 and 0.
 * Each candidate runs `trial_steps` driver-loop steps (graph replayed, after
 * one untimed cycle) from the same saved state; the state is restored
 * afterwards, so tuning does not advance the simulation.  The fastest
 * candidate stays selected. */
#define LBM_HIP_TUNE_CANDIDATES 11
typedef struct lbm_hip_tuning {
    int tile_width;       /* selected: 64, 128 or 256 nodes along x, or 0 (row runs) */
    int xcd_remap;        /* selected tile order: 0, 1 or 2 (above) */
    double ms_per_step;   /* selected candidate's time per step */
    int cand_tile_width[LBM_HIP_TUNE_CANDIDATES];
    int cand_xcd_remap[LBM_HIP_TUNE_CANDIDATES];
    double cand_ms_per_step[LBM_HIP_TUNE_CANDIDATES];
} lbm_hip_tuning;

int lbm_hip_autotune(lbm_hip_ctx* ctx, int trial_steps, int mass_every, lbm_hip_tuning* out);

/* Select a tile shape directly (tile_width 64 / 128 / 256 with xcd_remap
 * 0 / 1 / 2, or tile_width 0 with xcd_remap 0 / 1).  A new context uses
 * tiles of 64, 128 or 256 nodes (the smallest at least NX, else 256) in
 * order 2. */
int lbm_hip_set_tiling(lbm_hip_ctx* ctx, int tile_width, int xcd_remap);

/* Cap on the collide-stream blocks (256 threads each) resident per compute
 * unit, for even and odd steps: 2..8, or 0 for no cap (the registers then
 * allow 8).  No counterpart in the reference.  The in-place update is bound
 * by HBM, and fewer wavefronts in flight can move its bytes faster; the cap
 * is set by reserving LDS per block, so results do not change.  Defaults,
 * measured on the 512 x 256 x 128 channel: FP64 3 / 3, FP32 4 / 4, FP16
 * 0 / 6, BF16 0 / 7. */
int lbm_hip_set_occupancy(lbm_hip_ctx* ctx, int even_blocks, int odd_blocks);

/* ---- ensemble of lattices ----
 * No counterpart in the reference (one run per process there); the members
 * are what `count` separate lbm_hip_create calls would make, and each
 * advances exactly as lbm_hip_run would advance it alone.  All members share
 * one precision and device.  Every phase of a step (collide-stream, even-step
 * boundary chunks, particle links, squirmer update + remap, mass-correction
 * finalize) is one launch for all members, and one cycle of the driver loop
 * is captured in a HIP graph and replayed. */
typedef struct lbm_hip_batch lbm_hip_batch;

typedef struct lbm_hip_batch_info {
    int count;                /* members */
    int precision;
    long long steps_done;
    long long total_nodes;
    int total_tiles;          /* blocks of one collide-stream launch */
    int boundary_blocks;      /* blocks of one even-step boundary launch */
    int link_blocks;          /* blocks of one particle-link launch */
    int tile_width, xcd_remap;
    size_t device_bytes;
    int odd_row_blocks;       /* blocks of one odd-step launch in row blocks (0: odd steps in tiles) */
    int odd_row_block_waves;  /* wavefronts per row block */
} lbm_hip_batch_info;

int lbm_hip_batch_create(const lbm_hip_params* members, int count, lbm_hip_batch** out);
void lbm_hip_batch_destroy(lbm_hip_batch* b);
/* `steps` driver-loop steps of every member (step, advance_squirmer,
 * correct_mass every `mass_every` for members with a squirmer); the last
 * step carries `last_flags` (LBM_STEP_* as in lbm_hip_step). */
int lbm_hip_batch_run(lbm_hip_batch* b, int steps, int mass_every, int last_flags);
int lbm_hip_batch_synchronize(lbm_hip_batch* b);
/* Member i as a context for the getters and transfers (get_squirmer,
 * download, diagnostics, ...).  Waits for the batch's work first.  The
 * context stays owned by the batch (lbm_hip_destroy ignores it); advancing it
 * on its own makes lbm_hip_batch_run fail. */
int lbm_hip_batch_member(lbm_hip_batch* b, int i, lbm_hip_ctx** out);
int lbm_hip_batch_info_get(lbm_hip_batch* b, lbm_hip_batch_info* out);
/* Autotuning as lbm_hip_autotune, one tile shape for all members. */
int lbm_hip_batch_autotune(lbm_hip_batch* b, int trial_steps, int mass_every, lbm_hip_tuning* out);
/* One tile shape for all members, as lbm_hip_set_tiling. */
int lbm_hip_batch_set_tiling(lbm_hip_batch* b, int tile_width, int xcd_remap);

/* One occupancy cap for all members, as lbm_hip_set_occupancy.  A new
 * batch uses 6 / 6 for FP16 and BF16 storage (measured on the 64-run
 * ensemble), the single-lattice defaults otherwise. */
int lbm_hip_batch_set_occupancy(lbm_hip_batch* b, int even_blocks, int odd_blocks);
/* Bench support: time `steps` graph-replayed steps with HIP events. */
int lbm_hip_batch_time_run(lbm_hip_batch* b, int steps, int mass_every, double* total_ms);

/* ---- bandwidth probe ----
 * No counterpart in the reference: a measured ceiling for the roofline of
 * the collide-stream kernel beside the published HBM peak.  Copies
 * bytes / 2 bytes between two device buffers of `device` (16-byte
 * nontemporal loads and stores) `reps` times after two untimed copies and
 * reports the bandwidth in GB/s counting bytes read plus written.
 * Allocates the two buffers and frees them before returning. */
int lbm_hip_copy_bandwidth(int device, size_t bytes, int reps, double* gbs);
/* The same copy spread over `parts` separate allocations (1..4096) of
 * bytes / parts each, all parts in one launch, as the batched ensemble
 * covers its members: separates the cost of the total footprint from the
 * cost of many allocations.  parts = 1 is lbm_hip_copy_bandwidth. */
int lbm_hip_copy_bandwidth_parts(int device, size_t bytes, int parts, int reps, double* gbs);

/* Message of the last failed call on this thread ("" if none). */
const char* lbm_hip_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* LBM_HIP_H */
