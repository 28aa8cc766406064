// lbm_hip.hip -- the C ABI (include/lbm_hip.h): contexts, device memory,
// launch sequencing and host transfers.  All compute runs in the kernels of
// kernels_fp64.hip / kernels_fast.hip.
#include "lbm_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "launch.hpp"

using namespace lbmhip;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

#define HIP_TRY(expr)                                                                                  \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            return fail(LBM_HIP_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));          \
    } while (0)

size_t storage_bytes(int prec) {
    switch (prec) {
        case LBM_HIP_FP64: return 8;
        case LBM_HIP_FP32: return 4;
        default: return 2;
    }
}

}  // namespace

struct lbm_hip_ctx {
    lbm_hip_params p{};
    // when set, lbm_hip_step records these around the collide-stream update
    // (bulk kernel, plus the boundary kernel on even steps), for
    // lbm_hip_time_run
    hipEvent_t kev[2] = {nullptr, nullptr};
    int device = 0;
    hipStream_t stream = nullptr;
    int NX = 0, NY = 0, NZ = 0;
    uint32_t N = 0;
    uint32_t SS = 0;  // slot stride in elements (PX * NY * NZ plus padding)
    uint32_t PX = 0;  // population row pitch: NX, or NX rounded up to 64 nodes
    int prec = 0;
    size_t ssize = 8, msize = 8;
    void* f = nullptr;
    uint8_t* flag = nullptr;
    void* rho = nullptr;
    void* ux = nullptr;
    void* uy = nullptr;
    void* uz = nullptr;
    DevState* st = nullptr;
    MassPart* mass_part = nullptr;  // >= max(G, ceil(N/256)) fixed-point partials
    double* diag_part = nullptr;  // 5 * ceil(N/256) doubles
    double* diag_out = nullptr;   // 5 doubles
    double* red_part = nullptr;   // 8 * max(LINK_MAX_BLOCKS, ceil(N/256)) doubles (link / wall kernels)
    int cpl_blocks = 0;           // blocks of the link pass (LINK_THREADS threads each)
    // PARTICLE link list (links_kernel.hpp) and what the host knows of it:
    // LIST_CHECKED after a link pass or a build, LIST_UNCHECKED after a remap
    // that followed one (the next link pass removes the links it undid),
    // LIST_DIRTY when it must be rebuilt from the flags before the next pass
    uint2* links = nullptr;
    uint32_t link_cap = 0;
    int list_state = 0;
    // lbm_hip_step leaves the link pass to the coupling launch of
    // lbm_hip_advance_squirmer; anything that reads its results first runs it
    bool links_pending = false;
    uint32_t* wall_rows = nullptr;  // rows next to a wall (even-step boundary chunks)
    BoundaryArgs bargs{};
    int bblocks = 0;                // blocks of aa_boundary_kernel
    int gx = 0, gy = 0, gz = 0, tx_shift = 8, ty = 1, G = 0, nb256 = 0;
    int xcd_remap = 2;  // whole tile rows per XCD until set_tiling / autotune chooses
    // cap on the collide-stream blocks resident per CU on even / odd steps
    // (0: none) and the dynamic LDS per block that enforces it
    int occ_cap[2] = {0, 0};
    unsigned occ_lds[2] = {0, 0};
    // odd steps in row blocks (aa_odd_row_kernel; 0: tiles, 1: blocks in
    // order, 2: block ranges per XCD) and the dynamic LDS that holds those
    // blocks to the odd cap's wavefronts per CU
    int odd_rows = 0;
    int row_waves = 0, rows_per_block = 1, row_blocks = 0;  // PX / 64, rows per block, blocks per odd step
    int batch_row_blocks = 0;  // row blocks of this member in its batch's odd steps
    unsigned occ_lds_rows = 0;
    int G_max = 0;  // largest tile count over the tile shapes (mass_part holds G_max + bblocks)
    double omPlus = 0, omMinus = 0, nu = 0;
    FastCoef K{};
    long long t = 0;  // steps done
    bool has_squirmer = false;
    bool macro_valid = true;  // initial rho = 1, u = 0 are the reference's values
    bool wallf_valid = false;
    bool mass_valid = false;
    int last_parts = 0;   // collide-stream mass partials written by the last step (tiles or rows)
    int last_bparts = 0;  // boundary-kernel mass partials written by the last step
    size_t device_bytes = 0;
    std::vector<hipEvent_t> events;
    // one captured cycle of the driver loop (step, advance, mass correction),
    // and graph_reps cycles in one graph (graph_long): each graph launch
    // costs a gap of several microseconds on the device
    hipGraphExec_t graph = nullptr, graph_long = nullptr;
    int graph_cycle = 0, graph_mass_every = -1, graph_reps = 1;
    bool owned_by_batch = false;  // member of an lbm_hip_batch
    double radius_max = 0.0;      // radius the boundary work list is sized for
    int remaps_since_step = 0;    // lbm_hip_advance_squirmer calls since the last lbm_hip_step
};

// Staging of lbm_hip_download / lbm_hip_upload_populations: one array of
// STAGE_NODES x 19 doubles (40 MB) per device for the life of the process,
// allocated by the first transfer on that device and used under
// g_stage_mu, so a transfer of any lattice moves in chunks through it and
// allocates nothing once it exists.
namespace {

constexpr size_t STAGE_NODES = size_t(1) << 18;
std::mutex g_stage_mu;

int staging(int device, double** out) {
    static std::vector<double*> buf;
    if (device < 0) return fail(LBM_HIP_EINVAL, "device ordinal out of range");
    if ((size_t)device >= buf.size()) buf.resize((size_t)device + 1, nullptr);
    if (!buf[(size_t)device]) {
        if (hipSetDevice(device) != hipSuccess) return fail(LBM_HIP_EDEVICE, "hipSetDevice failed");
        if (hipMalloc((void**)&buf[(size_t)device], sizeof(double) * Q * STAGE_NODES) != hipSuccess) {
            buf[(size_t)device] = nullptr;
            return fail(LBM_HIP_ENOMEM, "transfer staging allocation failed");
        }
    }
    *out = buf[(size_t)device];
    return LBM_HIP_OK;
}

}  // namespace

// Error text of the other entry points of the library (the MLP surrogate).
void lbmhip_set_error_text(const char* msg) { g_err = msg; }

extern "C" {

const char* lbm_hip_last_error(void) { return g_err.c_str(); }

void lbm_hip_default_params(lbm_hip_params* p) {
    if (!p) return;
    std::memset(p, 0, sizeof *p);
    // struct Params defaults, params.hpp:16-56
    p->nx = 64; p->ny = 34; p->nz = 32;
    p->tau = 1.0; p->lambda = 3.0 / 16.0;
    p->walls_y = 1; p->walls_z = 0;
    p->squirmer = 0;
    p->radius = 4.0; p->b1 = 0.015; p->beta = 0.0; p->rho_p = 1.0;
    p->x0 = -1; p->y0 = -1; p->z0 = -1;
    p->angle = 30.0; p->planar = 1;
    p->eps = 0.2; p->hc = 2.0; p->hmin = 0.5;
    p->plane_xz = 0;
    p->precision = LBM_HIP_FP64;
    p->device = 0;
}

// Params::validate, params.cpp:120-140 (same rules and messages), plus the
// 32-bit node indexing limit of this implementation and a precision check.
int lbm_hip_validate(const lbm_hip_params* p) {
    if (!p) return fail(LBM_HIP_EINVAL, "null parameters");
    auto bad = [](const char* m) { return fail(LBM_HIP_EINVAL, m); };
    if (p->nx < 1 || p->ny < 1 || p->nz < 1) return bad("grid dimensions must be positive");
    if (p->walls_y && p->ny < 4) return bad("ny must be >= 4 with walls in y");
    if (p->walls_z && p->nz < 4) return bad("nz must be >= 4 with walls in z");
    if (p->tau <= 0.5) return bad("tau must be > 0.5");
    if (p->lambda <= 0.0) return bad("lambda must be > 0");
    if (p->precision < LBM_HIP_FP64 || p->precision > LBM_HIP_BF16) return bad("unknown precision");
    // node and in-slot indices are 32-bit (population i of a node sits at
    // i * SS + that index, formed in 64 bits)
    const double N = (double)p->nx * p->ny * p->nz + 65536.0;
    if (N >= 4294967296.0) return bad("grid too large for 32-bit node indexing (nx*ny*nz must be < 2^32 - 65536)");
    if (!p->squirmer) return LBM_HIP_OK;
    if (p->radius <= 1.0) return bad("squirmer radius must be > 1 lattice unit");
    if (std::fabs(p->angle) >= 45.0) return bad("initial angle must satisfy |angle| < 45 deg from the x-axis");
    if (p->rho_p <= 0.0) return bad("rho_p must be > 0");
    if (p->eps < 0.0 || p->hc <= 0.0 || p->hmin < 0.0) return bad("eps >= 0, hc > 0, hmin >= 0 required");
    const double R = p->radius;
    const double y = p->y0 >= 0 ? p->y0 : 0.5 * (p->ny - 1);
    if (p->walls_y && (y - R - 0.5 < p->hmin || (p->ny - 1.5) - y - R < p->hmin))
        return bad("squirmer initially overlaps (or is too close to) a y-wall");
    if (2.0 * R + 2.0 > p->nx || (!p->walls_y && 2.0 * R + 2.0 > p->ny) || 2.0 * R + 2.0 > p->nz)
        return bad("domain too small for the squirmer");
    if (2.0 * p->b1 / 3.0 > 0.05) return bad("swimming speed U0 = 2 B1 / 3 too large (Mach number); keep U0 < 0.05");
    return LBM_HIP_OK;
}

}  // extern "C"

namespace {

template <class F>
int dispatch(int prec, F&& fn) {
    switch (prec) {
        case LBM_HIP_FP64: return fn((double*)nullptr);
        case LBM_HIP_FP32: return fn((float*)nullptr);
        case LBM_HIP_FP16: return fn((_Float16*)nullptr);
        default: return fn((bf16_t*)nullptr);
    }
}

// Squirmer::init (squirmer.cpp:20-30) on the host, in FP64.
void init_squirmer(DevSquirmer& s, const lbm_hip_params& p, double nu, double xc, double yc, double zc) {
    std::memset(&s, 0, sizeof s);
    s.R = p.radius; s.B1 = p.b1; s.beta = p.beta; s.planar = p.planar; s.plane_xz = p.plane_xz;
    s.M = p.rho_p * 4.0 / 3.0 * kPi * s.R * s.R * s.R;
    s.I = 0.4 * s.M * s.R * s.R;
    s.X[0] = xc; s.X[1] = yc; s.X[2] = zc;
    const double a = p.angle * kPi / 180.0;
    if (p.plane_xz) { s.e[0] = std::cos(a); s.e[1] = 0.0; s.e[2] = std::sin(a); }
    else { s.e[0] = std::cos(a); s.e[1] = std::sin(a); s.e[2] = 0.0; }
    s.eps = p.eps; s.hc = p.hc; s.hmin = p.hmin;
    s.FS = 6.0 * kPi * nu * s.R * (2.0 * s.B1 / 3.0);
}

void band_of(const DevSquirmer& s, int NX, int NY, int NZ, int band[6]) {
    const int Nd[3] = {NX, NY, NZ};
    for (int d = 0; d < 3; ++d) {
        const int a = (int)std::floor(s.X[d] - s.R) - 4, b = (int)std::ceil(s.X[d] + s.R) + 4;
        int w = b - a + 1;
        if (w > Nd[d]) w = Nd[d];
        band[d] = ((a % Nd[d]) + Nd[d]) % Nd[d];
        band[3 + d] = w;
    }
}

// The lattice has a body force: its steps run the collision with the
// forcing terms (STEP_BODY_FORCE).
bool body_force(const lbm_hip_ctx* c) { return c->p.fx != 0.0 || c->p.fy != 0.0 || c->p.fz != 0.0; }

FastCoef make_coef(double omP, double omM, double Fx, double Fy, double Fz, double scale) {
    FastCoef k{};
    const double gP = 1.0 - 0.5 * omP, gM = 1.0 - 0.5 * omM;
    const double w1 = 1.0 / 18, w2 = 1.0 / 36, w0 = 1.0 / 3.0;
    k.inv_s = (float)(1.0 / scale);
    k.s = (float)scale;
    k.a = (float)(0.5 * omP);
    k.b = (float)(0.5 * omM);
    k.one_m_omp = (float)(1.0 - omP);
    k.kp_axis = (float)(omP * w1); k.kp_diag = (float)(omP * w2); k.kp_rest = (float)(omP * w0);
    k.km_axis = (float)(3.0 * omM * w1); k.km_diag = (float)(3.0 * omM * w2);
    k.hFx = (float)(0.5 * scale * Fx); k.hFy = (float)(0.5 * scale * Fy); k.hFz = (float)(0.5 * scale * Fz);
    k.gPw9[0] = (float)(9.0 * gP * w1); k.gPw9[1] = (float)(9.0 * gP * w2); k.gPw9[2] = (float)(9.0 * gP * w0);
    k.gPw3[0] = (float)(3.0 * gP * w1); k.gPw3[1] = (float)(3.0 * gP * w2); k.gPw3[2] = (float)(3.0 * gP * w0);
    k.sm_axis = (float)(3.0 * scale * gM * w1); k.sm_diag = (float)(3.0 * scale * gM * w2);
    k.Fx = (float)Fx; k.Fy = (float)Fy; k.Fz = (float)Fz;
    k.one = 1.0f;
    return k;
}

template <class S>
StepArgs<S> step_args(const lbm_hip_ctx* c) {
    StepArgs<S> a{};
    a.f = (S*)c->f; a.flag = c->flag;
    a.rho = c->rho; a.ux = c->ux; a.uy = c->uy; a.uz = c->uz;
    a.st = c->st; a.mass_part = c->mass_part;
    a.NX = c->NX; a.NY = c->NY; a.NZ = c->NZ; a.N = c->N; a.SS = c->SS; a.PX = c->PX;
    a.walls_y = c->p.walls_y; a.walls_z = c->p.walls_z; a.has_squirmer = c->has_squirmer;
    a.gx = c->gx; a.gy = c->gy; a.gz = c->gz; a.tx_shift = c->tx_shift; a.ty = c->ty;
    a.nchx = (c->NX + 63) / 64;
    a.div_gx = make_fast_div((uint32_t)a.gx); a.div_gy = make_fast_div((uint32_t)a.gy);
    a.div_nchx = make_fast_div((uint32_t)a.nchx); a.div_ny = make_fast_div((uint32_t)a.NY);
    a.xcd_remap = c->xcd_remap;
    a.odd_rows = c->odd_rows;
    a.row_waves = c->row_waves; a.rows_per_block = c->rows_per_block;
    a.omPlus = c->omPlus; a.omMinus = c->omMinus;
    a.Fx = c->p.fx; a.Fy = c->p.fy; a.Fz = c->p.fz;
    a.t = c->t;
    return a;
}

template <class S>
RemapArgs<S> remap_args(const lbm_hip_ctx* c) {
    RemapArgs<S> a{};
    a.f = (S*)c->f; a.flag = c->flag;
    a.rho = c->rho; a.ux = c->ux; a.uy = c->uy; a.uz = c->uz;
    a.st = c->st;
    a.NX = c->NX; a.NY = c->NY; a.NZ = c->NZ; a.N = c->N; a.SS = c->SS; a.PX = c->PX;
    a.walls_y = c->p.walls_y; a.walls_z = c->p.walls_z;
    a.shifted = (c->t > 0 && ((c->t - 1) & 1) == 0) ? 1 : 0;
    a.shift_slot = (int)(c->t & 1);
    a.accumulate = c->remaps_since_step > 0 ? 1 : 0;
    a.list = c->links; a.cap = c->link_cap;
    a.maintain = 0;
    return a;
}

template <class S>
LinkArgs<S> link_args(const lbm_hip_ctx* c) {
    LinkArgs<S> a{};
    a.f = (S*)c->f; a.flag = c->flag; a.rho = c->rho; a.st = c->st; a.part = c->red_part;
    a.list = c->links; a.cap = c->link_cap;
    a.NX = c->NX; a.NY = c->NY; a.NZ = c->NZ; a.SS = c->SS; a.PX = c->PX;
    a.walls_y = c->p.walls_y; a.walls_z = c->p.walls_z;
    return a;
}

enum { LIST_DIRTY = 0, LIST_CHECKED = 1, LIST_UNCHECKED = 2 };

// Blocks of the link pass.  A sphere of radius R has about 72 R^2 PARTICLE
// links (D3Q19); 100 R^2 list entries leave room for the dead ones, at one
// entry per thread (a longer list is strided over).
int cpl_blocks_for(double R) {
    const double entries = std::max(1.0, 100.0 * R * R);
    return std::min(LINK_MAX_BLOCKS, std::max(1, (int)std::ceil(entries / LINK_THREADS)));
}

// Capacity of the link list for spheres up to radius R.  Every link-owning
// FLUID node lies in the shell rf <= |r| < rf + sqrt(2) around the centre
// the flags were classified at, and owns at most 18 links; the unit cube
// around a lattice point of the shell lies in the shell widened by sqrt(3)/2
// on both sides, which bounds the points.  Twice that bound holds the dead
// entries the compaction tolerates, plus the appends of one full remap.
uint32_t link_capacity(double R) {
    const double a = std::max(0.0, R - 0.8660254037844386), b = R + 1.4142135623730951 + 0.8660254037844386;
    const double shell = 4.0 / 3.0 * M_PI * (b * b * b - a * a * a) + 64.0;
    const double cap = 2.0 * (Q - 1) * shell + 2.0 * (Q - 1) * REMAP_CAP + 4096.0;
    return (uint32_t)std::min(cap, 4.0e9);
}

// Population row pitch.  A wavefront covers 64 consecutive x of one row;
// when rows start on a multiple of 64 nodes each wavefront's access to a
// population is one aligned segment (not two partial ones), which measured
// 1.6x faster on rows of 500 or 456 nodes.  Lanes past NX load and store
// nothing, so the pad costs memory, not traffic: rows are padded when that
// adds at most half again (LBM_HIP_ROW_PAD=0 turns it off).
uint32_t row_pitch(int nx, int ny, int nz) {
    const uint32_t px = ((uint32_t)nx + 63u) / 64u * 64u;
    if (const char* e = std::getenv("LBM_HIP_ROW_PAD"))
        if (std::atoi(e) == 0) return (uint32_t)nx;
    if (px == (uint32_t)nx || 2 * px > 3 * (uint32_t)nx) return (uint32_t)nx;
    if ((double)px * ny * nz + 65536.0 >= 4294967296.0) return (uint32_t)nx;
    return px;
}

// Slot stride: N rounded up to 256 bytes, plus a pad so that the 19
// population streams do not sit at the same offset modulo large powers of
// two (LBM_HIP_SLOT_PAD overrides the pad, in bytes).
uint32_t slot_stride(uint32_t N, size_t ssize) {
    size_t bytes = ((size_t)N * ssize + 255) / 256 * 256;
    size_t pad = 33 * 256;
    if (const char* e = std::getenv("LBM_HIP_SLOT_PAD")) pad = (size_t)std::strtoull(e, nullptr, 10) / 256 * 256;
    return (uint32_t)((bytes + pad) / ssize);
}

WallForceArgs wall_force_args(const lbm_hip_ctx* c) {
    WallForceArgs w;
    w.wall_rows = c->bargs.wall_rows;
    w.n_wall_rows = c->bargs.n_wall_rows;
    w.nchx = c->bargs.nchx;
    w.walls_y = c->p.walls_y;
    w.walls_z = c->p.walls_z;
    return w;
}

bool layout_shifted(const lbm_hip_ctx* c) { return c->t > 0 && ((c->t - 1) & 1) == 0; }

void drop_graphs(lbm_hip_ctx* c) {
    for (hipGraphExec_t* g : {&c->graph, &c->graph_long})
        if (*g) { hipGraphExecDestroy(*g); *g = nullptr; }
}

// Tiles of a block of BLOCK threads covering 2^ts x (BLOCK >> ts) nodes, or
// (ts = 0) BLOCK / 64 consecutive 64-node row chunks.
int tiles_for(int nx, int ny, int nz, int ts) {
    if (ts == 0) {
        const long long chunks = (long long)((nx + 63) / 64) * ny * nz;
        return (int)((chunks + BLOCK / 64 - 1) / (BLOCK / 64));
    }
    const int ty = BLOCK >> ts;
    return ((nx + (1 << ts) - 1) >> ts) * ((ny + ty - 1) / ty) * nz;
}

// Tile shape of the collide-stream kernel (ts = log2 of the tile width, 6..8,
// or 0 for row runs).  The boundary kernel's mass partials follow the step
// kernel's.
// Dynamic LDS bytes per block that leave room for at most `blocks` blocks of
// the collide-stream kernel on a CU (0: no cap).  The kernel's own LDS
// (under 512 bytes) fits in the 512-byte margin below the CU's LDS / blocks,
// and the result stays above the CU's LDS / (blocks + 1).
int occupancy_lds(int device, int blocks, unsigned* bytes) {
    *bytes = 0;
    if (blocks == 0) return LBM_HIP_OK;
    int per_cu = 0, per_block = 0;
    if (hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) != hipSuccess ||
        hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
        return fail(LBM_HIP_EDEVICE, "LDS size query failed");
    const long b = std::min((long)(per_cu / blocks) / 512 * 512, (long)per_block) - 512;
    if (b <= (long)(per_cu / (blocks + 1))) return fail(LBM_HIP_EINVAL, "occupancy cap out of range for this device");
    *bytes = (unsigned)b;
    return LBM_HIP_OK;
}

// Measured on the 512 x 256 x 128 channel: the FP32 and FP64 updates move
// their bytes fastest with 4 and 3 blocks per CU on both parities, FP16 and
// BF16 with the even step uncapped and the odd step at 6 and 7 -- with the
// odd step in tiles.  The odd step in row blocks is fastest uncapped
// (FP16 225.5 us per step against 257.2 us at the 24 wavefronts per CU of
// 6 tile blocks).
void default_occupancy(int prec, bool odd_rows, int cap[2]) {
    switch (prec) {
        case LBM_HIP_FP64: cap[0] = cap[1] = 3; break;
        case LBM_HIP_FP32: cap[0] = cap[1] = 4; break;
        case LBM_HIP_BF16: cap[0] = 0; cap[1] = 7; break;
        default: cap[0] = 0; cap[1] = 6; break;
    }
    if (odd_rows && prec != LBM_HIP_FP64) cap[1] = 0;
}

// Dynamic LDS per block of the row kernel (ROW_LDS_WORDS words of static LDS,
// `block_waves` wavefronts per block) that holds a CU to the wavefronts of
// `cap` 256-thread blocks, where the LDS allocation (512-byte granules) can
// express it; else 0 (no cap).
int row_occupancy_lds(int device, int cap, int block_waves, unsigned* bytes) {
    *bytes = 0;
    const int rb = cap * (BLOCK / 64) / block_waves;
    if (cap == 0 || rb < 1 || rb * block_waves >= 32) return LBM_HIP_OK;
    int per_cu = 0, per_block = 0;
    if (hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) != hipSuccess ||
        hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, device) != hipSuccess)
        return fail(LBM_HIP_EDEVICE, "LDS size query failed");
    const long stat = (long)ROW_LDS_WORDS * 8;  // a multiple of 512
    const long total = std::min((long)(per_cu / rb) / 512 * 512, (long)per_block);
    if (total > stat && total * (rb + 1) > per_cu) *bytes = (unsigned)(total - stat);
    return LBM_HIP_OK;
}

int set_occupancy(lbm_hip_ctx* c, int even_blocks, int odd_blocks) {
    for (int b : {even_blocks, odd_blocks})
        if (b != 0 && (b < 2 || b > 8)) return fail(LBM_HIP_EINVAL, "blocks per CU must be 0 (no cap) or 2..8");
    unsigned lds[2], lds_rows = 0;
    if (int rc = occupancy_lds(c->device, even_blocks, &lds[0])) return rc;
    if (int rc = occupancy_lds(c->device, odd_blocks, &lds[1])) return rc;
    if (c->odd_rows)
        if (int rc = row_occupancy_lds(c->device, odd_blocks, c->row_waves * c->rows_per_block, &lds_rows)) return rc;
    if (c->stream && hipStreamSynchronize(c->stream) != hipSuccess) return fail(LBM_HIP_EDEVICE, "stream sync");
    c->occ_cap[0] = even_blocks; c->occ_cap[1] = odd_blocks;
    c->occ_lds[0] = lds[0]; c->occ_lds[1] = lds[1];
    c->occ_lds_rows = lds_rows;
    drop_graphs(c);  // captured launches carry the old LDS size
    return LBM_HIP_OK;
}

void set_tile_shape(lbm_hip_ctx* c, int ts) {
    c->tx_shift = ts;
    if (ts == 0) {
        c->ty = 1;
        c->gx = tiles_for(c->p.nx, c->p.ny, c->p.nz, 0);
        c->gy = c->gz = 1;
    } else {
        c->ty = BLOCK >> ts;
        c->gx = (c->p.nx + (1 << ts) - 1) >> ts;
        c->gy = (c->p.ny + c->ty - 1) / c->ty;
        c->gz = c->p.nz;
    }
    c->G = c->gx * c->gy * c->gz;
    c->bargs.mass_part = c->mass_part ? c->mass_part + c->G : nullptr;
    drop_graphs(c);
}

void free_ctx(lbm_hip_ctx* c) {
    if (!c) return;
    hipSetDevice(c->device);
    if (c->stream) hipStreamSynchronize(c->stream);
    for (void* q : {c->f, (void*)c->flag, c->rho, c->ux, c->uy, c->uz, (void*)c->st, (void*)c->mass_part,
                    (void*)c->diag_part, (void*)c->diag_out, (void*)c->red_part, (void*)c->wall_rows, (void*)c->links})
        if (q) hipFree(q);
    for (hipEvent_t e : c->events) hipEventDestroy(e);
    drop_graphs(c);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
}

int check_device_error(lbm_hip_ctx* c) {
    int err = 0;
    HIP_TRY(hipMemcpyAsync(&err, &c->st->err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (err & ERR_REMAP_OVERFLOW)
        return fail(LBM_HIP_EOVERFLOW, "remap changed more than " + std::to_string(REMAP_CAP) +
                                           " nodes in one step; the lattice state is no longer valid");
    if (err & ERR_LINK_OVERFLOW)
        return fail(LBM_HIP_EOVERFLOW, "the particle link list outgrew its capacity of " +
                                           std::to_string(c->link_cap) + " entries");
    return LBM_HIP_OK;
}

// Rebuild the link list from the flags if the host invalidated it.
int ensure_link_list(lbm_hip_ctx* c) {
    if (!c->has_squirmer || c->list_state != LIST_DIRTY) return LBM_HIP_OK;
    const hipError_t e = launch_link_build(c->flag, c->st, c->links, c->link_cap, c->NX, c->NY, c->NZ, c->stream);
    if (e != hipSuccess) return fail(LBM_HIP_EDEVICE, std::string("link list build: ") + hipGetErrorString(e));
    c->list_state = LIST_CHECKED;
    return LBM_HIP_OK;
}

// The coupling launch: the link pass left pending by lbm_hip_step (`links`)
// and / or the squirmer update with its remap (`remap`).
int launch_coupling(lbm_hip_ctx* c, bool links, bool remap) {
    int rc;
    if (links && (rc = ensure_link_list(c))) return rc;
    const bool maintain = links || c->list_state == LIST_CHECKED;
    hipError_t e = hipSuccess;
    dispatch(c->prec, [&](auto* tag) -> int {
        using S = std::remove_pointer_t<decltype(tag)>;
        RemapArgs<S> r = remap_args<S>(c);
        r.maintain = maintain ? 1 : 0;
        e = launch_couple<S>(link_args<S>(c), r, c->cpl_blocks, links ? 1 : 0, remap ? 1 : 0, c->stream);
        return 0;
    });
    if (e != hipSuccess) return fail(LBM_HIP_EDEVICE, std::string("coupling launch: ") + hipGetErrorString(e));
    if (links) c->links_pending = false;
    if (remap) c->list_state = maintain ? LIST_UNCHECKED : LIST_DIRTY;
    else c->list_state = LIST_CHECKED;
    return LBM_HIP_OK;
}

// Run the link pass of the last step if it is still pending.
int flush_links(lbm_hip_ctx* c) {
    return c->links_pending ? launch_coupling(c, true, false) : LBM_HIP_OK;
}

}  // namespace

extern "C" {

int lbm_hip_create(const lbm_hip_params* pin, lbm_hip_ctx** out) {
    if (!out) return fail(LBM_HIP_EINVAL, "null output pointer");
    *out = nullptr;
    int rc = lbm_hip_validate(pin);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(LBM_HIP_EDEVICE, "no HIP device available (liblbm_hip needs an AMD GPU)");
    if (pin->device < 0 || pin->device >= ndev) return fail(LBM_HIP_EINVAL, "device ordinal out of range");

    auto* c = new lbm_hip_ctx();
    c->p = *pin;
    c->device = pin->device;
    c->prec = pin->precision;
    c->NX = pin->nx; c->NY = pin->ny; c->NZ = pin->nz;
    c->N = (uint32_t)((size_t)pin->nx * pin->ny * pin->nz);
    c->ssize = storage_bytes(c->prec);
    c->PX = row_pitch(pin->nx, pin->ny, pin->nz);
    // Odd steps in row blocks where a stored row is whole wavefronts (padded
    // or not) and fits a block, for FP16 storage (measured on 512 x 256 x
    // 128: FP16 odd step 198-208 against 210-215 us in tiles; BF16 2 %
    // slower in rows, FP32 and FP64 within 1 %).  Rows of fewer than 4
    // wavefronts share a block.  LBM_HIP_ODD_ROWS: 0 off, 1 blocks in order,
    // 2 block ranges per XCD, for any storage.
    if (c->PX % 64 == 0 && c->PX <= 64 * ROW_MAX_WAVES) {
        c->odd_rows = c->prec == LBM_HIP_FP16 ? 1 : 0;
        if (const char* e = std::getenv("LBM_HIP_ODD_ROWS")) c->odd_rows = std::max(0, std::min(2, std::atoi(e)));
        c->row_waves = (int)(c->PX / 64);
        c->rows_per_block = c->row_waves >= 4 ? 1 : 4 / c->row_waves;
        c->row_blocks = (pin->ny * pin->nz + c->rows_per_block - 1) / c->rows_per_block;
    }
    c->SS = slot_stride(c->PX * (uint32_t)pin->ny * (uint32_t)pin->nz, c->ssize);
    c->msize = c->prec == LBM_HIP_FP64 ? 8 : 4;
    // lbm.cpp:67-69
    c->omPlus = 1.0 / pin->tau;
    c->omMinus = 1.0 / (0.5 + pin->lambda / (pin->tau - 0.5));
    c->nu = (pin->tau - 0.5) / 3.0;
    const double scale = c->prec == LBM_HIP_FP16 ? 32768.0 : 1.0;
    c->K = make_coef(c->omPlus, c->omMinus, pin->fx, pin->fy, pin->fz, scale);

    // tiles: one wavefront row-chunk along x, block of 256 threads
    for (int ts : {0, 6, 7, 8}) c->G_max = std::max(c->G_max, tiles_for(pin->nx, pin->ny, pin->nz, ts));
    // one dispatch holds fewer than 2^32 work-items
    if ((double)c->G_max * BLOCK >= 4294967296.0) {
        delete c;
        return fail(LBM_HIP_EINVAL, "grid too large: its collide-stream tiles exceed 2^32 work-items per launch");
    }
    set_tile_shape(c, pin->nx <= 64 ? 6 : (pin->nx <= 128 ? 7 : 8));
    c->nb256 = (int)(((size_t)c->N + 255) / 256);

    // even-step boundary chunks: rows next to a wall, and the squirmer's box rows
    std::vector<uint32_t> wall_rows;
    for (int z = 0; z < pin->nz; ++z)
        for (int y = 0; y < pin->ny; ++y) {
            const bool solid_row = (pin->walls_y && (y == 0 || y == pin->ny - 1)) ||
                                   (pin->walls_z && (z == 0 || z == pin->nz - 1));
            const bool next_to_wall = (pin->walls_y && (y == 1 || y == pin->ny - 2)) ||
                                      (pin->walls_z && (z == 1 || z == pin->nz - 2));
            if (!solid_row && next_to_wall) wall_rows.push_back((uint32_t)y + (uint32_t)pin->ny * (uint32_t)z);
        }
    {
        BoundaryArgs& b = c->bargs;
        b.n_wall_rows = (int)wall_rows.size();
        b.nchx = (pin->nx + 63) / 64;
        b.wy_lo = pin->walls_y ? 1 : -1; b.wy_hi = pin->walls_y ? pin->ny - 2 : -1;
        b.wz_lo = pin->walls_z ? 1 : -1; b.wz_hi = pin->walls_z ? pin->nz - 2 : -1;
        if (pin->squirmer) {
            // boundary box: at most ceil(2r) + 4 rows per axis for the covering sphere, r < R + 1
            const int w = (int)std::ceil(2.0 * pin->radius) + 7;
            b.box_rows_max = std::min(w, pin->ny) * std::min(w, pin->nz);
            b.box_chunks = std::min(b.nchx, (w + 63) / 64 + 2);
            c->radius_max = pin->radius;
        }
        const long long waves = (long long)b.n_wall_rows * b.nchx + (long long)b.box_rows_max * b.box_chunks;
        c->bblocks = (int)((waves + (BLOCK / 64) - 1) / (BLOCK / 64));
    }

    auto cleanup = [&](int code) {
        std::string m = g_err;
        free_ctx(c);
        g_err = m;
        return code;
    };
    if (hipSetDevice(c->device) != hipSuccess) return cleanup(fail(LBM_HIP_EDEVICE, "hipSetDevice failed"));
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess)
        return cleanup(fail(LBM_HIP_EDEVICE, "hipStreamCreate failed"));
    {
        int cap[2];
        default_occupancy(c->prec, c->odd_rows != 0, cap);
        if (int rc2 = set_occupancy(c, cap[0], cap[1])) return cleanup(rc2);
    }

    const size_t N = c->N;
    auto dmalloc = [&](void** q, size_t bytes) -> bool {
        if (hipMalloc(q, bytes) != hipSuccess) { *q = nullptr; return false; }
        c->device_bytes += bytes;
        return true;
    };
    const int nparts = std::max({c->G_max + c->bblocks, c->nb256, c->p.ny * c->p.nz});
    if (!dmalloc(&c->f, (size_t)Q * c->SS * c->ssize) || !dmalloc((void**)&c->flag, N) ||
        !dmalloc(&c->rho, N * c->msize) || !dmalloc(&c->ux, N * c->msize) || !dmalloc(&c->uy, N * c->msize) ||
        !dmalloc(&c->uz, N * c->msize) || !dmalloc((void**)&c->st, sizeof(DevState)) ||
        !dmalloc((void**)&c->mass_part, sizeof(MassPart) * nparts) ||
        !dmalloc((void**)&c->diag_part, sizeof(double) * 5 * c->nb256) ||
        !dmalloc((void**)&c->diag_out, sizeof(double) * 5) ||
        !dmalloc((void**)&c->red_part, sizeof(double) * 8 * std::max(LINK_MAX_BLOCKS, c->nb256)) ||
        !dmalloc((void**)&c->wall_rows, sizeof(uint32_t) * std::max<size_t>(1, wall_rows.size())))
        return cleanup(fail(LBM_HIP_ENOMEM, "device allocation failed"));

    // flags (lbm.cpp:72-93): WALL layers and BOUNDARY bits are set by a
    // device kernel; the host classifies the initial sphere's box and
    // uploads its PARTICLE nodes.  FLUID count: the nodes of the rows that
    // are not wall layers, less the PARTICLE nodes.
    auto wall_at = [&](int y, int z) {
        return (pin->walls_y && (y == 0 || y == c->NY - 1)) || (pin->walls_z && (z == 0 || z == c->NZ - 1));
    };
    long long fluid_rows = 0;
    for (int z = 0; z < c->NZ; ++z)
        for (int y = 0; y < c->NY; ++y) fluid_rows += wall_at(y, z) ? 0 : 1;
    std::vector<uint32_t> particle_nodes;
    DevState hs;
    std::memset(&hs, 0, sizeof hs);
    c->has_squirmer = pin->squirmer != 0;
    if (c->has_squirmer) {
        init_squirmer(hs.sq, *pin, c->nu, pin->x0 >= 0 ? pin->x0 : 0.5 * (c->NX - 1),
                      pin->y0 >= 0 ? pin->y0 : 0.5 * (c->NY - 1), pin->z0 >= 0 ? pin->z0 : 0.5 * (c->NZ - 1));
        // updateParticleFlags(true) (lbm.cpp:271-291): the box is cut to one
        // period per axis, so no node is listed twice
        const DevSquirmer& s = hs.sq;
        const double R = s.R;
        const int x0 = (int)std::floor(s.X[0] - R) - 2, x1 = (int)std::ceil(s.X[0] + R) + 2;
        int y0 = (int)std::floor(s.X[1] - R) - 2, y1 = (int)std::ceil(s.X[1] + R) + 2;
        const int z0 = (int)std::floor(s.X[2] - R) - 2, z1 = (int)std::ceil(s.X[2] + R) + 2;
        if (pin->walls_y) { y0 = std::max(0, y0); y1 = std::min(c->NY - 1, y1); }
        const int wx = std::min(x1 - x0 + 1, c->NX), wy = std::min(y1 - y0 + 1, c->NY), wz = std::min(z1 - z0 + 1, c->NZ);
        for (int kz = 0; kz < wz; ++kz)
            for (int ky = 0; ky < wy; ++ky)
                for (int kx = 0; kx < wx; ++kx) {
                    const int x = (((x0 + kx) % c->NX) + c->NX) % c->NX;
                    const int y = (((y0 + ky) % c->NY) + c->NY) % c->NY;
                    const int z = (((z0 + kz) % c->NZ) + c->NZ) % c->NZ;
                    if (wall_at(y, z)) continue;
                    double r[3];
                    r[0] = x - s.X[0]; r[0] -= c->NX * std::round(r[0] / c->NX);
                    r[1] = y - s.X[1]; if (!pin->walls_y) r[1] -= c->NY * std::round(r[1] / c->NY);
                    r[2] = z - s.X[2]; if (!pin->walls_z) r[2] -= c->NZ * std::round(r[2] / c->NZ);
                    if (r[0] * r[0] + r[1] * r[1] + r[2] * r[2] < R * R)
                        particle_nodes.push_back((uint32_t)x + (uint32_t)c->NX * ((uint32_t)y + (uint32_t)c->NY * (uint32_t)z));
                }
        for (int d = 0; d < 3; ++d) hs.xf[d] = s.X[d];
        hs.rf = R;
        band_of(hs.sq, c->NX, c->NY, c->NZ, hs.band);
        c->cpl_blocks = cpl_blocks_for(hs.sq.R);
        c->link_cap = link_capacity(hs.sq.R);
        if (!dmalloc((void**)&c->links, sizeof(uint2) * c->link_cap))
            return cleanup(fail(LBM_HIP_ENOMEM, "device allocation failed (link list)"));
    } else {
        hs.sq.R = 4.0; hs.sq.B1 = 0.015; hs.sq.planar = 1; hs.sq.e[0] = 1.0;
        hs.sq.eps = 0.2; hs.sq.hc = 2.0; hs.sq.hmin = 0.5;
    }
    hs.nfluid = fluid_rows * c->NX - (long long)particle_nodes.size();
    uint32_t* dparticles = nullptr;
    if (!particle_nodes.empty() &&
        (hipMalloc((void**)&dparticles, sizeof(uint32_t) * particle_nodes.size()) != hipSuccess ||
         hipMemcpyAsync(dparticles, particle_nodes.data(), sizeof(uint32_t) * particle_nodes.size(),
                        hipMemcpyHostToDevice, c->stream) != hipSuccess)) {
        if (dparticles) hipFree(dparticles);
        return cleanup(fail(LBM_HIP_ENOMEM, "context initialisation: particle list"));
    }
    const hipError_t fe = launch_init_flags(c->flag, c->NX, c->NY, c->NZ, c->N, pin->walls_y, pin->walls_z, dparticles,
                                            (int)particle_nodes.size(), c->stream);
    const hipError_t se = hipStreamSynchronize(c->stream);  // the list is freed next
    if (dparticles) hipFree(dparticles);
    if (fe != hipSuccess || se != hipSuccess) return cleanup(fail(LBM_HIP_EDEVICE, "context initialisation: flags"));

    c->bargs.wall_rows = c->wall_rows;
    c->bargs.mass_part = c->mass_part + c->G;
    if (!wall_rows.empty() &&
        hipMemcpyAsync(c->wall_rows, wall_rows.data(), sizeof(uint32_t) * wall_rows.size(), hipMemcpyHostToDevice,
                       c->stream) != hipSuccess)
        return cleanup(fail(LBM_HIP_EDEVICE, "context initialisation failed"));
    if (hipMemcpyAsync(c->st, &hs, sizeof hs, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        launch_fill_macro(c->rho, c->ux, c->uy, c->uz, c->prec == LBM_HIP_FP64, c->N, c->stream) != hipSuccess ||
        dispatch(c->prec, [&](auto* tag) -> int {
            using S = std::remove_pointer_t<decltype(tag)>;
            return launch_fill_populations<S>((S*)c->f, c->SS, c->stream) == hipSuccess ? 0 : -1;
        }) != 0 ||
        hipMemsetAsync(c->mass_part, 0, sizeof(MassPart) * nparts, c->stream) != hipSuccess)
        return cleanup(fail(LBM_HIP_EDEVICE, "context initialisation failed"));
    c->list_state = LIST_DIRTY;
    if ((rc = ensure_link_list(c))) return cleanup(rc);
    if (hipStreamSynchronize(c->stream) != hipSuccess)
        return cleanup(fail(LBM_HIP_EDEVICE, "context initialisation failed"));
    *out = c;
    return LBM_HIP_OK;
}

void lbm_hip_destroy(lbm_hip_ctx* ctx) {
    if (ctx && ctx->owned_by_batch) return;  // freed with its batch
    free_ctx(ctx);
}

int lbm_hip_step(lbm_hip_ctx* c, int flags) {
    if (!c) return fail(LBM_HIP_EINVAL, "null context");
    int mode = 0;
    if (flags & LBM_STEP_MACRO) mode |= STEP_MACRO;
    if (flags & LBM_STEP_MASS) mode |= STEP_MASS;
    if (body_force(c)) mode |= STEP_BODY_FORCE;
    const bool odd = (c->t & 1) != 0;
    if (int rc = flush_links(c)) return rc;
    hipError_t e = hipSuccess;
    dispatch(c->prec, [&](auto* tag) -> int {
        using S = std::remove_pointer_t<decltype(tag)>;
        const StepArgs<S> sa = step_args<S>(c);
        if (c->kev[0]) e = hipEventRecord(c->kev[0], c->stream);
        const unsigned lds = odd && c->odd_rows ? c->occ_lds_rows : c->occ_lds[odd];
        if (e == hipSuccess) e = launch_step<S>(sa, c->K, odd, mode, lds, c->stream);
        if (e == hipSuccess && !odd) e = launch_boundary<S>(sa, c->K, c->bargs, c->bblocks, mode, c->stream);
        if (e == hipSuccess && c->kev[1]) e = hipEventRecord(c->kev[1], c->stream);
        // the moving-boundary bounce-back and force/torque on the squirmer
        // run in the coupling launch (lbm_hip_advance_squirmer, or whatever
        // reads them first)
        if (e == hipSuccess && (flags & LBM_STEP_WALL_FORCE))
            e = launch_wall_force<S>((const S*)c->f, c->flag, wall_force_args(c), c->red_part, c->st, c->NX, c->NY,
                                     c->NZ, c->SS, c->PX, c->stream);
        return 0;
    });
    if (e != hipSuccess) return fail(LBM_HIP_EDEVICE, std::string("step launch: ") + hipGetErrorString(e));
    c->links_pending = c->has_squirmer;
    c->last_parts = odd && c->odd_rows ? c->row_blocks : c->G;
    c->last_bparts = odd ? 0 : c->bblocks;
    ++c->t;
    c->remaps_since_step = 0;
    c->macro_valid = (flags & LBM_STEP_MACRO) != 0;
    c->wallf_valid = (flags & LBM_STEP_WALL_FORCE) != 0;
    c->mass_valid = (flags & LBM_STEP_MASS) != 0;
    return LBM_HIP_OK;
}

int lbm_hip_advance_squirmer(lbm_hip_ctx* c) {
    if (!c) return fail(LBM_HIP_EINVAL, "null context");
    if (!c->has_squirmer) return LBM_HIP_OK;
    if (int rc = launch_coupling(c, c->links_pending, true)) return rc;
    ++c->remaps_since_step;
    return LBM_HIP_OK;
}

int lbm_hip_correct_mass(lbm_hip_ctx* c) {
    if (!c) return fail(LBM_HIP_EINVAL, "null context");
    if (int rc = flush_links(c)) return rc;
    hipError_t e;
    const int shift_slot = (int)(c->t & 1);
    if (c->mass_valid) {
        const int terms = MASS_TERM_LINKS | (c->remaps_since_step > 0 ? MASS_TERM_REMAP : 0);
        e = launch_mass_finalize(c->mass_part, c->last_parts + c->last_bparts, c->st, terms, shift_slot, c->stream);
    } else {
        const int shifted = layout_shifted(c) ? 1 : 0;
        e = hipSuccess;
        dispatch(c->prec, [&](auto* tag) -> int {
            using S = std::remove_pointer_t<decltype(tag)>;
            e = launch_mass<S>((const S*)c->f, c->flag, c->mass_part, c->NX, c->NY, c->NZ, c->N, c->SS, c->PX, shifted,
                               c->st, shift_slot, c->stream);
            return 0;
        });
        if (e == hipSuccess) e = launch_mass_finalize(c->mass_part, c->nb256, c->st, 0, shift_slot, c->stream);
    }
    if (e != hipSuccess) return fail(LBM_HIP_EDEVICE, std::string("mass correction: ") + hipGetErrorString(e));
    c->mass_valid = false;
    return LBM_HIP_OK;
}

// One driver-loop step (main.cpp:60-64) with the by-products it needs.
static int loop_step(lbm_hip_ctx* c, int mass_every, int extra_flags) {
    const long long tnext = c->t + 1;  // 1-based step number (main.cpp:60)
    const bool mass = c->has_squirmer && mass_every > 0 && tnext % mass_every == 0;
    int rc = lbm_hip_step(c, extra_flags | (mass ? LBM_STEP_MASS : 0));
    if (rc) return rc;
    if (c->has_squirmer) {
        if ((rc = lbm_hip_advance_squirmer(c))) return rc;
        if (mass && (rc = lbm_hip_correct_mass(c))) return rc;
    }
    return LBM_HIP_OK;
}

// Capture `cycles` cycles of the loop (a cycle: parity period 2 x mass
// period) into one graph.  The kernels see the step index only through its
// parity and the mass cadence, so a capture starting at any multiple of the
// cycle length replays the same work.
static int capture_cycles(lbm_hip_ctx* c, int mass_every, int cycles, hipGraphExec_t* out) {
    int rc0;
    if ((rc0 = flush_links(c)) || (rc0 = ensure_link_list(c))) return rc0;
    const long long t_save = c->t;
    const bool mv = c->macro_valid, wv = c->wallf_valid, ms = c->mass_valid;
    const int ls = c->list_state, rs = c->remaps_since_step;
    c->t = 0;  // capture with a cycle-aligned step index
    HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = LBM_HIP_OK;
    for (int s = 0; s < cycles * c->graph_cycle && rc == LBM_HIP_OK; ++s) rc = loop_step(c, mass_every, 0);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    c->t = t_save; c->macro_valid = mv; c->wallf_valid = wv; c->mass_valid = ms;
    c->list_state = ls; c->remaps_since_step = rs; c->links_pending = false;
    if (rc) { if (g) hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return fail(LBM_HIP_EDEVICE, std::string("graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) { *out = nullptr; return fail(LBM_HIP_EDEVICE, std::string("graph instantiate: ") + hipGetErrorString(e)); }
    return LBM_HIP_OK;
}

// The one-cycle graph for `mass_every`, and the multi-cycle one when a run
// of `steps` can use it (about 64 steps per launch).
static int ensure_graph(lbm_hip_ctx* c, int mass_every, int steps) {
    if (!(c->graph && c->graph_mass_every == mass_every)) {
        drop_graphs(c);
        c->graph_cycle = (c->has_squirmer && mass_every > 0) ? ((mass_every % 2 == 0) ? mass_every : 2 * mass_every) : 2;
        c->graph_reps = std::max(1, 64 / c->graph_cycle);
        const int rc = capture_cycles(c, mass_every, 1, &c->graph);
        if (rc) return rc;
        c->graph_mass_every = mass_every;
    }
    if (!c->graph_long && c->graph_reps > 1 && steps > c->graph_reps * c->graph_cycle)
        return capture_cycles(c, mass_every, c->graph_reps, &c->graph_long);
    return LBM_HIP_OK;
}

static int run_loop(lbm_hip_ctx* c, int steps, int mass_every, int last_flags, bool use_graph) {
    if (steps < 0) return fail(LBM_HIP_EINVAL, "steps must be >= 0");
    int rc;
    int left = steps;
    // eager until the step index is cycle-aligned; the last step stays eager
    // when it has flags of its own
    const int tail = last_flags ? 1 : 0;
    if (use_graph && steps > 1) {
        if ((rc = ensure_graph(c, mass_every, steps))) return rc;
        const int C = c->graph_cycle;
        while (left > 1 && (c->t % C) != 0) {
            if ((rc = loop_step(c, mass_every, 0))) return rc;
            --left;
        }
        // the graphs start with a step: no link pass pending, the list valid
        if ((rc = flush_links(c)) || (rc = ensure_link_list(c))) return rc;
        int done = 0;
        if (c->graph_long) {
            const int CL = C * c->graph_reps;
            for (; left - tail >= CL; left -= CL, done += CL) HIP_TRY(hipGraphLaunch(c->graph_long, c->stream));
        }
        for (; left - tail >= C; left -= C, done += C) HIP_TRY(hipGraphLaunch(c->graph, c->stream));
        if (done > 0) {
            c->t += done;
            c->macro_valid = c->wallf_valid = false;
            c->mass_valid = false;
            c->remaps_since_step = c->has_squirmer ? 1 : 0;
            if (c->has_squirmer) c->list_state = LIST_UNCHECKED;
        }
    }
    while (left > 0) {
        if ((rc = loop_step(c, mass_every, left == 1 ? last_flags : 0))) return rc;
        --left;
    }
    return LBM_HIP_OK;
}

int lbm_hip_run(lbm_hip_ctx* c, int steps, int mass_every, int last_flags) {
    if (!c) return fail(LBM_HIP_EINVAL, "null context");
    return run_loop(c, steps, mass_every, last_flags, true);
}

int lbm_hip_synchronize(lbm_hip_ctx* c) {
    if (!c) return fail(LBM_HIP_EINVAL, "null context");
    if (int rc = flush_links(c)) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return check_device_error(c);
}

int lbm_hip_get_squirmer(lbm_hip_ctx* c, lbm_squirmer_state* o) {
    if (!c || !o) return fail(LBM_HIP_EINVAL, "null argument");
    if (int rc = flush_links(c)) return rc;
    DevState hs;
    HIP_TRY(hipMemcpyAsync(&hs, c->st, sizeof hs, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (hs.err & ERR_REMAP_OVERFLOW) return check_device_error(c);
    const DevSquirmer& s = hs.sq;
    o->R = s.R; o->B1 = s.B1; o->beta = s.beta; o->M = s.M; o->I = s.I; o->FS = s.FS;
    for (int d = 0; d < 3; ++d) {
        o->X[d] = s.X[d]; o->e[d] = s.e[d]; o->U[d] = s.U[d]; o->W[d] = s.W[d];
        o->Fh[d] = s.Fh[d]; o->Th[d] = s.Th[d]; o->Fc[d] = s.Fc[d]; o->Tc[d] = s.Tc[d]; o->Fw[d] = s.Fw[d];
    }
    o->eps = s.eps; o->hc = s.hc; o->hmin = s.hmin;
    o->contacts = s.contacts;
    o->planar = s.planar; o->plane_xz = s.plane_xz;
    return LBM_HIP_OK;
}

int lbm_hip_set_squirmer(lbm_hip_ctx* c, const lbm_squirmer_state* in) {
    if (!c || !in) return fail(LBM_HIP_EINVAL, "null argument");
    DevSquirmer s;
    std::memset(&s, 0, sizeof s);
    s.R = in->R; s.B1 = in->B1; s.beta = in->beta; s.M = in->M; s.I = in->I; s.FS = in->FS;
    for (int d = 0; d < 3; ++d) {
        s.X[d] = in->X[d]; s.e[d] = in->e[d]; s.U[d] = in->U[d]; s.W[d] = in->W[d];
        s.Fh[d] = in->Fh[d]; s.Th[d] = in->Th[d]; s.Fc[d] = in->Fc[d]; s.Tc[d] = in->Tc[d]; s.Fw[d] = in->Fw[d];
    }
    s.eps = in->eps; s.hc = in->hc; s.hmin = in->hmin;
    s.contacts = in->contacts;
    s.planar = in->planar; s.plane_xz = in->plane_xz;
    if (c->has_squirmer) {
        if (!(s.R > 0.0) || s.R > c->radius_max)
            return fail(LBM_HIP_EINVAL, "squirmer radius must be positive and at most the radius the context was "
                                        "created with (the boundary work list and the link list are sized for it)");
        // the pending link pass belongs to the state it was stepped with
        if (int rc = flush_links(c)) return rc;
        // The PARTICLE flags stay those of the centre they were classified
        // at until the next remap, which only reclassifies around the new
        // centre (as lbm.cpp:271-291 does): the centre may move by less than
        // a node per axis, as in a time step.
        double xf[3];
        HIP_TRY(hipMemcpyAsync(xf, c->st->xf, sizeof xf, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        const int Nd[3] = {c->NX, c->NY, c->NZ};
        const bool per[3] = {true, !c->p.walls_y, !c->p.walls_z};
        for (int d = 0; d < 3; ++d) {
            double dx = s.X[d] - xf[d];
            if (per[d]) dx -= Nd[d] * std::round(dx / Nd[d]);
            if (!(std::fabs(dx) < 1.0))
                return fail(LBM_HIP_EINVAL, "the squirmer centre may move by less than one node per axis from "
                                            "where its flags were classified");
        }
    }
    HIP_TRY(hipMemcpyAsync(&c->st->sq, &s, sizeof s, hipMemcpyHostToDevice, c->stream));
    int band[6];
    band_of(s, c->NX, c->NY, c->NZ, band);
    HIP_TRY(hipMemcpyAsync(c->st->band, band, sizeof band, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LBM_HIP_OK;
}

int lbm_hip_get_wall_force(lbm_hip_ctx* c, double out[3]) {
    if (!c || !out) return fail(LBM_HIP_EINVAL, "null argument");
    if (c->t == 0) { out[0] = out[1] = out[2] = 0.0; return LBM_HIP_OK; }
    if (!c->wallf_valid) return fail(LBM_HIP_ESTATE, "the last step did not accumulate the wall force");
    HIP_TRY(hipMemcpyAsync(out, c->st->wallForce, 3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return LBM_HIP_OK;
}

int lbm_hip_diagnostics(lbm_hip_ctx* c, lbm_hip_diag* o) {
    if (!c || !o) return fail(LBM_HIP_EINVAL, "null argument");
    if (!c->macro_valid) return fail(LBM_HIP_ESTATE, "the last step did not write the macroscopic fields");
    HIP_TRY(launch_diag(c->flag, c->rho, c->ux, c->uy, c->uz, c->prec == LBM_HIP_FP64, c->N, c->diag_part, c->nb256,
                        c->diag_out, c->stream));
    double h[5];
    HIP_TRY(hipMemcpyAsync(h, c->diag_out, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    o->count_fluid = (int64_t)h[0];
    o->mean_density = h[0] > 0 ? h[1] / h[0] : 0.0;
    o->max_speed = h[2];
    o->sum_ux = h[3];
    o->unstable_nodes = (int64_t)h[4];
    return LBM_HIP_OK;
}

int lbm_hip_download(lbm_hip_ctx* c, double* f, double* rho, double* ux, double* uy, double* uz, uint8_t* flag) {
    if (!c) return fail(LBM_HIP_EINVAL, "null context");
    if (int rc = flush_links(c)) return rc;
    const size_t N = c->N;
    if ((rho || ux || uy || uz) && !c->macro_valid)
        return fail(LBM_HIP_ESTATE, "the last step did not write the macroscopic fields");
    double* outs[4] = {rho, ux, uy, uz};
    void* srcs[4] = {c->rho, c->ux, c->uy, c->uz};
    const bool macro_double = c->prec == LBM_HIP_FP64;
    if (f || ((rho || ux || uy || uz) && !macro_double)) {
        std::lock_guard<std::mutex> lock(g_stage_mu);
        double* stage = nullptr;
        if (int rc = staging(c->device, &stage)) return rc;
        hipError_t e = hipSuccess;
        // chunks of STAGE_NODES nodes; stream order keeps a chunk's copies
        // ahead of the kernel that refills the staging array
        if (f) {
            const int shifted = layout_shifted(c) ? 1 : 0;
            for (size_t n0 = 0; n0 < N && e == hipSuccess; n0 += STAGE_NODES) {
                const uint32_t cnt = (uint32_t)std::min<size_t>(STAGE_NODES, N - n0);
                dispatch(c->prec, [&](auto* tag) -> int {
                    using S = std::remove_pointer_t<decltype(tag)>;
                    e = launch_gather<S>((const S*)c->f, c->flag, stage, c->NX, c->NY, c->NZ, (uint32_t)n0, cnt, c->SS,
                                         c->PX, shifted, c->st, (int)(c->t & 1), c->stream);
                    return 0;
                });
                for (int i = 0; i < Q && e == hipSuccess; ++i)
                    e = hipMemcpyAsync(f + (size_t)i * N + n0, stage + (size_t)i * cnt, sizeof(double) * cnt,
                                       hipMemcpyDeviceToHost, c->stream);
            }
        }
        if (!macro_double)
            for (int k = 0; k < 4 && e == hipSuccess; ++k) {
                if (!outs[k]) continue;
                for (size_t n0 = 0; n0 < N && e == hipSuccess; n0 += STAGE_NODES * Q) {
                    const uint32_t cnt = (uint32_t)std::min<size_t>(STAGE_NODES * Q, N - n0);
                    e = launch_macro_to_double((const float*)srcs[k] + n0, 0, stage, cnt, c->stream);
                    if (e == hipSuccess)
                        e = hipMemcpyAsync(outs[k] + n0, stage, sizeof(double) * cnt, hipMemcpyDeviceToHost, c->stream);
                }
            }
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return fail(LBM_HIP_EDEVICE, std::string("download: ") + hipGetErrorString(e));
    }
    if (macro_double)
        for (int k = 0; k < 4; ++k)
            if (outs[k]) HIP_TRY(hipMemcpyAsync(outs[k], srcs[k], sizeof(double) * N, hipMemcpyDeviceToHost, c->stream));
    if (flag) HIP_TRY(hipMemcpyAsync(flag, c->flag, N, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (flag)
        for (size_t n = 0; n < N; ++n) flag[n] &= TYPE_MASK;
    return check_device_error(c);
}

int lbm_hip_upload_populations(lbm_hip_ctx* c, const double* f) {
    if (!c || !f) return fail(LBM_HIP_EINVAL, "null argument");
    if (int rc = flush_links(c)) return rc;
    const size_t N = c->N;
    std::lock_guard<std::mutex> lock(g_stage_mu);
    double* stage = nullptr;
    if (int rc = staging(c->device, &stage)) return rc;
    const int shifted = layout_shifted(c) ? 1 : 0;
    hipError_t e = hipSuccess;
    for (size_t n0 = 0; n0 < N && e == hipSuccess; n0 += STAGE_NODES) {
        const uint32_t cnt = (uint32_t)std::min<size_t>(STAGE_NODES, N - n0);
        for (int i = 0; i < Q && e == hipSuccess; ++i)
            e = hipMemcpyAsync(stage + (size_t)i * cnt, f + (size_t)i * N + n0, sizeof(double) * cnt,
                               hipMemcpyHostToDevice, c->stream);
        if (e == hipSuccess)
            dispatch(c->prec, [&](auto* tag) -> int {
                using S = std::remove_pointer_t<decltype(tag)>;
                e = launch_scatter<S>((S*)c->f, c->flag, stage, c->NX, c->NY, c->NZ, (uint32_t)n0, cnt, c->SS, c->PX,
                                      shifted, c->stream);
                return 0;
            });
    }
    const double zero = 0.0;
    if (e == hipSuccess) e = hipMemcpyAsync(&c->st->shift[c->t & 1], &zero, sizeof zero, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) return fail(LBM_HIP_EDEVICE, std::string("upload: ") + hipGetErrorString(e));
    c->mass_valid = false;
    return LBM_HIP_OK;
}

int lbm_hip_get_info(lbm_hip_ctx* c, lbm_hip_info* o) {
    if (!c || !o) return fail(LBM_HIP_EINVAL, "null argument");
    o->nx = c->NX; o->ny = c->NY; o->nz = c->NZ; o->precision = c->prec;
    o->n = c->N;
    o->steps_done = c->t;
    o->tau = c->p.tau; o->om_plus = c->omPlus; o->om_minus = c->omMinus; o->nu = c->nu;
    o->bytes_per_lup = 2.0 * Q * (double)c->ssize;
    o->device_bytes = c->device_bytes;
    o->has_squirmer = c->has_squirmer; o->walls_y = c->p.walls_y; o->walls_z = c->p.walls_z;
    long long nf = 0;
    HIP_TRY(hipMemcpyAsync(&nf, &c->st->nfluid, sizeof nf, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    o->count_fluid = nf;
    o->row_pitch = (int)c->PX;
    o->tile_width = c->tx_shift ? 1 << c->tx_shift : 0;
    o->xcd_remap = c->xcd_remap;
    o->occupancy_even = c->occ_cap[0];
    o->occupancy_odd = c->occ_cap[1];
    o->odd_row_waves = c->odd_rows ? c->row_waves : 0;
    hipError_t e = hipSuccess;
    dispatch(c->prec, [&](auto* tag) -> int {
        using S = std::remove_pointer_t<decltype(tag)>;
        for (int odd = 0; odd < 2 && e == hipSuccess; ++odd)
            e = step_blocks_per_cu<S>(odd != 0, c->occ_lds[odd], odd ? &o->resident_odd : &o->resident_even);
        return 0;
    });
    if (e != hipSuccess) return fail(LBM_HIP_EDEVICE, std::string("occupancy query: ") + hipGetErrorString(e));
    return LBM_HIP_OK;
}

int lbm_hip_particle_links(lbm_hip_ctx* c, uint32_t* node, uint8_t* dir, int64_t cap, int64_t* count) {
    if (!c || !count || cap < 0 || (cap > 0 && (!node || !dir))) return fail(LBM_HIP_EINVAL, "invalid argument");
    *count = 0;
    if (!c->has_squirmer) return LBM_HIP_OK;
    int rc;
    if ((rc = flush_links(c)) || (rc = ensure_link_list(c))) return rc;
    unsigned nl = 0;
    HIP_TRY(hipMemcpyAsync(&nl, &c->st->nlinks, sizeof nl, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    std::vector<uint2> list(nl);
    std::vector<uint8_t> fl(c->N);
    if (nl) HIP_TRY(hipMemcpyAsync(list.data(), c->links, sizeof(uint2) * nl, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(fl.data(), c->flag, c->N, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    // live entries whose link the flags still hold (after a remap the list
    // also holds the links it removed, until the next link pass marks them)
    const uint32_t NX = (uint32_t)c->NX, NY = (uint32_t)c->NY, NZ = (uint32_t)c->NZ;
    int64_t k = 0;
    for (const uint2& e : list) {
        if (e.y == LINK_DEAD || e.y >= (uint32_t)Q || e.x >= c->N) continue;
        const uint32_t x = e.x % NX, y = (e.x / NX) % NY, z = e.x / (NX * NY);
        const uint32_t xn = (x + NX + CX[e.y]) % NX, yn = (y + NY + CY[e.y]) % NY, zn = (z + NZ + CZ[e.y]) % NZ;
        if ((fl[e.x] & TYPE_MASK) != FLUID || (fl[xn + NX * (yn + NY * zn)] & TYPE_MASK) != PARTICLE) continue;
        if (k < cap) { node[k] = e.x; dir[k] = (uint8_t)e.y; }
        ++k;
    }
    *count = k;
    return check_device_error(c);
}

int lbm_hip_time_run(lbm_hip_ctx* c, int steps, int mass_every, double* total_ms, double* step_kernel_ms) {
    if (!c) return fail(LBM_HIP_EINVAL, "null context");
    if (steps <= 0) return fail(LBM_HIP_EINVAL, "steps must be > 0");
    const int ksteps = steps < 64 ? steps : 64;
    const size_t need = 2 * (size_t)ksteps + 2;
    while (c->events.size() < need) {
        hipEvent_t e;
        HIP_TRY(hipEventCreate(&e));
        c->events.push_back(e);
    }
    hipEvent_t* ev = c->events.data();
    int rc;
    // the graphs are built outside the timed region
    if ((rc = ensure_graph(c, mass_every, steps))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    // pass 1: `steps` full driver-loop steps (graph replay)
    HIP_TRY(hipEventRecord(ev[0], c->stream));
    if ((rc = run_loop(c, steps, mass_every, 0, true))) return rc;
    HIP_TRY(hipEventRecord(ev[1], c->stream));
    HIP_TRY(hipEventSynchronize(ev[1]));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, ev[0], ev[1]));
    // pass 2: the collide-stream update alone (bulk + even-step boundary
    // kernel), bracketed by events on its stream (the driver loop around it
    // runs as usual)
    double ksum = 0.0;
    for (int s = 0; s < ksteps; ++s) {
        const long long tnext = c->t + 1;
        const bool mass = c->has_squirmer && mass_every > 0 && tnext % mass_every == 0;
        c->kev[0] = ev[2 + 2 * s];
        c->kev[1] = ev[3 + 2 * s];
        rc = lbm_hip_step(c, mass ? LBM_STEP_MASS : 0);
        c->kev[0] = c->kev[1] = nullptr;
        if (rc) return rc;
        if (c->has_squirmer) {
            if ((rc = lbm_hip_advance_squirmer(c))) return rc;
            if (mass && (rc = lbm_hip_correct_mass(c))) return rc;
        }
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    for (int s = 0; s < ksteps; ++s) {
        float k = 0.f;
        HIP_TRY(hipEventElapsedTime(&k, ev[2 + 2 * s], ev[3 + 2 * s]));
        ksum += k;
    }
    if (total_ms) *total_ms = ms;
    if (step_kernel_ms) *step_kernel_ms = ksum * (double)steps / ksteps;
    return check_device_error(c);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Saved device state (autotuning runs candidates from the same state).
// ---------------------------------------------------------------------------
namespace {

constexpr int SNAP_PARTS = 8;

struct Snapshot {
    void* buf[SNAP_PARTS] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    size_t bytes[SNAP_PARTS] = {0, 0, 0, 0, 0, 0, 0, 0};
    long long t = 0;
    bool macro_valid = false, wallf_valid = false, mass_valid = false;
    int last_parts = 0, last_bparts = 0;
    int remaps_since_step = 0;
    int list_state = 0;
};

// Device arrays of the state: populations, flags, moments, DevState and the
// link list (whose order the later link sums follow, so it is copied rather
// than rebuilt).
void snapshot_parts(lbm_hip_ctx* c, void* (&src)[SNAP_PARTS], size_t (&bytes)[SNAP_PARTS]) {
    src[0] = c->f; bytes[0] = (size_t)Q * c->SS * c->ssize;
    src[1] = c->flag; bytes[1] = c->N;
    src[2] = c->rho; src[3] = c->ux; src[4] = c->uy; src[5] = c->uz;
    for (int k = 2; k < 6; ++k) bytes[k] = (size_t)c->N * c->msize;
    src[6] = c->st; bytes[6] = sizeof(DevState);
    src[7] = c->links; bytes[7] = c->links ? sizeof(uint2) * c->link_cap : 0;
}

void snapshot_free(Snapshot& s) {
    for (void*& q : s.buf)
        if (q) { hipFree(q); q = nullptr; }
}

int snapshot_take(lbm_hip_ctx* c, Snapshot& s, hipStream_t stream) {
    if (int rc = flush_links(c)) return rc;
    void* src[SNAP_PARTS];
    snapshot_parts(c, src, s.bytes);
    for (int k = 0; k < SNAP_PARTS; ++k) {
        if (s.bytes[k] == 0) continue;
        if (hipMalloc(&s.buf[k], s.bytes[k]) != hipSuccess) {
            s.buf[k] = nullptr;
            snapshot_free(s);
            return fail(LBM_HIP_ENOMEM, "autotune: cannot allocate the state copy");
        }
        HIP_TRY(hipMemcpyAsync(s.buf[k], src[k], s.bytes[k], hipMemcpyDeviceToDevice, stream));
    }
    s.t = c->t;
    s.macro_valid = c->macro_valid; s.wallf_valid = c->wallf_valid; s.mass_valid = c->mass_valid;
    s.last_parts = c->last_parts; s.last_bparts = c->last_bparts;
    s.remaps_since_step = c->remaps_since_step;
    s.list_state = c->list_state;
    return LBM_HIP_OK;
}

int snapshot_restore(lbm_hip_ctx* c, const Snapshot& s, hipStream_t stream) {
    void* dst[SNAP_PARTS];
    size_t bytes[SNAP_PARTS];
    snapshot_parts(c, dst, bytes);
    for (int k = 0; k < SNAP_PARTS; ++k)
        if (bytes[k]) HIP_TRY(hipMemcpyAsync(dst[k], s.buf[k], bytes[k], hipMemcpyDeviceToDevice, stream));
    c->t = s.t;
    c->macro_valid = s.macro_valid; c->wallf_valid = s.wallf_valid;
    // The per-block mass partials of a pending mass step are not part of the
    // copy (the trials overwrite them, possibly with another tile shape):
    // a following correct_mass sums the populations instead.
    c->mass_valid = false;
    c->last_parts = s.last_parts; c->last_bparts = s.last_bparts;
    c->remaps_since_step = s.remaps_since_step;
    c->list_state = s.list_state;
    c->links_pending = false;
    return LBM_HIP_OK;
}

int tile_shift_of(int width) {
    return width == 0 ? 0 : (width == 64 ? 6 : (width == 128 ? 7 : (width == 256 ? 8 : -1)));
}

const int kTuneWidth[LBM_HIP_TUNE_CANDIDATES] = {64, 128, 256, 64, 128, 256, 64, 128, 256, 0, 0};
const int kTuneXcd[LBM_HIP_TUNE_CANDIDATES] = {1, 1, 1, 0, 0, 0, 2, 2, 2, 1, 0};

int graph_cycle_of(bool squirmer, int mass_every) {
    return (squirmer && mass_every > 0) ? ((mass_every % 2 == 0) ? mass_every : 2 * mass_every) : 2;
}

}  // namespace

extern "C" {

int lbm_hip_set_tiling(lbm_hip_ctx* c, int tile_width, int xcd_remap) {
    if (!c) return fail(LBM_HIP_EINVAL, "null context");
    const int ts = tile_shift_of(tile_width);
    if (ts < 0) return fail(LBM_HIP_EINVAL, "tile_width must be 0 (row runs), 64, 128 or 256");
    if (xcd_remap < 0 || xcd_remap > 2) return fail(LBM_HIP_EINVAL, "xcd_remap must be 0, 1 or 2");
    if (ts == 0 && xcd_remap == 2) return fail(LBM_HIP_EINVAL, "row-run tiles (tile_width 0) take xcd_remap 0 or 1");
    set_tile_shape(c, ts);
    c->xcd_remap = xcd_remap;
    c->mass_valid = false;  // partials of a pending mass step follow the old tiles
    return LBM_HIP_OK;
}

int lbm_hip_set_occupancy(lbm_hip_ctx* c, int even_blocks, int odd_blocks) {
    if (!c) return fail(LBM_HIP_EINVAL, "null context");
    return set_occupancy(c, even_blocks, odd_blocks);
}

int lbm_hip_autotune(lbm_hip_ctx* c, int trial_steps, int mass_every, lbm_hip_tuning* out) {
    if (!c) return fail(LBM_HIP_EINVAL, "null context");
    if (trial_steps < 1) return fail(LBM_HIP_EINVAL, "trial_steps must be >= 1");
    HIP_TRY(hipStreamSynchronize(c->stream));
    Snapshot snap;
    int rc = snapshot_take(c, snap, c->stream);
    if (rc) return rc;
    while (c->events.size() < 2) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) { snapshot_free(snap); return fail(LBM_HIP_EDEVICE, "hipEventCreate"); }
        c->events.push_back(e);
    }
    lbm_hip_tuning res{};
    int best = -1;
    const int warm = graph_cycle_of(c->has_squirmer, mass_every);
    for (int k = 0; k < LBM_HIP_TUNE_CANDIDATES && rc == LBM_HIP_OK; ++k) {
        res.cand_tile_width[k] = kTuneWidth[k];
        res.cand_xcd_remap[k] = kTuneXcd[k];
        if ((rc = lbm_hip_set_tiling(c, kTuneWidth[k], kTuneXcd[k]))) break;
        if ((rc = snapshot_restore(c, snap, c->stream))) break;
        if ((rc = run_loop(c, warm, mass_every, 0, true))) break;
        float ms = 0.f;
        if (hipEventRecord(c->events[0], c->stream) != hipSuccess) { rc = fail(LBM_HIP_EDEVICE, "event"); break; }
        if ((rc = run_loop(c, trial_steps, mass_every, 0, true))) break;
        if (hipEventRecord(c->events[1], c->stream) != hipSuccess || hipEventSynchronize(c->events[1]) != hipSuccess ||
            hipEventElapsedTime(&ms, c->events[0], c->events[1]) != hipSuccess) {
            rc = fail(LBM_HIP_EDEVICE, "autotune timing");
            break;
        }
        res.cand_ms_per_step[k] = ms / trial_steps;
        if (best < 0 || res.cand_ms_per_step[k] < res.cand_ms_per_step[best]) best = k;
    }
    if (rc == LBM_HIP_OK) rc = check_device_error(c);
    if (best >= 0) {
        const int r2 = lbm_hip_set_tiling(c, kTuneWidth[best], kTuneXcd[best]);
        if (rc == LBM_HIP_OK) rc = r2;
        res.tile_width = kTuneWidth[best];
        res.xcd_remap = kTuneXcd[best];
        res.ms_per_step = res.cand_ms_per_step[best];
    }
    const int r3 = snapshot_restore(c, snap, c->stream);
    if (rc == LBM_HIP_OK) rc = r3;
    hipStreamSynchronize(c->stream);
    snapshot_free(snap);
    if (rc == LBM_HIP_OK && out) *out = res;
    return rc;
}

}  // extern "C"


// ---------------------------------------------------------------------------
// Ensemble of lattices: one launch per phase for all members.
// ---------------------------------------------------------------------------
struct lbm_hip_batch {
    int device = 0, prec = 0, count = 0;
    hipStream_t stream = nullptr;
    std::vector<lbm_hip_ctx*> mem;
    void* desc = nullptr;             // BatchMember<S>[count]
    BatchFinalize* fin = nullptr;     // [count]
    uint32_t* tile_map = nullptr;     // member << 24 | tile, sum of G_max entries
    uint32_t* bmap = nullptr;         // member << 24 | boundary block
    uint32_t* lmap = nullptr;         // member << 24 | link block
    uint32_t* row_map = nullptr;      // member << 24 | row block (odd steps in row blocks)
    int Gtot = 0, Btot = 0, Ltot = 0, Rtot = 0;
    bool odd_rows = false;            // odd steps in row blocks (every member's rows whole wavefronts)
    int row_block_waves = 0;          // wavefronts per row block
    unsigned occ_lds_rows = 0;        // dynamic LDS of the row blocks for the odd cap
    bool any_squirmer = false;
    bool any_force = false;           // some member has a body force (all run the forcing terms)
    long long t = 0;
    hipGraphExec_t graph = nullptr;
    int graph_cycle = 0, graph_mass_every = -1;
    size_t device_bytes = 0;
    hipEvent_t ev[2] = {nullptr, nullptr};
};

namespace {

void free_batch(lbm_hip_batch* b) {
    if (!b) return;
    hipSetDevice(b->device);
    if (b->stream) hipStreamSynchronize(b->stream);
    for (void* q : {b->desc, (void*)b->fin, (void*)b->tile_map, (void*)b->bmap, (void*)b->lmap, (void*)b->row_map})
        if (q) hipFree(q);
    if (b->graph) hipGraphExecDestroy(b->graph);
    for (hipEvent_t e : {b->ev[0], b->ev[1]})
        if (e) hipEventDestroy(e);
    for (lbm_hip_ctx* c : b->mem) {
        if (!c) continue;
        c->owned_by_batch = false;
        free_ctx(c);
    }
    if (b->stream) hipStreamDestroy(b->stream);
    delete b;
}

// (Re)build the member argument table and the block maps for the members'
// current tile shapes.
int batch_tables(lbm_hip_batch* b) {
    std::vector<uint32_t> tmap, bm, lm, rmap;
    if (b->odd_rows) {
        // wavefronts per row block: the smallest power of two that holds the
        // longest row (whole blocks fill a CU's 32 wavefront slots), each
        // member's rows packed floor(block / row) to a block
        int lo = 1;
        for (lbm_hip_ctx* c : b->mem) lo = std::max(lo, c->row_waves);
        b->row_block_waves = 1;
        while (b->row_block_waves < lo) b->row_block_waves *= 2;
        for (int m = 0; m < b->count; ++m) {
            lbm_hip_ctx* c = b->mem[(size_t)m];
            const int rpb = b->row_block_waves / c->row_waves;
            c->batch_row_blocks = (c->NY * c->NZ + rpb - 1) / rpb;
            for (int k = 0; k < c->batch_row_blocks; ++k) rmap.push_back((uint32_t)m << 24 | (uint32_t)k);
        }
        if (int rc = row_occupancy_lds(b->device, b->mem[0]->occ_cap[1], b->row_block_waves, &b->occ_lds_rows))
            return rc;
    }
    std::vector<BatchFinalize> fin((size_t)b->count);
    const bool row_order = b->mem[0]->xcd_remap == 2;
    std::vector<uint32_t> queue[8];  // tile order 2: whole tile rows per XCD
    size_t qlen[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int m = 0; m < b->count; ++m) {
        lbm_hip_ctx* c = b->mem[(size_t)m];
        const uint32_t tag = (uint32_t)m << 24;
        if (row_order) {
            // each tile row to the XCD queue with the fewest tiles so far
            for (int r = 0; r < c->G / c->gx; ++r) {
                const int q = (int)(std::min_element(qlen, qlen + 8) - qlen);
                for (int k = 0; k < c->gx; ++k) queue[q].push_back(tag | (uint32_t)(r * c->gx + k));
                qlen[q] += (size_t)c->gx;
            }
        } else {
            for (int k = 0; k < c->G; ++k) tmap.push_back(tag | (uint32_t)k);
        }
        for (int k = 0; k < c->bblocks; ++k) bm.push_back(tag | (uint32_t)k);
        if (c->has_squirmer)
            for (int k = 0; k < c->cpl_blocks; ++k) lm.push_back(tag | (uint32_t)k);
        fin[(size_t)m] = BatchFinalize{c->mass_part, c->st, c->G, c->bblocks, c->has_squirmer ? 1 : 0,
                                       b->odd_rows ? c->batch_row_blocks : c->G};
    }
    if (row_order) {
        // block i runs on XCD i mod 8: interleave the queues
        for (size_t j = 0, left = 1; left; ++j) {
            left = 0;
            for (int q = 0; q < 8; ++q)
                if (j < queue[q].size()) { tmap.push_back(queue[q][j]); ++left; }
        }
    }
    b->Gtot = (int)tmap.size(); b->Btot = (int)bm.size(); b->Ltot = (int)lm.size(); b->Rtot = (int)rmap.size();
    if (!rmap.empty())
        HIP_TRY(hipMemcpyAsync(b->row_map, rmap.data(), rmap.size() * sizeof(uint32_t), hipMemcpyHostToDevice,
                               b->stream));
    HIP_TRY(hipMemcpyAsync(b->tile_map, tmap.data(), tmap.size() * sizeof(uint32_t), hipMemcpyHostToDevice, b->stream));
    if (!bm.empty())
        HIP_TRY(hipMemcpyAsync(b->bmap, bm.data(), bm.size() * sizeof(uint32_t), hipMemcpyHostToDevice, b->stream));
    if (!lm.empty())
        HIP_TRY(hipMemcpyAsync(b->lmap, lm.data(), lm.size() * sizeof(uint32_t), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(b->fin, fin.data(), fin.size() * sizeof(BatchFinalize), hipMemcpyHostToDevice, b->stream));
    int rc = LBM_HIP_OK;
    dispatch(b->prec, [&](auto* tag) -> int {
        using S = std::remove_pointer_t<decltype(tag)>;
        std::vector<BatchMember<S>> d((size_t)b->count);
        for (int m = 0; m < b->count; ++m) {
            lbm_hip_ctx* c = b->mem[(size_t)m];
            BatchMember<S>& x = d[(size_t)m];
            x.A = step_args<S>(c);
            if (b->odd_rows) x.A.rows_per_block = b->row_block_waves / c->row_waves;
            x.K = c->K;
            x.B = c->bargs;
            x.L = link_args<S>(c);
            x.R = remap_args<S>(c);
            x.lblocks = c->cpl_blocks;
            x.has_squirmer = c->has_squirmer ? 1 : 0;
        }
        if (hipMemcpyAsync(b->desc, d.data(), d.size() * sizeof(BatchMember<S>), hipMemcpyHostToDevice, b->stream) !=
            hipSuccess)
            rc = fail(LBM_HIP_EDEVICE, "batch: upload of member arguments failed");
        // the host vector dies here: wait for the copy
        if (rc == LBM_HIP_OK && hipStreamSynchronize(b->stream) != hipSuccess)
            rc = fail(LBM_HIP_EDEVICE, "batch: synchronize failed");
        return 0;
    });
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (b->graph) { hipGraphExecDestroy(b->graph); b->graph = nullptr; }
    return LBM_HIP_OK;
}

// One driver-loop step of every member (main.cpp:60-64).
int batch_step(lbm_hip_batch* b, int mass_every, int extra_flags) {
    const long long tnext = b->t + 1;
    const bool mass = b->any_squirmer && mass_every > 0 && tnext % mass_every == 0;
    int mode = mass ? STEP_MASS : 0;
    if (extra_flags & LBM_STEP_MACRO) mode |= STEP_MACRO;
    if (b->any_force) mode |= STEP_BODY_FORCE;
    const bool odd = (b->t & 1) != 0;
    hipError_t e = hipSuccess;
    dispatch(b->prec, [&](auto* tag) -> int {
        using S = std::remove_pointer_t<decltype(tag)>;
        const auto* d = static_cast<const BatchMember<S>*>(b->desc);
        if (odd && b->odd_rows)
            e = launch_batch_odd_rows<S>(d, b->row_map, b->Rtot, 64 * b->row_block_waves, b->t, mode, b->occ_lds_rows,
                                         b->stream);
        else
            e = launch_batch_step<S>(d, b->tile_map, b->Gtot, b->t, b->mem[0]->xcd_remap == 1, odd, mode,
                                     b->mem[0]->occ_lds[odd], b->stream);
        if (e == hipSuccess && !odd) e = launch_batch_boundary<S>(d, b->bmap, b->Btot, b->t, mode, b->stream);
        if (e == hipSuccess && (extra_flags & LBM_STEP_WALL_FORCE))
            for (lbm_hip_ctx* c : b->mem)
                if (e == hipSuccess)
                    e = launch_wall_force<S>((const S*)c->f, c->flag, wall_force_args(c), c->red_part, c->st, c->NX,
                                             c->NY, c->NZ, c->SS, c->PX, b->stream);
        // link pass, rigid-body update, remap and link-list upkeep of every member
        if (e == hipSuccess && b->any_squirmer)
            e = launch_batch_couple<S>(d, b->lmap, b->Ltot, b->count, tnext, b->stream);
        return 0;
    });
    if (e == hipSuccess && mass) e = launch_batch_finalize(b->fin, b->count, tnext, b->stream);
    if (e != hipSuccess) return fail(LBM_HIP_EDEVICE, std::string("batch step launch: ") + hipGetErrorString(e));
    b->t = tnext;
    for (lbm_hip_ctx* c : b->mem) {
        c->last_parts = odd && b->odd_rows ? c->batch_row_blocks : c->G;
        c->last_bparts = odd ? 0 : c->bblocks;
        c->t = tnext;
        c->macro_valid = (extra_flags & LBM_STEP_MACRO) != 0;
        c->wallf_valid = (extra_flags & LBM_STEP_WALL_FORCE) != 0;
        c->mass_valid = false;
        if (c->has_squirmer) c->list_state = LIST_UNCHECKED;
    }
    return LBM_HIP_OK;
}

int batch_graph(lbm_hip_batch* b, int mass_every) {
    if (b->graph && b->graph_mass_every == mass_every) return LBM_HIP_OK;
    if (b->graph) { hipGraphExecDestroy(b->graph); b->graph = nullptr; }
    const int cycle = graph_cycle_of(b->any_squirmer, mass_every);
    const long long t_save = b->t;
    std::vector<Snapshot> flags((size_t)b->count);  // host-side member fields only
    for (int m = 0; m < b->count; ++m) {
        lbm_hip_ctx* c = b->mem[(size_t)m];
        flags[(size_t)m].t = c->t; flags[(size_t)m].last_bparts = c->last_bparts;
        flags[(size_t)m].macro_valid = c->macro_valid; flags[(size_t)m].wallf_valid = c->wallf_valid;
        flags[(size_t)m].mass_valid = c->mass_valid;
        flags[(size_t)m].list_state = c->list_state;
    }
    b->t = 0;  // capture with a cycle-aligned step index (the kernels use its parity only)
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    int rc = LBM_HIP_OK;
    for (int s = 0; s < cycle && rc == LBM_HIP_OK; ++s) rc = batch_step(b, mass_every, 0);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(b->stream, &g);
    b->t = t_save;
    for (int m = 0; m < b->count; ++m) {
        lbm_hip_ctx* c = b->mem[(size_t)m];
        c->t = flags[(size_t)m].t; c->last_bparts = flags[(size_t)m].last_bparts;
        c->macro_valid = flags[(size_t)m].macro_valid; c->wallf_valid = flags[(size_t)m].wallf_valid;
        c->mass_valid = flags[(size_t)m].mass_valid;
        c->list_state = flags[(size_t)m].list_state;
    }
    if (rc) { if (g) hipGraphDestroy(g); return rc; }
    if (e != hipSuccess) return fail(LBM_HIP_EDEVICE, std::string("batch graph capture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) {
        b->graph = nullptr;
        return fail(LBM_HIP_EDEVICE, std::string("batch graph instantiate: ") + hipGetErrorString(e));
    }
    b->graph_cycle = cycle;
    b->graph_mass_every = mass_every;
    return LBM_HIP_OK;
}

int batch_loop(lbm_hip_batch* b, int steps, int mass_every, int last_flags) {
    if (steps < 0) return fail(LBM_HIP_EINVAL, "steps must be >= 0");
    for (lbm_hip_ctx* c : b->mem)
        if (c->t != b->t) return fail(LBM_HIP_EINVAL, "a batch member was advanced outside the batch");
    int rc, left = steps;
    // every member's link list valid before the batch's coupling launches
    // (a member's own remap call may have invalidated it)
    for (lbm_hip_ctx* c : b->mem)
        if (c->links_pending || c->list_state == LIST_DIRTY) {
            if ((rc = flush_links(c)) || (rc = ensure_link_list(c))) return rc;
            HIP_TRY(hipStreamSynchronize(c->stream));
        }
    if (steps > 1) {
        if ((rc = batch_graph(b, mass_every))) return rc;
        const int C = b->graph_cycle;
        while (left > 1 && (b->t % C) != 0) {
            if ((rc = batch_step(b, mass_every, 0))) return rc;
            --left;
        }
        const int reps = (left - (last_flags ? 1 : 0)) / C;  // the last step eager if it has flags
        for (int r = 0; r < reps; ++r) HIP_TRY(hipGraphLaunch(b->graph, b->stream));
        if (reps > 0) {
            b->t += (long long)reps * C;
            left -= reps * C;
            for (lbm_hip_ctx* c : b->mem) {
                c->t = b->t;
                c->macro_valid = c->wallf_valid = c->mass_valid = false;
                c->last_parts = b->odd_rows ? c->batch_row_blocks : c->G;
                c->last_bparts = 0;  // a cycle ends on an odd step
            }
        }
    }
    while (left > 0) {
        if ((rc = batch_step(b, mass_every, left == 1 ? last_flags : 0))) return rc;
        --left;
    }
    return LBM_HIP_OK;
}

int batch_check(lbm_hip_batch* b) {
    HIP_TRY(hipStreamSynchronize(b->stream));
    for (lbm_hip_ctx* c : b->mem) {
        const int rc = check_device_error(c);
        if (rc) return rc;
    }
    return LBM_HIP_OK;
}

}  // namespace

extern "C" {

int lbm_hip_batch_create(const lbm_hip_params* members, int count, lbm_hip_batch** out) {
    if (!out) return fail(LBM_HIP_EINVAL, "null output pointer");
    *out = nullptr;
    if (!members || count < 1) return fail(LBM_HIP_EINVAL, "a batch needs at least one member");
    if (count > 256) return fail(LBM_HIP_EINVAL, "a batch holds at most 256 members");
    for (int m = 0; m < count; ++m) {
        if (members[m].precision != members[0].precision)
            return fail(LBM_HIP_EINVAL, "batch members must share one precision");
        if (members[m].device != members[0].device) return fail(LBM_HIP_EINVAL, "batch members must share one device");
        const int rc = lbm_hip_validate(&members[m]);
        if (rc) return fail(rc, "member " + std::to_string(m) + ": " + g_err);
    }
    auto* b = new lbm_hip_batch();
    b->device = members[0].device;
    b->prec = members[0].precision;
    b->count = count;
    auto bail = [&](int code) {
        std::string msg = g_err;
        free_batch(b);
        g_err = msg;
        return code;
    };
    for (int m = 0; m < count; ++m) {
        lbm_hip_ctx* c = nullptr;
        const int rc = lbm_hip_create(&members[m], &c);
        if (rc) { g_err = "member " + std::to_string(m) + ": " + g_err; return bail(rc); }
        c->owned_by_batch = true;
        b->mem.push_back(c);
        b->device_bytes += c->device_bytes;
        b->any_squirmer = b->any_squirmer || c->has_squirmer;
        b->any_force = b->any_force || body_force(c);
        {
            // the ensemble's launches cover many lattices: measured on the
            // 64-run workload, 2-byte storage runs fastest at 6 blocks per CU
            // on both parities
            int cap[2];
            default_occupancy(c->prec, c->odd_rows != 0, cap);
            if (c->prec == LBM_HIP_FP16 || c->prec == LBM_HIP_BF16) cap[0] = cap[1] = 6;
            if (const int rc2 = set_occupancy(c, cap[0], cap[1])) return bail(rc2);
        }
        if (c->G_max >= (1 << 24) || c->bblocks >= (1 << 24) || c->cpl_blocks >= (1 << 24))
            return bail(fail(LBM_HIP_EINVAL, "member " + std::to_string(m) + " too large for a batch"));
    }
    if (hipSetDevice(b->device) != hipSuccess) return bail(fail(LBM_HIP_EDEVICE, "hipSetDevice failed"));
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(fail(LBM_HIP_EDEVICE, "hipStreamCreate failed"));
    for (hipEvent_t& e : b->ev)
        if (hipEventCreate(&e) != hipSuccess) return bail(fail(LBM_HIP_EDEVICE, "hipEventCreate failed"));
    // members finished their set-up on their own streams
    for (lbm_hip_ctx* c : b->mem)
        if (hipStreamSynchronize(c->stream) != hipSuccess) return bail(fail(LBM_HIP_EDEVICE, "member set-up failed"));
    size_t gmax = 0, bsum = 0, lsum = 0, rsum = 0;
    // Odd steps in row blocks only on request (LBM_HIP_ODD_ROWS = 1 or 2) and
    // where every member's rows are whole wavefronts: measured on the 64-run
    // ensemble (FP16, tuned tiles), row blocks of 8 / 12 / 16 wavefronts took
    // 3.42 / 3.40 / 3.59 ms per step against 3.27 ms with the odd step in tiles.
    {
        const char* e = std::getenv("LBM_HIP_ODD_ROWS");
        b->odd_rows = e && std::atoi(e) > 0;
    }
    for (lbm_hip_ctx* c : b->mem) {
        gmax += (size_t)c->G_max;
        bsum += (size_t)c->bblocks;
        lsum += c->has_squirmer ? (size_t)c->cpl_blocks : 0;
        rsum += (size_t)c->NY * (size_t)c->NZ;
        b->odd_rows = b->odd_rows && c->odd_rows != 0;
    }
    if (b->odd_rows)  // the row blocks run uncapped
        for (lbm_hip_ctx* c : b->mem)
            if (const int rc2 = set_occupancy(c, c->occ_cap[0], 0)) return bail(rc2);
    const size_t dsz = b->prec == LBM_HIP_FP64 ? sizeof(BatchMember<double>)
                       : b->prec == LBM_HIP_FP32 ? sizeof(BatchMember<float>)
                       : b->prec == LBM_HIP_FP16 ? sizeof(BatchMember<_Float16>)
                                                 : sizeof(BatchMember<bf16_t>);
    auto dmalloc = [&](void** q, size_t bytes) -> bool {
        if (hipMalloc(q, bytes ? bytes : 4) != hipSuccess) { *q = nullptr; return false; }
        b->device_bytes += bytes;
        return true;
    };
    if (!dmalloc(&b->desc, dsz * (size_t)count) || !dmalloc((void**)&b->fin, sizeof(BatchFinalize) * (size_t)count) ||
        !dmalloc((void**)&b->tile_map, sizeof(uint32_t) * gmax) || !dmalloc((void**)&b->bmap, sizeof(uint32_t) * bsum) ||
        !dmalloc((void**)&b->lmap, sizeof(uint32_t) * lsum) ||
        (b->odd_rows && !dmalloc((void**)&b->row_map, sizeof(uint32_t) * rsum)))
        return bail(fail(LBM_HIP_ENOMEM, "batch: device allocation failed"));
    const int rc = batch_tables(b);
    if (rc) return bail(rc);
    *out = b;
    return LBM_HIP_OK;
}

void lbm_hip_batch_destroy(lbm_hip_batch* b) { free_batch(b); }

int lbm_hip_batch_run(lbm_hip_batch* b, int steps, int mass_every, int last_flags) {
    if (!b) return fail(LBM_HIP_EINVAL, "null batch");
    return batch_loop(b, steps, mass_every, last_flags);
}

int lbm_hip_batch_synchronize(lbm_hip_batch* b) {
    if (!b) return fail(LBM_HIP_EINVAL, "null batch");
    return batch_check(b);
}

int lbm_hip_batch_member(lbm_hip_batch* b, int i, lbm_hip_ctx** out) {
    if (!b || !out) return fail(LBM_HIP_EINVAL, "null argument");
    if (i < 0 || i >= b->count) return fail(LBM_HIP_EINVAL, "member index out of range");
    HIP_TRY(hipStreamSynchronize(b->stream));
    *out = b->mem[(size_t)i];
    return LBM_HIP_OK;
}

int lbm_hip_batch_info_get(lbm_hip_batch* b, lbm_hip_batch_info* o) {
    if (!b || !o) return fail(LBM_HIP_EINVAL, "null argument");
    std::memset(o, 0, sizeof *o);
    o->count = b->count;
    o->precision = b->prec;
    o->steps_done = b->t;
    for (lbm_hip_ctx* c : b->mem) o->total_nodes += (long long)c->N;
    o->total_tiles = b->Gtot;
    o->boundary_blocks = b->Btot;
    o->link_blocks = b->Ltot;
    o->tile_width = b->mem[0]->tx_shift ? 1 << b->mem[0]->tx_shift : 0;
    o->xcd_remap = b->mem[0]->xcd_remap;
    o->device_bytes = b->device_bytes;
    o->odd_row_blocks = b->odd_rows ? b->Rtot : 0;
    o->odd_row_block_waves = b->odd_rows ? b->row_block_waves : 0;
    return LBM_HIP_OK;
}

int lbm_hip_batch_set_tiling(lbm_hip_batch* b, int tile_width, int xcd_remap) {
    if (!b) return fail(LBM_HIP_EINVAL, "null batch");
    int rc = batch_check(b);
    if (rc) return rc;
    for (lbm_hip_ctx* c : b->mem)
        if ((rc = lbm_hip_set_tiling(c, tile_width, xcd_remap))) return rc;
    return batch_tables(b);
}

int lbm_hip_batch_set_occupancy(lbm_hip_batch* b, int even_blocks, int odd_blocks) {
    if (!b) return fail(LBM_HIP_EINVAL, "null batch");
    int rc = batch_check(b);
    if (rc) return rc;
    for (lbm_hip_ctx* c : b->mem)
        if ((rc = set_occupancy(c, even_blocks, odd_blocks))) return rc;
    if (b->odd_rows && (rc = row_occupancy_lds(b->device, odd_blocks, b->row_block_waves, &b->occ_lds_rows))) return rc;
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (b->graph) { hipGraphExecDestroy(b->graph); b->graph = nullptr; }
    return LBM_HIP_OK;
}

int lbm_hip_batch_autotune(lbm_hip_batch* b, int trial_steps, int mass_every, lbm_hip_tuning* out) {
    if (!b) return fail(LBM_HIP_EINVAL, "null batch");
    if (trial_steps < 1) return fail(LBM_HIP_EINVAL, "trial_steps must be >= 1");
    int rc = batch_check(b);
    if (rc) return rc;
    std::vector<Snapshot> snaps((size_t)b->count);
    for (int m = 0; m < b->count && rc == LBM_HIP_OK; ++m) rc = snapshot_take(b->mem[(size_t)m], snaps[(size_t)m], b->stream);
    const long long t_save = b->t;
    auto restore_all = [&]() -> int {
        for (int m = 0; m < b->count; ++m) {
            const int r = snapshot_restore(b->mem[(size_t)m], snaps[(size_t)m], b->stream);
            if (r) return r;
        }
        b->t = t_save;
        return LBM_HIP_OK;
    };
    auto set_all = [&](int width, int xcd) -> int {
        for (lbm_hip_ctx* c : b->mem) {
            const int r = lbm_hip_set_tiling(c, width, xcd);
            if (r) return r;
        }
        return batch_tables(b);
    };
    lbm_hip_tuning res{};
    int best = -1;
    const int warm = graph_cycle_of(b->any_squirmer, mass_every);
    for (int k = 0; k < LBM_HIP_TUNE_CANDIDATES && rc == LBM_HIP_OK; ++k) {
        res.cand_tile_width[k] = kTuneWidth[k];
        res.cand_xcd_remap[k] = kTuneXcd[k];
        if ((rc = set_all(kTuneWidth[k], kTuneXcd[k]))) break;
        if ((rc = restore_all())) break;
        if ((rc = batch_loop(b, warm, mass_every, 0))) break;
        float ms = 0.f;
        if (hipEventRecord(b->ev[0], b->stream) != hipSuccess) { rc = fail(LBM_HIP_EDEVICE, "event"); break; }
        if ((rc = batch_loop(b, trial_steps, mass_every, 0))) break;
        if (hipEventRecord(b->ev[1], b->stream) != hipSuccess || hipEventSynchronize(b->ev[1]) != hipSuccess ||
            hipEventElapsedTime(&ms, b->ev[0], b->ev[1]) != hipSuccess) {
            rc = fail(LBM_HIP_EDEVICE, "batch autotune timing");
            break;
        }
        res.cand_ms_per_step[k] = ms / trial_steps;
        if (best < 0 || res.cand_ms_per_step[k] < res.cand_ms_per_step[best]) best = k;
    }
    if (rc == LBM_HIP_OK) rc = batch_check(b);
    if (best >= 0) {
        const int r2 = set_all(kTuneWidth[best], kTuneXcd[best]);
        if (rc == LBM_HIP_OK) rc = r2;
        res.tile_width = kTuneWidth[best];
        res.xcd_remap = kTuneXcd[best];
        res.ms_per_step = res.cand_ms_per_step[best];
    }
    const int r3 = restore_all();
    if (rc == LBM_HIP_OK) rc = r3;
    hipStreamSynchronize(b->stream);
    for (Snapshot& s : snaps) snapshot_free(s);
    if (rc == LBM_HIP_OK && out) *out = res;
    return rc;
}

int lbm_hip_batch_time_run(lbm_hip_batch* b, int steps, int mass_every, double* total_ms) {
    if (!b) return fail(LBM_HIP_EINVAL, "null batch");
    if (steps <= 0) return fail(LBM_HIP_EINVAL, "steps must be > 0");
    int rc = batch_graph(b, mass_every);
    if (rc) return rc;
    HIP_TRY(hipStreamSynchronize(b->stream));
    HIP_TRY(hipEventRecord(b->ev[0], b->stream));
    if ((rc = batch_loop(b, steps, mass_every, 0))) return rc;
    HIP_TRY(hipEventRecord(b->ev[1], b->stream));
    HIP_TRY(hipEventSynchronize(b->ev[1]));
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
    if (total_ms) *total_ms = ms;
    return batch_check(b);
}

int lbm_hip_copy_bandwidth(int device, size_t bytes, int reps, double* gbs) {
    return lbm_hip_copy_bandwidth_parts(device, bytes, 1, reps, gbs);
}

int lbm_hip_copy_bandwidth_parts(int device, size_t bytes, int parts, int reps, double* gbs) {
    if (!gbs || reps <= 0 || parts < 1 || parts > 4096 || bytes < 512 * (size_t)parts)
        return fail(LBM_HIP_EINVAL, "copy_bandwidth: bytes >= 512 * parts, 1 <= parts <= 4096, reps > 0, gbs");
    const size_t n16 = bytes / (size_t)parts / 2 / 16;  // 16-byte words per part and direction
    HIP_TRY(hipSetDevice(device));
    std::vector<void*> ptr(2 * (size_t)parts, nullptr);  // sources, then destinations
    void** table = nullptr;
    hipStream_t s = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    float ms = 0.f;
    hipError_t e = hipSuccess;
    for (size_t k = 0; k < ptr.size() && e == hipSuccess; ++k) e = hipMalloc(&ptr[k], n16 * 16);
    if (e == hipSuccess) e = hipMalloc((void**)&table, sizeof(void*) * ptr.size());
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipMemcpyAsync(table, ptr.data(), sizeof(void*) * ptr.size(), hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipEventCreate(&ev[0]);
    if (e == hipSuccess) e = hipEventCreate(&ev[1]);
    for (int k = 0; k < parts && e == hipSuccess; ++k) e = hipMemsetAsync(ptr[(size_t)k], 0, n16 * 16, s);
    auto copy = [&]() { return launch_copy_parts(table, table + parts, parts, n16, s); };
    for (int r = 0; r < 2 && e == hipSuccess; ++r) e = copy();
    if (e == hipSuccess) e = hipEventRecord(ev[0], s);
    for (int r = 0; r < reps && e == hipSuccess; ++r) e = copy();
    if (e == hipSuccess) e = hipEventRecord(ev[1], s);
    if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
    if (s) hipStreamSynchronize(s);
    for (hipEvent_t v : ev)
        if (v) hipEventDestroy(v);
    if (s) hipStreamDestroy(s);
    for (void* q : ptr)
        if (q) hipFree(q);
    if (table) hipFree(table);
    if (e != hipSuccess) return fail(LBM_HIP_EDEVICE, std::string("copy_bandwidth: ") + hipGetErrorString(e));
    *gbs = 2.0 * (double)(n16 * 16) * parts * reps / (ms * 1e-3) / 1e9;
    return LBM_HIP_OK;
}

}  // extern "C"
