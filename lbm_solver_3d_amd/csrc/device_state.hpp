// device_state.hpp -- device-resident squirmer state, coupling accumulators
// and the argument blocks of the kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "fast_div.hpp"
#include "lattice.hpp"

namespace lbmhip {

// Data that no block writes while a kernel runs (the per-member argument
// tables of an ensemble) is read through the constant address space, so
// the loads are scalar and the values stay in scalar registers, as kernel
// arguments do, instead of vector registers (which cost the collide-stream
// kernel a third of its occupancy).
template <class T>
__device__ __forceinline__ const __attribute__((address_space(4))) T* as_constant(const T* p) {
    return (const __attribute__((address_space(4))) T*)p;
}

template <class T>
__device__ __forceinline__ T from_constant(const T* p) {
    T out;
    __builtin_memcpy(&out, as_constant(p), sizeof(T));
    return out;
}

// struct Squirmer (reference squirmer.hpp:14-37), kept on the device so a
// whole time step (fluid update + rigid-body update + remap) runs without a
// host round trip.
struct DevSquirmer {
    double R, B1, beta, M, I, FS;
    double X[3], e[3], U[3], W[3];
    double Fh[3], Th[3], Fc[3], Tc[3], Fw[3];
    double eps, hc, hmin;
    long long contacts;
    int planar, plane_xz;
};

// Excess-mass partial sum in 128-bit fixed point (fixed_point.hpp): value
// hi + lo * 2^-64.  Integer addition is associative, so the total of the
// partials does not depend on how the nodes were grouped into tiles, blocks
// and kernels (tile shape, XCD mapping, bulk / boundary split, ensemble or
// single lattice); each node's excess is truncated toward zero to a
// multiple of 2^-64 (an error below 5.5e-20 per node).
struct MassPart {
    long long hi;
    unsigned long long lo;
};

constexpr int FIN_MAX_BLOCKS = 128;  // blocks of the mass-correction finalize

// Everything the kernels exchange through device memory besides the fields.
// The mass shift is double-buffered by step parity: step t reads slot t&1
// and clears slot (t+1)&1, so no kernel clears what another block of the
// same launch still reads, and no memset is needed (the step can be
// captured into a graph).
struct DevState {
    DevSquirmer sq;        // Fh, Th published by particle_links_kernel after each step
    double wallForce[3];   // wall momentum exchange of the last step that computed it
    // Pending mass-correction shift d read by step t from slot t&1
    // (lbm.cpp:333, 337: f_i += w_i d).
    double shift[2];
    double link_dmass;     // mass change of the moving-boundary corrections of the last step
    unsigned link_ticket;  // last-block counters of the coupling and wall-force kernels
    unsigned wall_ticket;
    // PARTICLE link list (links_kernel.hpp): entries in use (live and dead)
    // and the live ones the last link pass found
    unsigned nlinks;
    unsigned nvalid;
    // Mass change made by the remaps since the last step, as excess over one
    // unit per fluid node (cover removes sum_i f_i - 1, uncover adds
    // sum_i feq_i - 1).  The first remap after a step sets it, later ones
    // add to it (the host tells the remap which).
    double remap_dmass;
    long long nfluid;   // number of FLUID nodes, maintained by the remap
    int band[6];        // x0, y0, z0 (wrapped) and widths of the rho band
    int err;            // device error bits (ERR_*)
    int changed;        // nodes changed by the last remap
    // Centre and radius the PARTICLE flags were last classified with: every
    // non-WALL node of the remap box is PARTICLE exactly when it lies within
    // rf of xf (set at creation and by each remap; a host-side squirmer
    // update leaves it alone).
    double xf[3];
    double rf;
    // block sums of the mass-correction finalize and its last-block counter
    MassPart fin_part[FIN_MAX_BLOCKS];
    unsigned fin_ticket;
};

constexpr int ERR_REMAP_OVERFLOW = 1;
constexpr int ERR_LINK_OVERFLOW = 2;  // the link list outgrew its capacity

// Arguments of the fused collide-stream kernel.
template <class S>
struct StepArgs {
    S* f;                 // 19 * N populations, AA in-place layout
    uint8_t* flag;        // N flag bytes (NodeType | BOUNDARY)
    void* rho;            // N (double in FP64 mode, float otherwise)
    void* ux;
    void* uy;
    void* uz;
    DevState* st;
    MassPart* mass_part;  // one excess-mass partial per block
    int NX, NY, NZ;
    uint32_t N;
    uint32_t SS;          // slot stride: population i of node (x, y, z) is f[i * SS + x + PX (y + NY z)]
    uint32_t PX;          // population row pitch (NX, or NX rounded up to 64 nodes)
    int walls_y, walls_z, has_squirmer;
    int gx, gy, gz;       // grid in tiles
    int tx_shift;         // log2 of the tile width in x; 0: row-run tiles (see step_tile)
    int nchx;             // 64-node chunks per row
    int ty;               // tile height in y
    int xcd_remap;        // 1: give each XCD a contiguous range of tiles; 2: whole tile rows per XCD
    int odd_rows;         // odd steps in row blocks (aa_odd_row_kernel): 0 no, 1 blocks in order, 2 block ranges per XCD
    int row_waves;        // wavefronts per row (PX / 64) of the row kernel
    int rows_per_block;   // rows per block of the row kernel
    FastDiv div_gx, div_gy, div_nchx, div_ny;  // division by gx, gy, nchx and NY (tile decoding)
    double omPlus, omMinus, Fx, Fy, Fz;
    long long t;          // step index (selects double-buffer slots)
};

}  // namespace lbmhip
