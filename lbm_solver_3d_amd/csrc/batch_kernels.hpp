// batch_kernels.hpp -- one launch per phase for an ensemble of lattices.
//
// Every member lattice keeps its own arrays and device state; a table of
// per-member argument blocks (BatchMember) sits in device memory.  Each
// launch covers the work of all members: a block reads which member and
// which of that member's tiles / blocks it is from a small map (member << 24
// | local index), loads that member's arguments with uniform (scalar) loads
// and runs the same body as the single-lattice kernel.  Members therefore
// compute exactly what their own kernels compute; only the launch is shared.
#pragma once
#include "device_state.hpp"
#include "links_kernel.hpp"
#include "squirmer_kernel.hpp"
#include "step_kernel.hpp"

namespace lbmhip {

template <class S>
struct BatchMember {
    StepArgs<S> A;      // step index set per launch
    FastCoef K;
    BoundaryArgs B;
    LinkArgs<S> L;
    RemapArgs<S> R;     // layout parity and shift slot set per launch
    int lblocks;        // blocks of this member's link pass (LINK_THREADS threads each)
    int has_squirmer;
};

// Per-member arguments of the mass-correction finalize.
struct BatchFinalize {
    const MassPart* part;
    DevState* st;
    int G, bblocks, has_squirmer;
    int G_odd;  // collide-stream partials of an odd step (G, or the member's row blocks)
};

constexpr uint32_t BATCH_LOCAL_MASK = 0xFFFFFFu;

// The member tables are not written during a launch: the kernels read them
// with from_constant (scalar loads, see device_state.hpp).

template <class S, bool ODD, int MODE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(LBM_STEP_WAVES_PER_EU)))
void batch_step_kernel(const BatchMember<S>* __restrict__ M, const uint32_t* __restrict__ tile_map, int Gtot,
                       long long t, int xcd) {
    const int g = xcd ? xcd_tile(blockIdx.x, Gtot) : (int)blockIdx.x;
    const uint32_t e = as_constant(tile_map)[g];
    const BatchMember<S>* mb = M + (e >> 24);
    StepArgs<S> A = from_constant(&mb->A);
    const FastCoef K = from_constant(&mb->K);
    A.t = t;
    step_tile<S, ODD, MODE>(A, K, (int)(e & BATCH_LOCAL_MASK));
}

// Odd step in row blocks (odd_rows_block) for every member: block map
// member << 24 | the member's row block; the member's A.rows_per_block rows
// of A.row_waves wavefronts per block, the launch's block holding the most
// wavefronts any member needs.
template <class S, int MODE>
__global__ __launch_bounds__(64 * ROW_MAX_WAVES) __attribute__((amdgpu_waves_per_eu(LBM_STEP_WAVES_PER_EU)))
void batch_odd_row_kernel(const BatchMember<S>* __restrict__ M, const uint32_t* __restrict__ row_map, long long t) {
    const uint32_t e = as_constant(row_map)[blockIdx.x];
    const BatchMember<S>* mb = M + (e >> 24);
    StepArgs<S> A = from_constant(&mb->A);
    const FastCoef K = from_constant(&mb->K);
    A.t = t;
    const int lb = (int)(e & BATCH_LOCAL_MASK);
    odd_rows_block<S, MODE, true>(A, K, lb * A.rows_per_block, A.rows_per_block, &A.mass_part[lb]);
}

template <class S, int MODE>
__global__ __launch_bounds__(BLOCK) void batch_boundary_kernel(const BatchMember<S>* __restrict__ M,
                                                               const uint32_t* __restrict__ block_map, long long t) {
    const uint32_t e = as_constant(block_map)[blockIdx.x];
    const BatchMember<S>* mb = M + (e >> 24);
    StepArgs<S> A = from_constant(&mb->A);
    const FastCoef K = from_constant(&mb->K);
    const BoundaryArgs B = from_constant(&mb->B);
    A.t = t;
    boundary_block<S, MODE>(A, K, B, (int)(e & BATCH_LOCAL_MASK));
}

// The link pass of every member with a squirmer (block map: member << 24 |
// link block), one partial per block into the member's own array.
template <class S>
__global__ __launch_bounds__(LINK_THREADS) void batch_links_kernel(const BatchMember<S>* __restrict__ M,
                                                                   const uint32_t* __restrict__ block_map) {
    const uint32_t e = as_constant(block_map)[blockIdx.x];
    const BatchMember<S>* mb = M + (e >> 24);
    link_pass_block<S>(from_constant(&mb->L), (int)(e & BATCH_LOCAL_MASK), from_constant(&mb->lblocks));
}

// One block per member: the link sums, rigid-body update, remap and
// link-list upkeep.  `t_after` is the number of steps done.
template <class S>
__global__ __launch_bounds__(CPL_THREADS) void batch_couple_kernel(const BatchMember<S>* __restrict__ M,
                                                                   long long t_after) {
    const BatchMember<S>* mb = M + blockIdx.x;
    if (!from_constant(&mb->has_squirmer)) return;
    RemapArgs<S> R = from_constant(&mb->R);
    R.shifted = (t_after > 0 && ((t_after - 1) & 1) == 0) ? 1 : 0;
    R.shift_slot = (int)(t_after & 1);
    R.accumulate = 0;
    R.maintain = 1;
    couple_block<S>(from_constant(&mb->L), R, from_constant(&mb->lblocks), 1, 1);
}

}  // namespace lbmhip
