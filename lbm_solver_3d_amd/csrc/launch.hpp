// launch.hpp -- host-side launchers of the device kernels.  The FP64 kernels
// and the squirmer/remap kernels are compiled in kernels_fp64.hip with
// floating-point contraction off (reference operation order); the FP32 /
// FP16 / BF16 step kernels in kernels_fast.hip, also without contraction:
// their multiply-adds are explicit fmaf, so every kernel that updates a node
// (bulk, boundary, ensemble) rounds the same way.
#pragma once
#include <hip/hip_runtime.h>

#include "batch_kernels.hpp"
#include "device_state.hpp"
#include "links_kernel.hpp"
#include "squirmer_kernel.hpp"
#include "step_kernel.hpp"
#include "storage.hpp"

namespace lbmhip {

// `mode`: STEP_MACRO | STEP_MASS as requested, plus STEP_BODY_FORCE when the
// lattice (for an ensemble: any member) has a body force.  `lds`: dynamic
// LDS bytes per block, which caps the blocks resident per CU (0: no cap).
template <class S>
hipError_t launch_step(const StepArgs<S>& a, const FastCoef& k, bool odd, int mode, unsigned lds, hipStream_t s);

template <class S>
hipError_t launch_boundary(const StepArgs<S>& a, const FastCoef& k, const BoundaryArgs& b, int blocks, int mode,
                           hipStream_t s);

// Blocks of the collide-stream kernel (no moments, no mass partials) that
// can reside on one CU with `lds` dynamic LDS bytes per block, as HIP's
// occupancy calculator reports it.
template <class S>
hipError_t step_blocks_per_cu(bool odd, unsigned lds, int* blocks);

// The coupling of a step: the link pass over the link list (`links`;
// `link_blocks` blocks of LINK_THREADS threads), then one block that adds
// its partial sums and runs the rigid-body update and remap (`remap`).
template <class S>
hipError_t launch_couple(const LinkArgs<S>& l, const RemapArgs<S>& r, int link_blocks, int links, int remap,
                         hipStream_t s);

// Link list from the flags (one block).
hipError_t launch_link_build(const uint8_t* flag, DevState* st, uint2* list, uint32_t cap, int NX, int NY, int NZ,
                             hipStream_t s);

template <class S>
hipError_t launch_wall_force(const S* f, const uint8_t* flag, const WallForceArgs& w, double* part, DevState* st,
                             int NX, int NY, int NZ, uint32_t SS, uint32_t PX, hipStream_t s);

// Logical populations of the nodes n0 .. n0 + cnt - 1 to / from a staging
// array [19][cnt] of doubles.
template <class S>
hipError_t launch_gather(const S* f, const uint8_t* flag, double* out, int NX, int NY, int NZ, uint32_t n0,
                         uint32_t cnt, uint32_t SS, uint32_t PX, int shifted, const DevState* st, int shift_slot,
                         hipStream_t s);

template <class S>
hipError_t launch_scatter(S* f, const uint8_t* flag, const double* in, int NX, int NY, int NZ, uint32_t n0,
                          uint32_t cnt, uint32_t SS, uint32_t PX, int shifted, hipStream_t s);

template <class S>
hipError_t launch_mass(const S* f, const uint8_t* flag, MassPart* part, int NX, int NY, int NZ, uint32_t N,
                       uint32_t SS, uint32_t PX, int shifted, const DevState* st, int shift_slot, hipStream_t s);

template <class S>
hipError_t launch_fill_populations(S* f, uint32_t SS, hipStream_t s);

// Ensemble launches (batch_kernels.hpp): all members in one launch per phase.
template <class S>
hipError_t launch_batch_step(const BatchMember<S>* m, const uint32_t* tile_map, int Gtot, long long t, int xcd,
                             bool odd, int mode, unsigned lds, hipStream_t s);
template <class S>
hipError_t launch_batch_boundary(const BatchMember<S>* m, const uint32_t* block_map, int blocks, long long t,
                                 int mode, hipStream_t s);
// Odd step in row blocks of `threads` threads (row_map: member << 24 | row block).
template <class S>
hipError_t launch_batch_odd_rows(const BatchMember<S>* m, const uint32_t* row_map, int blocks, int threads,
                                 long long t, int mode, unsigned lds, hipStream_t s);
template <class S>
hipError_t launch_batch_couple(const BatchMember<S>* m, const uint32_t* block_map, int link_blocks, int count,
                               long long t_after, hipStream_t s);
hipError_t launch_batch_finalize(const BatchFinalize* fa, int count, long long t_after, hipStream_t s);

// Storage-independent kernels (kernels_fp64.hip).
constexpr int MASS_TERM_LINKS = 1;  // moving-boundary corrections of the last step (link_dmass)
constexpr int MASS_TERM_REMAP = 2;  // remaps since the last step (remap_dmass)
hipError_t launch_mass_finalize(const MassPart* part, int nparts, DevState* st, int terms, int shift_slot,
                                hipStream_t s);
// Diagnostics: out5 = {fluid nodes, sum rho, max speed, sum ux, unstable nodes};
// `part` holds 5 doubles per block of 256 nodes.
hipError_t launch_diag(const uint8_t* flag, const void* rho, const void* ux, const void* uy, const void* uz,
                       int is_double, uint32_t N, double* part, int nparts, double* out5, hipStream_t s);
hipError_t launch_fill_macro(void* rho, void* ux, void* uy, void* uz, int is_double, uint32_t N, hipStream_t s);
// Initial flags: WALL layers, the listed PARTICLE nodes, BOUNDARY bits.
hipError_t launch_init_flags(uint8_t* flag, int NX, int NY, int NZ, uint32_t N, int walls_y, int walls_z,
                             const uint32_t* particle_nodes, int nparticle, hipStream_t s);
hipError_t launch_macro_to_double(const void* src, int is_double, double* dst, uint32_t N, hipStream_t s);
hipError_t launch_copy_probe(const void* a, void* b, size_t n16, hipStream_t s);
// The copy probe over `parts` separate allocations (a[p] -> b[p], n16
// 16-byte words each; parts <= 65536), one launch.
hipError_t launch_copy_parts(const void* const* a, void* const* b, int parts, size_t n16, hipStream_t s);

#define LBMHIP_DECLARE(S)                                                                                          \
    extern template hipError_t launch_step<S>(const StepArgs<S>&, const FastCoef&, bool, int, unsigned,         \
                                              hipStream_t);                                                     \
    extern template hipError_t launch_boundary<S>(const StepArgs<S>&, const FastCoef&, const BoundaryArgs&, int, \
                                                  int, hipStream_t);                                             \
    extern template hipError_t step_blocks_per_cu<S>(bool, unsigned, int*);                                     \
    extern template hipError_t launch_couple<S>(const LinkArgs<S>&, const RemapArgs<S>&, int, int, int,         \
                                                hipStream_t);                                                   \
    extern template hipError_t launch_wall_force<S>(const S*, const uint8_t*, const WallForceArgs&, double*,      \
                                                    DevState*, int, int, int, uint32_t, uint32_t, hipStream_t);           \
    extern template hipError_t launch_gather<S>(const S*, const uint8_t*, double*, int, int, int, uint32_t,      \
                                                uint32_t, uint32_t, uint32_t, int, const DevState*, int, hipStream_t);    \
    extern template hipError_t launch_scatter<S>(S*, const uint8_t*, const double*, int, int, int, uint32_t,     \
                                                 uint32_t, uint32_t, uint32_t, int, hipStream_t);                         \
    extern template hipError_t launch_mass<S>(const S*, const uint8_t*, MassPart*, int, int, int, uint32_t,      \
                                              uint32_t, uint32_t, int, const DevState*, int, hipStream_t);                \
    extern template hipError_t launch_fill_populations<S>(S*, uint32_t, hipStream_t);                         \
    extern template hipError_t launch_batch_step<S>(const BatchMember<S>*, const uint32_t*, int, long long, int, \
                                                    bool, int, unsigned, hipStream_t);                          \
    extern template hipError_t launch_batch_boundary<S>(const BatchMember<S>*, const uint32_t*, int, long long,  \
                                                        int, hipStream_t);                                      \
    extern template hipError_t launch_batch_odd_rows<S>(const BatchMember<S>*, const uint32_t*, int, int,        \
                                                        long long, int, unsigned, hipStream_t);                 \
    extern template hipError_t launch_batch_couple<S>(const BatchMember<S>*, const uint32_t*, int, int, long long, \
                                                      hipStream_t);

LBMHIP_DECLARE(double)
LBMHIP_DECLARE(float)
LBMHIP_DECLARE(_Float16)
LBMHIP_DECLARE(bf16_t)
#undef LBMHIP_DECLARE

}  // namespace lbmhip
