// kernels_fp64.hip -- FP64 step kernels (reference operation order), the
// squirmer / remap kernel for every storage format, and the auxiliary
// kernels.  Compiled with -ffp-contract=off: no multiply-add is fused, so the
// FP64 results round exactly like the reference's scalar code.
#include <algorithm>

#include "launch_step_impl.hpp"
#include "lbm_hip.h"

namespace lbmhip {

template <class S>
hipError_t launch_couple(const LinkArgs<S>& l, const RemapArgs<S>& r, int link_blocks, int links, int remap,
                         hipStream_t s) {
    if (links) {
        hipLaunchKernelGGL((particle_links_kernel<S>), dim3(link_blocks), dim3(LINK_THREADS), 0, s, l);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    hipLaunchKernelGGL((couple_kernel<S>), dim3(1), dim3(CPL_THREADS), 0, s, l, r, link_blocks, links, remap);
    return hipGetLastError();
}

__global__ __launch_bounds__(CPL_THREADS) void link_build_kernel(const uint8_t* flag, DevState* st, uint2* list,
                                                                 uint32_t cap, int NX, int NY, int NZ) {
    link_build_block(flag, st, list, cap, NX, NY, NZ);
}

hipError_t launch_link_build(const uint8_t* flag, DevState* st, uint2* list, uint32_t cap, int NX, int NY, int NZ,
                             hipStream_t s) {
    hipLaunchKernelGGL(link_build_kernel, dim3(1), dim3(CPL_THREADS), 0, s, flag, st, list, cap, NX, NY, NZ);
    return hipGetLastError();
}

template <class S>
hipError_t launch_wall_force(const S* f, const uint8_t* flag, const WallForceArgs& w, double* part, DevState* st,
                             int NX, int NY, int NZ, uint32_t SS, uint32_t PX, hipStream_t s) {
    // at least one block, so that a lattice without walls publishes a zero force
    const long long waves = (long long)w.n_wall_rows * w.nchx;
    const unsigned blocks = (unsigned)std::max(1LL, (waves + 3) / 4);
    hipLaunchKernelGGL((wall_force_kernel<S>), dim3(blocks), dim3(256), 0, s, f, flag, w, part, st, NX, NY, NZ, SS, PX);
    return hipGetLastError();
}

template <class S>
hipError_t launch_gather(const S* f, const uint8_t* flag, double* out, int NX, int NY, int NZ, uint32_t n0,
                         uint32_t cnt, uint32_t SS, uint32_t PX, int shifted, const DevState* st, int shift_slot,
                         hipStream_t s) {
    hipLaunchKernelGGL((gather_kernel<S>), dim3((cnt + 255) / 256), dim3(256), 0, s, f, flag, out, NX, NY, NZ, n0, cnt,
                       SS, PX, shifted, st, shift_slot);
    return hipGetLastError();
}

template <class S>
hipError_t launch_scatter(S* f, const uint8_t* flag, const double* in, int NX, int NY, int NZ, uint32_t n0,
                          uint32_t cnt, uint32_t SS, uint32_t PX, int shifted, hipStream_t s) {
    hipLaunchKernelGGL((scatter_kernel<S>), dim3((cnt + 255) / 256), dim3(256), 0, s, f, flag, in, NX, NY, NZ, n0, cnt,
                       SS, PX, shifted);
    return hipGetLastError();
}

template <class S>
hipError_t launch_mass(const S* f, const uint8_t* flag, MassPart* part, int NX, int NY, int NZ, uint32_t N,
                       uint32_t SS, uint32_t PX, int shifted, const DevState* st, int shift_slot, hipStream_t s) {
    hipLaunchKernelGGL((mass_kernel<S>), dim3((N + 255) / 256), dim3(256), 0, s, f, flag, part, NX, NY, NZ, N, SS,
                       PX, shifted, st, shift_slot);
    return hipGetLastError();
}

// f_i = w_i everywhere (lbm.cpp:82-84), in storage format S (slot padding
// too).  19 * SS elements can exceed the 2^32 work-items of one dispatch, so
// a bounded grid strides over them.
template <class S>
__global__ __launch_bounds__(256) void fill_populations_kernel(S* f, uint32_t SS) {
    const size_t tot = (size_t)Q * SS, stride = (size_t)gridDim.x * 256u;
    for (size_t e = (size_t)blockIdx.x * 256u + threadIdx.x; e < tot; e += stride) {
        const int i = (int)(e / SS);
        f[e] = from_real<S>(W[i], i);
    }
}

template <class S>
hipError_t launch_fill_populations(S* f, uint32_t SS, hipStream_t s) {
    const size_t want = ((size_t)Q * SS + 255) / 256;
    const unsigned grid = (unsigned)std::min<size_t>(want, 1u << 20);
    hipLaunchKernelGGL((fill_populations_kernel<S>), dim3(grid), dim3(256), 0, s, f, SS);
    return hipGetLastError();
}

template <class S>
hipError_t launch_batch_couple(const BatchMember<S>* m, const uint32_t* map, int link_blocks, int count,
                               long long t_after, hipStream_t s) {
    if (link_blocks > 0) {
        hipLaunchKernelGGL((batch_links_kernel<S>), dim3(link_blocks), dim3(LINK_THREADS), 0, s, m, map);
        const hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL((batch_couple_kernel<S>), dim3(count), dim3(CPL_THREADS), 0, s, m, t_after);
    return hipGetLastError();
}

#define LBMHIP_INSTANTIATE(S)                                                                                     \
    template hipError_t launch_batch_couple<S>(const BatchMember<S>*, const uint32_t*, int, int, long long,       \
                                               hipStream_t);                                                      \
    template hipError_t launch_couple<S>(const LinkArgs<S>&, const RemapArgs<S>&, int, int, int, hipStream_t);   \
    template hipError_t launch_wall_force<S>(const S*, const uint8_t*, const WallForceArgs&, double*, DevState*, \
                                             int, int, int, uint32_t, uint32_t, hipStream_t);                             \
    template hipError_t launch_gather<S>(const S*, const uint8_t*, double*, int, int, int, uint32_t, uint32_t, uint32_t, \
                                         uint32_t, int, const DevState*, int, hipStream_t);                                \
    template hipError_t launch_scatter<S>(S*, const uint8_t*, const double*, int, int, int, uint32_t, uint32_t, uint32_t, \
                                          uint32_t, int, hipStream_t);                                                     \
    template hipError_t launch_mass<S>(const S*, const uint8_t*, MassPart*, int, int, int, uint32_t, uint32_t,   \
                                       uint32_t, int, const DevState*, int, hipStream_t);                                  \
    template hipError_t launch_fill_populations<S>(S*, uint32_t, hipStream_t);

LBMHIP_INSTANTIATE(double)
LBMHIP_INSTANTIATE(float)
LBMHIP_INSTANTIATE(_Float16)
LBMHIP_INSTANTIATE(bf16_t)
#undef LBMHIP_INSTANTIATE

template hipError_t launch_step<double>(const StepArgs<double>&, const FastCoef&, bool, int, unsigned, hipStream_t);
template hipError_t step_blocks_per_cu<double>(bool, unsigned, int*);
template hipError_t launch_boundary<double>(const StepArgs<double>&, const FastCoef&, const BoundaryArgs&, int, int,
                                            hipStream_t);
template hipError_t launch_batch_step<double>(const BatchMember<double>*, const uint32_t*, int, long long, int, bool,
                                              int, unsigned, hipStream_t);
template hipError_t launch_batch_boundary<double>(const BatchMember<double>*, const uint32_t*, int, long long, int,
                                                  hipStream_t);
template hipError_t launch_batch_odd_rows<double>(const BatchMember<double>*, const uint32_t*, int, int, long long,
                                                  int, unsigned, hipStream_t);

// ---------------------------------------------------------------------------
// Mass correction (lbm.cpp:324-339): d = 1 - m / n_fluid with
// m = n_fluid + excess, so d = -excess / n_fluid; the shift is added to the
// slot the next step reads.  One block; the partials are fixed-point
// (MassPart), so their total is exact and independent of the grouping.
// `terms` (MASS_TERM_*): the mass changes made after the partials were
// taken that the excess includes.
// ---------------------------------------------------------------------------
// Block sum of fixed-point values (blockDim.x a multiple of 64, at most
// 1024); thread 0 gets the total.
__device__ u128 block_sum_fixed(u128 v) {
    __shared__ unsigned long long lds[2 * 16];
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long h = __shfl_xor((unsigned long long)fixed_hi(v), off, 64);
        const unsigned long long l = __shfl_xor((unsigned long long)fixed_lo(v), off, 64);
        v += fixed_join(h, l);
    }
    if ((threadIdx.x & 63) == 0) {
        lds[2 * (threadIdx.x >> 6)] = (unsigned long long)fixed_hi(v);
        lds[2 * (threadIdx.x >> 6) + 1] = fixed_lo(v);
    }
    __syncthreads();
    u128 s = 0;
    if (threadIdx.x == 0)
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += fixed_join(lds[2 * w], lds[2 * w + 1]);
    return s;
}

// The shift d = -excess / n_fluid from the fixed-point excess total.
__device__ void apply_mass_total(u128 s, DevState* st, int terms, int shift_slot) {
    double e = fixed_to_double(s);
    if (terms & MASS_TERM_REMAP) e += st->remap_dmass;
    if (terms & MASS_TERM_LINKS) e += st->link_dmass;
    const long long nf = st->nfluid;
    if (nf > 0) st->shift[shift_slot] += -e / (double)nf;
}

__device__ __forceinline__ u128 load_part(const MassPart* p) {
    return fixed_join((unsigned long long)p->hi, p->lo);
}

// One block of 1024 threads over all partials (an ensemble member's).
__device__ void mass_finalize_block(const MassPart* part, int nparts, DevState* st, int terms, int shift_slot) {
    u128 a[4] = {0, 0, 0, 0};  // four loads in flight per thread
    int k = threadIdx.x;
    for (; k + 3 * 1024 < nparts; k += 4 * 1024) {
#pragma unroll
        for (int j = 0; j < 4; ++j) a[j] += load_part(&part[k + j * 1024]);
    }
    for (; k < nparts; k += 1024) a[0] += load_part(&part[k]);
    const u128 s = block_sum_fixed((a[0] + a[1]) + (a[2] + a[3]));
    if (threadIdx.x == 0) apply_mass_total(s, st, terms, shift_slot);
}

// Single lattice: FIN blocks each sum a strided share of the partials into
// st->fin_part; the last block to finish adds those and sets the shift.
constexpr int FIN_BLOCK = 256;
__global__ __launch_bounds__(FIN_BLOCK) void mass_finalize_kernel(const MassPart* part, int nparts, DevState* st,
                                                                  int terms, int shift_slot) {
    __shared__ bool is_last;
    u128 v = 0;
    for (int k = blockIdx.x * FIN_BLOCK + threadIdx.x; k < nparts; k += gridDim.x * FIN_BLOCK) v += load_part(&part[k]);
    v = block_sum_fixed(v);
    if (threadIdx.x == 0) {
        st->fin_part[blockIdx.x].hi = fixed_hi(v);
        st->fin_part[blockIdx.x].lo = fixed_lo(v);
        __threadfence();
        is_last = atomicAdd(&st->fin_ticket, 1u) == gridDim.x - 1;
    }
    __syncthreads();
    if (!is_last) return;
    __threadfence();
    u128 w = 0;
    for (unsigned b = threadIdx.x; b < gridDim.x; b += FIN_BLOCK)
        w += fixed_join((unsigned long long)__builtin_nontemporal_load(&st->fin_part[b].hi),
                       __builtin_nontemporal_load(&st->fin_part[b].lo));
    w = block_sum_fixed(w);
    if (threadIdx.x == 0) {
        apply_mass_total(w, st, terms, shift_slot);
        st->fin_ticket = 0u;  // ready for the next launch
    }
}

// One block per ensemble member; members without a squirmer take no mass
// correction (main.cpp:62-64).  `t_after`: steps done, the last one with mass
// partials.
__global__ __launch_bounds__(1024) void batch_finalize_kernel(const BatchFinalize* fa, long long t_after) {
    const BatchFinalize& f = fa[blockIdx.x];
    if (!f.has_squirmer) return;
    const bool last_odd = ((t_after - 1) & 1) != 0;
    mass_finalize_block(f.part, last_odd ? f.G_odd : f.G + f.bblocks, f.st, MASS_TERM_LINKS | MASS_TERM_REMAP,
                        (int)(t_after & 1));
}

hipError_t launch_batch_finalize(const BatchFinalize* fa, int count, long long t_after, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    hipLaunchKernelGGL(batch_finalize_kernel, dim3(count), dim3(1024), 0, s, fa, t_after);
    return hipGetLastError();
}

hipError_t launch_mass_finalize(const MassPart* part, int nparts, DevState* st, int terms, int shift_slot,
                                hipStream_t s) {
    int blocks = (nparts + 4 * FIN_BLOCK - 1) / (4 * FIN_BLOCK);  // about four partials per thread
    blocks = blocks < 1 ? 1 : (blocks > FIN_MAX_BLOCKS ? FIN_MAX_BLOCKS : blocks);
    hipLaunchKernelGGL(mass_finalize_kernel, dim3(blocks), dim3(FIN_BLOCK), 0, s, part, nparts, st, terms, shift_slot);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Diagnostics (lbm.cpp:342-364): count of FLUID nodes, sum of rho, maximum
// speed and sum of ux over FLUID nodes, and the count of unstable FLUID
// nodes (non-finite moments, rho outside (0, LBM_HIP_RHO_MAX), speed at or
// above 1/sqrt(3)); per-block partials, then one block.  The maximum
// propagates a NaN speed (a comparison with NaN would drop it).
// ---------------------------------------------------------------------------
constexpr int NDIAG = 5;

__device__ __forceinline__ double diag_max(double a, double b) { return (a != a || a > b) ? a : b; }

__device__ __forceinline__ void diag_combine(double (&v)[NDIAG], const double* o) {
    v[0] += o[0];
    v[1] += o[1];
    v[2] = diag_max(v[2], o[2]);
    v[3] += o[3];
    v[4] += o[4];
}

__device__ __forceinline__ void diag_wave(double (&v)[NDIAG]) {
    for (int off = 32; off > 0; off >>= 1) {
        double o[NDIAG];
        for (int k = 0; k < NDIAG; ++k) o[k] = __shfl_xor(v[k], off, 64);
        diag_combine(v, o);
    }
}

__global__ __launch_bounds__(256) void diag_kernel(const uint8_t* flag, const void* rho, const void* ux,
                                                   const void* uy, const void* uz, int is_double, uint32_t N,
                                                   double* part) {
    __shared__ double lds[4 * NDIAG];
    const uint32_t n = blockIdx.x * 256u + threadIdx.x;
    double v[NDIAG] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (n < N && (flag[n] & TYPE_MASK) == FLUID) {
        double r, a, b, c;
        if (is_double) {
            r = ((const double*)rho)[n]; a = ((const double*)ux)[n];
            b = ((const double*)uy)[n]; c = ((const double*)uz)[n];
        } else {
            r = ((const float*)rho)[n]; a = ((const float*)ux)[n];
            b = ((const float*)uy)[n]; c = ((const float*)uz)[n];
        }
        const double u2 = a * a + b * b + c * c;
        v[0] = 1.0; v[1] = r; v[2] = sqrt(u2); v[3] = a;
        // the negated test also catches NaN: every comparison with NaN is false
        const bool ok = r > 0.0 && r < LBM_HIP_RHO_MAX && u2 < 1.0 / 3.0;
        v[4] = ok ? 0.0 : 1.0;
    }
    diag_wave(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < NDIAG; ++k) lds[w * NDIAG + k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        double o[NDIAG];
        for (int k = 0; k < NDIAG; ++k) o[k] = lds[k];
        for (int ww = 1; ww < 4; ++ww) diag_combine(o, &lds[ww * NDIAG]);
        for (int k = 0; k < NDIAG; ++k) part[(size_t)blockIdx.x * NDIAG + k] = o[k];
    }
}

__global__ __launch_bounds__(1024) void diag_finalize_kernel(const double* part, int nparts, double* out) {
    __shared__ double lds[16 * NDIAG];
    double v[NDIAG] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = threadIdx.x; k < nparts; k += 1024) diag_combine(v, &part[(size_t)k * NDIAG]);
    diag_wave(v);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < NDIAG; ++k) lds[w * NDIAG + k] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
        double o[NDIAG];
        for (int k = 0; k < NDIAG; ++k) o[k] = lds[k];
        for (int ww = 1; ww < 16; ++ww) diag_combine(o, &lds[ww * NDIAG]);
        for (int k = 0; k < NDIAG; ++k) out[k] = o[k];
    }
}

hipError_t launch_diag(const uint8_t* flag, const void* rho, const void* ux, const void* uy, const void* uz,
                       int is_double, uint32_t N, double* part, int nparts, double* out5, hipStream_t s) {
    hipLaunchKernelGGL(diag_kernel, dim3(nparts), dim3(256), 0, s, flag, rho, ux, uy, uz, is_double, N, part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(diag_finalize_kernel, dim3(1), dim3(1024), 0, s, part, nparts, out5);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Initial node flags (lbm.cpp:72-93) on the device: WALL layers, the
// PARTICLE nodes of the initial sphere (classified on the host over the
// sphere's box, uploaded as a list), then the BOUNDARY bit of every FLUID
// node with a non-FLUID neighbour.  One thread per node.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void init_flags_kernel(uint8_t* flag, int NX, int NY, int NZ, uint32_t N,
                                                         int walls_y, int walls_z) {
    const uint32_t n = blockIdx.x * 256u + threadIdx.x;
    if (n >= N) return;
    const uint32_t r = n / (uint32_t)NX;
    const int y = (int)(r % (uint32_t)NY), z = (int)(r / (uint32_t)NY);
    const bool wall = (walls_y && (y == 0 || y == NY - 1)) || (walls_z && (z == 0 || z == NZ - 1));
    flag[n] = wall ? (uint8_t)WALL : (uint8_t)FLUID;
}

__global__ __launch_bounds__(256) void set_flags_kernel(uint8_t* flag, const uint32_t* nodes, int count,
                                                        uint8_t value) {
    const int k = (int)(blockIdx.x * 256u + threadIdx.x);
    if (k < count) flag[nodes[k]] = value;
}

__global__ __launch_bounds__(256) void boundary_bits_kernel(uint8_t* flag, int NX, int NY, int NZ, uint32_t N) {
    const uint32_t n = blockIdx.x * 256u + threadIdx.x;
    if (n >= N || flag[n] != FLUID) return;
    const uint32_t r = n / (uint32_t)NX;
    const int x = (int)(n - r * (uint32_t)NX), y = (int)(r % (uint32_t)NY), z = (int)(r / (uint32_t)NY);
    const int xs[3] = {x == 0 ? NX - 1 : x - 1, x, x + 1 == NX ? 0 : x + 1};
    const int ys[3] = {y == 0 ? NY - 1 : y - 1, y, y + 1 == NY ? 0 : y + 1};
    const int zs[3] = {z == 0 ? NZ - 1 : z - 1, z, z + 1 == NZ ? 0 : z + 1};
    bool b = false;
#pragma unroll
    for (int i = 1; i < Q; ++i) {
        const uint32_t m = (uint32_t)xs[1 + CX[i]] + (uint32_t)NX * ((uint32_t)ys[1 + CY[i]] + (uint32_t)NY * (uint32_t)zs[1 + CZ[i]]);
        b |= (flag[m] & TYPE_MASK) != FLUID;
    }
    // only this node's own byte is written; neighbours are read-only here
    // (the BOUNDARY bit lives in bits the type test masks off)
    if (b) flag[n] = (uint8_t)(FLUID | BOUNDARY);
}

hipError_t launch_init_flags(uint8_t* flag, int NX, int NY, int NZ, uint32_t N, int walls_y, int walls_z,
                             const uint32_t* particle_nodes, int nparticle, hipStream_t s) {
    const unsigned blocks = (unsigned)(((size_t)N + 255) / 256);
    hipLaunchKernelGGL(init_flags_kernel, dim3(blocks), dim3(256), 0, s, flag, NX, NY, NZ, N, walls_y, walls_z);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && nparticle > 0) {
        hipLaunchKernelGGL(set_flags_kernel, dim3((unsigned)((nparticle + 255) / 256)), dim3(256), 0, s, flag,
                           particle_nodes, nparticle, (uint8_t)PARTICLE);
        e = hipGetLastError();
    }
    if (e == hipSuccess) {
        hipLaunchKernelGGL(boundary_bits_kernel, dim3(blocks), dim3(256), 0, s, flag, NX, NY, NZ, N);
        e = hipGetLastError();
    }
    return e;
}

__global__ __launch_bounds__(256) void fill_macro_kernel(void* rho, void* ux, void* uy, void* uz, int is_double,
                                                         uint32_t N) {
    const uint32_t n = blockIdx.x * 256u + threadIdx.x;
    if (n >= N) return;
    if (is_double) {
        ((double*)rho)[n] = 1.0; ((double*)ux)[n] = 0.0; ((double*)uy)[n] = 0.0; ((double*)uz)[n] = 0.0;
    } else {
        ((float*)rho)[n] = 1.0f; ((float*)ux)[n] = 0.0f; ((float*)uy)[n] = 0.0f; ((float*)uz)[n] = 0.0f;
    }
}

hipError_t launch_fill_macro(void* rho, void* ux, void* uy, void* uz, int is_double, uint32_t N, hipStream_t s) {
    hipLaunchKernelGGL(fill_macro_kernel, dim3((N + 255) / 256), dim3(256), 0, s, rho, ux, uy, uz, is_double, N);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void macro_to_double_kernel(const float* src, double* dst, uint32_t N) {
    const uint32_t n = blockIdx.x * 256u + threadIdx.x;
    if (n < N) dst[n] = (double)src[n];
}

hipError_t launch_macro_to_double(const void* src, int is_double, double* dst, uint32_t N, hipStream_t s) {
    if (is_double) return hipMemcpyAsync(dst, src, sizeof(double) * N, hipMemcpyDeviceToDevice, s);
    hipLaunchKernelGGL(macro_to_double_kernel, dim3((N + 255) / 256), dim3(256), 0, s, (const float*)src, dst, N);
    return hipGetLastError();
}

// Streaming copy of n16 16-byte words with nontemporal loads and stores
// (the bandwidth probe).
typedef uint32_t word4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void copy_probe_kernel(const word4* __restrict__ a, word4* __restrict__ b,
                                                         size_t n16) {
    const size_t stride = (size_t)gridDim.x * 256u;
    for (size_t i = (size_t)blockIdx.x * 256u + threadIdx.x; i < n16; i += stride)
        __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}

hipError_t launch_copy_probe(const void* a, void* b, size_t n16, hipStream_t s) {
    const size_t want = (n16 + 255) / 256;
    const unsigned grid = (unsigned)(want < 65536 ? (want ? want : 1) : 65536);
    hipLaunchKernelGGL(copy_probe_kernel, dim3(grid), dim3(256), 0, s, (const word4*)a, (word4*)b, n16);
    return hipGetLastError();
}

// The same copy over `parts` separate allocations in one launch, as the
// ensemble kernels cover their members: blocks [p * bpp, (p + 1) * bpp)
// stream part p (a[p] -> b[p], n16 16-byte words each).
__global__ __launch_bounds__(256) void copy_parts_kernel(const word4* const* __restrict__ a, word4* const* __restrict__ b,
                                                         size_t n16, unsigned bpp) {
    const unsigned p = blockIdx.x / bpp, k = blockIdx.x % bpp;
    const word4* src = a[p];
    word4* dst = b[p];
    const size_t stride = (size_t)bpp * 256u;
    for (size_t i = (size_t)k * 256u + threadIdx.x; i < n16; i += stride)
        __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

hipError_t launch_copy_parts(const void* const* a, void* const* b, int parts, size_t n16, hipStream_t s) {
    const size_t want = (n16 + 255) / 256;
    const unsigned per = 65536u / (unsigned)parts;
    const unsigned bpp = (unsigned)(want < per ? (want ? want : 1) : per);
    hipLaunchKernelGGL(copy_parts_kernel, dim3(bpp * (unsigned)parts), dim3(256), 0, s, (const word4* const*)a,
                       (word4* const*)b, n16, bpp);
    return hipGetLastError();
}

}  // namespace lbmhip
