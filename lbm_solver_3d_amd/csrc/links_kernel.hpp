// links_kernel.hpp -- the moving-boundary bounce-back and momentum exchange
// of PARTICLE links, and the wall momentum exchange, applied after the fused
// step to the slots where the step left the bounced populations.
//
// After either step parity the bounced population of a link (n, i) with a
// solid node n + c_i is in slot i of n + c_i, holding the post-collision
// value f*_i of node n (the wall bounce-back of lbm.cpp:197-199 with u_b = 0).
// particle_links_kernel turns it into
//     f_b = f*_i - 6 w_i rho_n (c_i . u_b)          (lbm.cpp:189-199)
// with u_b the squirmer surface velocity (rigid body + two-mode slip) at the
// link midpoint, and sums the momentum exchange (Wen et al. 2014)
//     dp = (f* + f_b) c_i - (f* - f_b) u_b,   torque r_b x dp   (lbm.cpp:202-215)
// in FP64.  The sums are deterministic: per thread in direction order, per
// block in a fixed tree, then the last block to finish adds the block
// partials in block order and publishes Fh, Th (lbm.cpp:220-223).
#pragma once
#include "device_state.hpp"
#include "step_kernel.hpp"
#include "storage.hpp"

namespace lbmhip {

constexpr int LINK_BLOCK = 256;
constexpr int LINK_MAX_BLOCKS = 8192;

template <class S>
struct LinkArgs {
    S* f;
    const uint8_t* flag;
    const void* rho;     // density of this step (double in FP64 mode, float otherwise)
    DevState* st;
    double* part;        // LINK_MAX_BLOCKS x 8 partial sums
    int NX, NY, NZ;
    uint32_t SS, PX;  // slot stride, population row pitch
    int walls_y, walls_z;
};

// Box of nodes that can own a particle link: the sphere's nodes lie within
// R of X, their neighbours within R + 1 (lbm.cpp:273-275 without the margin).
struct LinkBox {
    int x0, y0, z0, wx, wy, wz;
};
__device__ inline LinkBox link_box(const double X[3], double R, int NX, int NY, int NZ) {
    LinkBox b;
    const int Nd[3] = {NX, NY, NZ};
    int lo[3], wd[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = (int)floor(X[d] - R) - 1;
        const int hi = (int)ceil(X[d] + R) + 1;
        int w = hi - lo[d] + 1;
        wd[d] = w > Nd[d] ? Nd[d] : w;
    }
    b.x0 = lo[0]; b.y0 = lo[1]; b.z0 = lo[2];
    b.wx = wd[0]; b.wy = wd[1]; b.wz = wd[2];
    return b;
}

__device__ inline int wrap_once(int a, int n) {
    a %= n;
    return a < 0 ? a + n : a;
}

// Fixed-order combination of per-block partials by the last block to finish.
// Returns true in the last block; `out` then holds the sums (thread 0).
// `block` / `nblocks`: this block's index among the blocks that share the
// ticket and the partials.
template <int K>
__device__ bool last_block_sum(double* part, unsigned* ticket, const double (&mine)[K], double (&out)[K],
                               double* lds, unsigned block, unsigned nblocks) {
    __shared__ bool is_last;
    if (threadIdx.x == 0) {
        for (int k = 0; k < K; ++k) part[(size_t)block * 8 + k] = mine[k];
        __threadfence();
        const unsigned t = atomicAdd(ticket, 1u);
        is_last = t == nblocks - 1;
    }
    __syncthreads();
    if (!is_last) return false;
    __threadfence();
    double v[K];
    for (int k = 0; k < K; ++k) v[k] = 0.0;
    for (unsigned b = threadIdx.x; b < nblocks; b += blockDim.x)
        for (int k = 0; k < K; ++k) v[k] += __builtin_nontemporal_load(&part[(size_t)b * 8 + k]);
    block_reduce<K>(v, lds);
    if (threadIdx.x == 0) {
        for (int k = 0; k < K; ++k) out[k] = v[k];
        *ticket = 0u;  // ready for the next launch
    }
    return true;
}

// One block covers LINK_BLOCK consecutive nodes of the box.  Its threads
// first find their node's PARTICLE links (one flag load per neighbour, all in
// flight together) and compact them into LDS in (node, direction) order;
// then every thread takes links of that list, so a wavefront's lanes all do
// link work although only a thin shell of the box owns links.
template <class S>
__device__ __forceinline__ void links_block(const LinkArgs<S>& A, int block, int nblocks) {
    constexpr bool EXACT = std::is_same<S, double>::value;
    __shared__ double lds[(LINK_BLOCK / 64) * 7];
    __shared__ uint16_t lk[LINK_BLOCK * (Q - 1)];  // thread << 5 | direction
    __shared__ double nrho[LINK_BLOCK];
    __shared__ int nx[LINK_BLOCK], ny[LINK_BLOCK], nz[LINK_BLOCK];
    __shared__ int wsum[LINK_BLOCK / 64];
    DevState* st = A.st;
    const DevSquirmer& sq = st->sq;
    const int NX = A.NX, NY = A.NY, NZ = A.NZ;
    double cs[3];
    const double rs = covering_sphere(st, NX, NY, NZ, cs);
    const LinkBox b = link_box(cs, rs, NX, NY, NZ);
    const int M = b.wx * b.wy * b.wz;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int k = block * LINK_BLOCK + tid;
    auto node_of = [&](int x, int y, int z) -> uint32_t {
        return (uint32_t)x + (uint32_t)NX * ((uint32_t)y + (uint32_t)NY * (uint32_t)z);
    };
    uint32_t pmask = 0;
    if (k < M) {
        const int ix = k % b.wx, iy = (k / b.wx) % b.wy, iz = k / (b.wx * b.wy);
        const int x = wrap_once(b.x0 + ix, NX), y = wrap_once(b.y0 + iy, NY), z = wrap_once(b.z0 + iz, NZ);
        const uint32_t n = node_of(x, y, z);
        const uint8_t fl = A.flag[n];
        if ((fl & TYPE_MASK) == FLUID && (fl & BOUNDARY)) {
            if constexpr (EXACT) nrho[tid] = static_cast<const double*>(A.rho)[n];
            else nrho[tid] = (double)static_cast<const float*>(A.rho)[n];
            nx[tid] = x; ny[tid] = y; nz[tid] = z;
            const int xm = x == 0 ? NX - 1 : x - 1, xp = x + 1 == NX ? 0 : x + 1;
            const int ym = y == 0 ? NY - 1 : y - 1, yp = y + 1 == NY ? 0 : y + 1;
            const int zm = z == 0 ? NZ - 1 : z - 1, zp = z + 1 == NZ ? 0 : z + 1;
            const int xs[3] = {xm, x, xp}, ys[3] = {ym, y, yp}, zs[3] = {zm, z, zp};
            static_for<1, Q>([&](auto I) {
                constexpr int i = decltype(I)::value;
                const uint32_t m = node_of(xs[1 + CX[i]], ys[1 + CY[i]], zs[1 + CZ[i]]);
                pmask |= (uint32_t)((A.flag[m] & TYPE_MASK) == PARTICLE) << i;
            });
        }
    }
    // block exclusive scan of the link counts, then the compacted list
    const int cnt = __builtin_popcount(pmask);
    int incl = cnt;
    for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int base = 0, total = 0;
    for (int w = 0; w < LINK_BLOCK / 64; ++w) {
        const int v = wsum[w];
        base += w < wave ? v : 0;
        total += v;
    }
    int pos = base + incl - cnt;
    while (pmask) {
        const int i = __builtin_ctz(pmask);
        pmask &= pmask - 1;
        lk[pos++] = (uint16_t)(tid << 5 | i);
    }
    __syncthreads();

    double acc[7] = {0, 0, 0, 0, 0, 0, 0};  // Fx Fy Fz Tx Ty Tz dmass
    for (int q = tid; q < total; q += LINK_BLOCK) {
        const int e = lk[q], t = e >> 5, i = e & 31;
        const int x = nx[t], y = ny[t], z = nz[t];
        const int xn = x + CX[i] < 0 ? NX - 1 : (x + CX[i] >= NX ? 0 : x + CX[i]);
        const int yn = y + CY[i] < 0 ? NY - 1 : (y + CY[i] >= NY ? 0 : y + CY[i]);
        const int zn = z + CZ[i] < 0 ? NZ - 1 : (z + CZ[i] >= NZ ? 0 : z + CZ[i]);
        S* slot = A.f + (size_t)i * A.SS + pop_index(xn, yn, zn, A.PX, NY);
        const double fpost = to_real<S>(*slot, i);
        double rb[3], ub[3];
        link_velocity(sq, NX, NY, NZ, A.walls_y, A.walls_z, x + 0.5 * CX[i], y + 0.5 * CY[i], z + 0.5 * CZ[i],
                      rb, ub);
        const double cub = CX[i] * ub[0] + CY[i] * ub[1] + CZ[i] * ub[2];
        const double fb = fpost - 6.0 * W[i] * nrho[t] * cub;
        *slot = from_real<S>(fb, i);
        const double s = fpost + fb, dl = fpost - fb;
        const double dpx = s * CX[i] - dl * ub[0];
        const double dpy = s * CY[i] - dl * ub[1];
        const double dpz = s * CZ[i] - dl * ub[2];
        acc[0] += dpx; acc[1] += dpy; acc[2] += dpz;
        acc[3] += rb[1] * dpz - rb[2] * dpy;
        acc[4] += rb[2] * dpx - rb[0] * dpz;
        acc[5] += rb[0] * dpy - rb[1] * dpx;
        acc[6] += fb - fpost;
    }
    block_reduce<7>(acc, lds);
    double tot[7];
    if (last_block_sum<7>(A.part, &st->link_ticket, acc, tot, lds, (unsigned)block, (unsigned)nblocks) &&
        threadIdx.x == 0) {
        for (int d = 0; d < 3; ++d) { st->sq.Fh[d] = tot[d]; st->sq.Th[d] = tot[3 + d]; }
        st->link_dmass = tot[6];
    }
}

template <class S>
__global__ __launch_bounds__(LINK_BLOCK) void particle_links_kernel(const LinkArgs<S> A) {
    links_block<S>(A, (int)blockIdx.x, (int)gridDim.x);
}

// Wall momentum exchange (lbm.cpp:209, 219): for every WALL link the bounced
// population equals f*_i, so dp = 2 f*_i c_i.  Only the rows next to a wall
// have WALL links, and their wall directions follow from the row's
// coordinates (the walls are the planes y = 0, NY - 1 and z = 0, NZ - 1), so
// one wavefront takes one 64-node chunk of such a row: a flag load per node
// (the squirmer may cover part of the row) and one load per wall direction
// from the wall node's slot.  On request.
struct WallForceArgs {
    const uint32_t* wall_rows;  // y + NY * z of each row next to a wall
    int n_wall_rows, nchx;      // rows and 64-node chunks per row
    int walls_y, walls_z;
};

template <class S>
__global__ __launch_bounds__(256) void wall_force_kernel(const S* __restrict__ f, const uint8_t* __restrict__ flag,
                                                         const WallForceArgs W, double* part, DevState* st, int NX,
                                                         int NY, int NZ, uint32_t SS, uint32_t PX) {
    __shared__ double lds[4 * 3];
    const int wave = (int)(blockIdx.x * 4u + (threadIdx.x >> 6));
    double acc[3] = {0.0, 0.0, 0.0};
    if (wave < W.n_wall_rows * W.nchx) {
        const uint32_t r = W.wall_rows[wave / W.nchx];
        const int y = (int)(r % (uint32_t)NY), z = (int)(r / (uint32_t)NY);
        const int x = (wave % W.nchx) * 64 + (int)(threadIdx.x & 63);
        if (x < NX && (flag[(size_t)x + (size_t)NX * r] & TYPE_MASK) == FLUID) {
            const int xs[3] = {x == 0 ? NX - 1 : x - 1, x, x + 1 == NX ? 0 : x + 1};
            static_for<1, Q>([&](auto I) {
                constexpr int i = decltype(I)::value;
                const int yn = y + CY[i], zn = z + CZ[i];
                const bool wall = (W.walls_y && (yn == 0 || yn == NY - 1)) || (W.walls_z && (zn == 0 || zn == NZ - 1));
                if (wall) {  // wavefront-uniform
                    const int yw = yn < 0 ? NY - 1 : (yn >= NY ? 0 : yn), zw = zn < 0 ? NZ - 1 : (zn >= NZ ? 0 : zn);
                    const double fpost = to_real<S>(f[(size_t)i * SS + pop_index(xs[1 + CX[i]], yw, zw, PX, NY)], i);
                    const double s = fpost + fpost;
                    acc[0] += s * CX[i]; acc[1] += s * CY[i]; acc[2] += s * CZ[i];
                }
            });
        }
    }
    block_reduce<3>(acc, lds);
    double tot[3];
    if (last_block_sum<3>(part, &st->wall_ticket, acc, tot, lds, blockIdx.x, gridDim.x) && threadIdx.x == 0)
        for (int d = 0; d < 3; ++d) st->wallForce[d] = tot[d];
}

}  // namespace lbmhip
