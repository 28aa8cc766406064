// links_kernel.hpp -- the moving-boundary bounce-back and momentum exchange
// of PARTICLE links over a device-resident link list, and the wall momentum
// exchange, applied after the fused step to the slots where the step left
// the bounced populations.
//
// After either step parity the bounced population of a link (n, i) with a
// solid node n + c_i is in slot i of n + c_i, holding the post-collision
// value f*_i of node n (the wall bounce-back of lbm.cpp:197-199 with u_b = 0).
// The link pass turns it into
//     f_b = f*_i - 6 w_i rho_n (c_i . u_b)          (lbm.cpp:189-199)
// with u_b the squirmer surface velocity (rigid body + two-mode slip) at the
// link midpoint, and sums the momentum exchange (Wen et al. 2014)
//     dp = (f* + f_b) c_i - (f* - f_b) u_b,   torque r_b x dp   (lbm.cpp:202-215)
// in FP64.  The sums are deterministic: per thread in list order, per block
// in a fixed tree into one partial per block; the squirmer kernel that runs
// next adds the block partials in a fixed order and publishes Fh, Th
// (lbm.cpp:220-223).  No block waits for another, and the launch boundary
// makes the corrected populations visible to the remap.
//
// The link list.  The reference finds its PARTICLE links by testing the 18
// neighbour flags of every node of the sphere's box each step
// (lbm.cpp:179-199).  Here the links are kept in a list of (node, direction)
// entries that lives on the device with the lattice:
//   - link_build_block writes it from the flags (box scan; at creation and
//     whenever the host invalidated it);
//   - the remap appends the links its flag changes create (squirmer_kernel.hpp);
//   - the link pass checks each entry against the flags it loads anyway
//     (node FLUID, neighbour PARTICLE) and marks an entry whose link a remap
//     removed as dead (direction 0), so a dead entry never comes back to
//     life: a later remap that recreates the link appends a new entry;
//   - the remap compacts the list in place when dead entries pile up.
// Invariant: after a link pass, the live entries are exactly the links of
// the current flags, each once.  A remap keeps it if a link pass ran since
// the previous remap (the host tracks this and rebuilds the list otherwise).
#pragma once
#include "device_state.hpp"
#include "step_kernel.hpp"
#include "storage.hpp"

namespace lbmhip {

constexpr int CPL_THREADS = 1024;    // threads of the squirmer / remap block
constexpr int LINK_THREADS = 256;    // threads of a link-pass block
constexpr int LINK_MAX_BLOCKS = 8192;
constexpr uint32_t LINK_DEAD = 0u;   // direction field of a removed list entry

template <class S>
struct LinkArgs {
    S* f;
    const uint8_t* flag;
    const void* rho;     // density of this step (double in FP64 mode, float otherwise)
    DevState* st;
    double* part;        // LINK_MAX_BLOCKS x 8 partial sums
    uint2* list;         // PARTICLE links: {node, direction}
    uint32_t cap;        // entries the list can hold
    int NX, NY, NZ;
    uint32_t SS, PX;     // slot stride, population row pitch
    int walls_y, walls_z;
};

__device__ inline int wrap_once(int a, int n) {
    a %= n;
    return a < 0 ? a + n : a;
}
// a + d for d in {-1, 0, 1} on a periodic axis of length n, 0 <= a < n
__device__ inline int step1(int a, int d, int n) {
    const int b = a + d;
    return b < 0 ? n - 1 : (b >= n ? 0 : b);
}

// Exclusive block scan of per-thread counts over NT threads; `wsum` holds
// NT / 64 ints of LDS.  Ends with a barrier, so `wsum` may be reused.
template <int NT>
__device__ __forceinline__ int block_exclusive_scan(int cnt, int* wsum, int& total) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int incl = cnt;
    for (int off = 1; off < 64; off <<= 1) {
        const int v = __shfl_up(incl, off, 64);
        if (lane >= off) incl += v;
    }
    if (lane == 63) wsum[wave] = incl;
    __syncthreads();
    int base = 0;
    total = 0;
    for (int w = 0; w < NT / 64; ++w) {
        const int v = wsum[w];
        base += w < wave ? v : 0;
        total += v;
    }
    __syncthreads();
    return base + incl - cnt;
}

// Fixed-order combination of per-block partials by the last block to finish.
// Returns true in the last block; `out` then holds the sums (thread 0).
// `block` / `nblocks`: this block's index among the blocks that share the
// ticket and the partials.  `lds` holds (NT / 64) * K doubles.
template <int K, int NT = BLOCK>
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
    block_reduce<K, NT>(v, lds);
    if (threadIdx.x == 0) {
        for (int k = 0; k < K; ++k) out[k] = v[k];
        *ticket = 0u;  // ready for the next launch
    }
    return true;
}

// The link pass of one block.  The list is dealt to the blocks in chunks of
// 64 entries, round robin (chunk c to block c mod nblocks, then to its waves
// in turn), so a short list still spreads over every block of the launch,
// each wavefront taking a few chunks.  acc = {Fx, Fy, Fz, Tx, Ty, Tz, mass
// change, live entries}.
template <class S>
__device__ __forceinline__ void link_pass(const LinkArgs<S>& A, int block, int nblocks, double (&acc)[8]) {
    DevState* st = A.st;
    const DevSquirmer sq = st->sq;  // read once: the list and slot stores below may alias it
    const int NX = A.NX, NY = A.NY, NZ = A.NZ;
    const uint32_t uNX = (uint32_t)NX, uNXY = (uint32_t)NX * (uint32_t)NY;
    for (int k = 0; k < 8; ++k) acc[k] = 0.0;
    const uint32_t nl = st->nlinks;
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint32_t cstride = (uint32_t)nblocks * (LINK_THREADS / 64);  // chunks per round
    for (uint32_t c = (uint32_t)block + (uint32_t)nblocks * wave;; c += cstride) {
        const uint32_t q = c * 64u + lane;
        // the entry load does not wait for the length (entries past it are ignored)
        uint2 e = make_uint2(0u, LINK_DEAD);
        if (q < A.cap) e = A.list[q];
        if (c * 64u >= nl) break;  // wavefront-uniform
        const int i = (int)e.y;
        if (q >= nl || i == (int)LINK_DEAD) continue;
        const uint32_t n = e.x;
        const int x = (int)(n % uNX), y = (int)((n / uNX) % (uint32_t)NY), z = (int)(n / uNXY);
        const int xn = step1(x, CX[i], NX), yn = step1(y, CY[i], NY), zn = step1(z, CZ[i], NZ);
        const uint32_t nb = (uint32_t)xn + uNX * (uint32_t)yn + uNXY * (uint32_t)zn;
        S* slot = A.f + (size_t)i * A.SS + pop_index(xn, yn, zn, A.PX, NY);
        // four independent loads
        const uint8_t fn = A.flag[n], fb = A.flag[nb];
        double rho;
        if constexpr (std::is_same<S, double>::value) rho = static_cast<const double*>(A.rho)[n];
        else rho = (double)static_cast<const float*>(A.rho)[n];
        const S sv = *slot;
        if ((fn & TYPE_MASK) != FLUID || (fb & TYPE_MASK) != PARTICLE) {
            A.list[q].y = LINK_DEAD;  // a remap removed this link
            continue;
        }
        const double fpost = to_real<S>(sv, i);
        double rb[3], ub[3];
        link_velocity(sq, NX, NY, NZ, A.walls_y, A.walls_z, x + 0.5 * CX[i], y + 0.5 * CY[i], z + 0.5 * CZ[i], rb,
                      ub);
        const double cub = CX[i] * ub[0] + CY[i] * ub[1] + CZ[i] * ub[2];
        const double fbv = fpost - 6.0 * W[i] * rho * cub;
        *slot = from_real<S>(fbv, i);
        const double s = fpost + fbv, dl = fpost - fbv;
        const double dpx = s * CX[i] - dl * ub[0];
        const double dpy = s * CY[i] - dl * ub[1];
        const double dpz = s * CZ[i] - dl * ub[2];
        acc[0] += dpx; acc[1] += dpy; acc[2] += dpz;
        acc[3] += rb[1] * dpz - rb[2] * dpy;
        acc[4] += rb[2] * dpx - rb[0] * dpz;
        acc[5] += rb[0] * dpy - rb[1] * dpx;
        acc[6] += fbv - fpost;
        acc[7] += 1.0;
    }
}

// One block of the link-pass launch: its partial sums go to part[block].
template <class S>
__device__ __forceinline__ void link_pass_block(const LinkArgs<S>& A, int block, int nblocks) {
    __shared__ double lds[(LINK_THREADS / 64) * 8];
    double acc[8];
    link_pass<S>(A, block, nblocks, acc);
    block_reduce<8, LINK_THREADS>(acc, lds);
    if (threadIdx.x == 0)
        for (int k = 0; k < 8; ++k) A.part[(size_t)block * 8 + k] = acc[k];
}

template <class S>
__global__ __launch_bounds__(LINK_THREADS) void particle_links_kernel(const LinkArgs<S> A) {
    link_pass_block<S>(A, (int)blockIdx.x, (int)gridDim.x);
}

// Link-list build from the flags by one block of CPL_THREADS threads.  The
// PARTICLE nodes are the nodes within rf of xf (the classification the
// flags hold), so every link-owning FLUID node lies within rf + sqrt(2) of
// xf: the box [floor(xf - rf) - 2, ceil(xf + rf) + 2] per axis (one period
// at most) holds them all.  Entries in box order (z, y, x), directions
// ascending.  Sets nlinks and nvalid.
__device__ inline void link_build_block(const uint8_t* __restrict__ flag, DevState* st, uint2* list, uint32_t cap,
                                        int NX, int NY, int NZ) {
    __shared__ int wsum[CPL_THREADS / 64];
    const int Nd[3] = {NX, NY, NZ};
    int lo[3], wd[3];
    for (int d = 0; d < 3; ++d) {
        lo[d] = (int)floor(st->xf[d] - st->rf) - 2;
        const int hi = (int)ceil(st->xf[d] + st->rf) + 2;
        wd[d] = hi - lo[d] + 1 > Nd[d] ? Nd[d] : hi - lo[d] + 1;
    }
    const uint32_t M = (uint32_t)wd[0] * (uint32_t)wd[1] * (uint32_t)wd[2];
    const uint32_t uNX = (uint32_t)NX, uNXY = (uint32_t)NX * (uint32_t)NY;
    uint32_t nl = 0;
    bool over = false;
    for (uint32_t r0 = 0; r0 < M; r0 += CPL_THREADS) {
        const uint32_t k = r0 + threadIdx.x;
        uint32_t mask = 0, n = 0;
        if (k < M) {
            const int ix = (int)(k % (uint32_t)wd[0]), iy = (int)((k / (uint32_t)wd[0]) % (uint32_t)wd[1]),
                      iz = (int)(k / ((uint32_t)wd[0] * (uint32_t)wd[1]));
            const int x = wrap_once(lo[0] + ix, NX), y = wrap_once(lo[1] + iy, NY), z = wrap_once(lo[2] + iz, NZ);
            n = (uint32_t)x + uNX * (uint32_t)y + uNXY * (uint32_t)z;
            const uint8_t fl = flag[n];
            if ((fl & TYPE_MASK) == FLUID && (fl & BOUNDARY)) {
                static_for<1, Q>([&](auto I) {
                    constexpr int i = decltype(I)::value;
                    const uint32_t m = (uint32_t)step1(x, CX[i], NX) + uNX * (uint32_t)step1(y, CY[i], NY) +
                                       uNXY * (uint32_t)step1(z, CZ[i], NZ);
                    mask |= (uint32_t)((flag[m] & TYPE_MASK) == PARTICLE) << i;
                });
            }
        }
        int total;
        uint32_t pos = nl + (uint32_t)block_exclusive_scan<CPL_THREADS>(__builtin_popcount(mask), wsum, total);
        while (mask) {
            const int i = __builtin_ctz(mask);
            mask &= mask - 1;
            if (pos < cap) list[pos] = make_uint2(n, (uint32_t)i);
            ++pos;
        }
        nl += (uint32_t)total;
        if (nl > cap) over = true;
    }
    if (threadIdx.x == 0) {
        if (over) {
            atomicOr(&st->err, ERR_LINK_OVERFLOW);
            nl = 0;
        }
        st->nlinks = nl;
        st->nvalid = nl;
    }
}

// In-place compaction of the list by one block: live entries keep their
// order.  Returns the new length (every thread).  A round reads its entries
// before the scan's barrier and writes them to positions at or below their
// own, which this or an earlier round has read.
__device__ inline uint32_t link_compact_block(uint2* list, uint32_t nl) {
    __shared__ int wsum[CPL_THREADS / 64];
    uint32_t w = 0;
    for (uint32_t r0 = 0; r0 < nl; r0 += CPL_THREADS) {
        const uint32_t q = r0 + threadIdx.x;
        uint2 e = make_uint2(0u, LINK_DEAD);
        if (q < nl) e = list[q];
        const int keep = e.y != LINK_DEAD ? 1 : 0;
        int total;
        const int pos = block_exclusive_scan<CPL_THREADS>(keep, wsum, total);
        if (keep) list[w + (uint32_t)pos] = e;
        w += (uint32_t)total;
    }
    return w;
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
