// squirmer_kernel.hpp -- LBM::advanceSquirmer on the device (one block):
// soft wall potential, Newton-Euler update and hard gap in FP64 by one
// thread, then the remap of the sphere onto the lattice.
//
// The reference remap (src/utils/lbm.cpp:271-320) is a sequential sweep over
// the sphere's bounding box in z, y, x order, in place: an uncovered node's
// refill density averages the FLUID neighbours as they are at that point of
// the sweep, and the cover/uncover momentum sums are order dependent.  Here
//   B. all threads classify the box in parallel and compact the changed
//      nodes in sweep order (per-thread chunks + one block scan).  A node
//      can only change if it lies within |X - xf| of the sphere surface
//      (the flags are the inside test at xf), so only those nodes load
//      their flag;
//   C. all threads gather what the sweep reads at each changed node (its
//      populations, its neighbours' flags and densities before the remap,
//      and which neighbours are themselves in the list) into LDS; then one
//      wavefront walks the list in order from LDS, one lane per direction,
//      summing in the reference's order on every lane, so the FP64 results
//      are the reference's bit for bit;
//   D. all threads move populations whose AA storage slot depends on the
//      flags that just changed, write the equilibrium of uncovered nodes
//      into the slots the next step reads, and refresh BOUNDARY bits.
//
// Storage slot of the population f_j(n) that the next step reads
// ("loc(n, j)", see step_kernel.hpp): after an even step slot opp(j) at
// n - c_j whatever the flags, so no slot moves; after an odd step slot j at n
// when n - c_j is FLUID, else slot opp(j) at n - c_j.  After an odd step a
// flag change therefore moves the slots of the links between the changed
// node and its unchanged FLUID neighbours; phase D applies exactly those
// moves.  Every such link owns a pair of slots no other link touches, so
// phase D runs without ordering.
#pragma once
#include "device_state.hpp"
#include "links_kernel.hpp"
#include "step_kernel.hpp"
#include "storage.hpp"

namespace lbmhip {

constexpr int SQ_THREADS = CPL_THREADS;
constexpr int REMAP_CAP = 512;   // changed nodes per remap held in LDS
constexpr int GATHER = 128;      // changed nodes per gather pass of phase C
constexpr int BOX_MAX = 128;     // remap box columns per axis held in LDS

template <class S>
struct RemapArgs {
    S* f;
    uint8_t* flag;
    void* rho;
    void* ux;
    void* uy;
    void* uz;
    DevState* st;
    int NX, NY, NZ;
    uint32_t N;
    uint32_t SS;     // slot stride (elements between populations i and i+1)
    uint32_t PX;     // population row pitch
    int walls_y, walls_z;
    int shifted;     // layout after the last step: 1 after an even step
    int shift_slot;  // mass-shift slot the next step reads
    int accumulate;  // 1: another remap ran since the last step (add to remap_dmass, do not set it)
    uint2* list;     // PARTICLE link list (links_kernel.hpp)
    uint32_t cap;    // its capacity in entries
    int maintain;    // 1: append the links this remap creates (a link pass ran since the last remap)
};

// The momentum-exchange sums of the link pass that ran just before the
// remap block (thread 0: Fx Fy Fz Tx Ty Tz, mass change, live list entries)
// and the list length (every thread).
struct LinkSums {
    double v[8];
    uint32_t nl;
    int valid;
};

// squirmer.cpp:73-77
__device__ inline double soft_wall_force(double h, double eps, double FS, double hc) {
    if (h >= hc) return 0.0;
    const double hp = (h < 0.0) ? 0.0 : h;
    const double s = 1.0 - hp / hc;
    return eps * FS * s * s;
}

// lattice.hpp:42-47
__device__ inline double equilibrium(int i, double rho, double ux, double uy, double uz) {
    const double cu = CX[i] * ux + CY[i] * uy + CZ[i] * uz;
    const double usq = ux * ux + uy * uy + uz * uz;
    return W[i] * rho * (1.0 + 3.0 * cu + 4.5 * cu * cu - 1.5 * usq);
}

__device__ inline int wrapi(int a, int n) { return ((a % n) + n) % n; }

// Read a device-side macroscopic value as double.
template <class S>
__device__ inline double load_macro(const void* p, uint32_t n) {
    if constexpr (std::is_same<S, double>::value) return static_cast<const double*>(p)[n];
    else return (double)static_cast<const float*>(p)[n];
}
// The value a density stored by store_macro reads back as.
template <class S>
__device__ inline double macro_round(double v) {
    if constexpr (std::is_same<S, double>::value) return v;
    else return (double)(float)v;
}
template <class S>
__device__ inline void store_macro(void* p, uint32_t n, double v) {
    if constexpr (std::is_same<S, double>::value) static_cast<double*>(p)[n] = v;
    else static_cast<float*>(p)[n] = (float)v;
}

// Link-list maintenance after a remap (links_kernel.hpp), by the remap
// block once the flags are final and FLIPPED still marks the K changed
// nodes.  Every link a flag change creates involves a changed node n: its
// own links to PARTICLE neighbours when it was uncovered, or the links of
// its unchanged FLUID neighbours m = n - c_i into it when it was covered (a
// link between two changed nodes comes from the uncovered one only, so none
// is appended twice).  The links a change removes stay in the list until the
// next link pass marks them dead.  The list is compacted first when dead
// entries make up more than a fifth of it or the appends could overflow it.
// `gfl` / `gidx` (when the remap gathered all K nodes in one pass, else
// nullptr): for changed node k and direction i the type before the remap of
// the neighbour its gather looked at (n - c_i if covered, n + c_i if
// uncovered) and that neighbour's index in the list or -1; with them the new
// links follow from LDS without loading a flag.
template <class S>
__device__ void append_links(const RemapArgs<S>& A, const uint32_t* changed, const uint8_t* cover, int K,
                             const LinkSums& hs, const uint8_t (*gfl)[Q], const int16_t (*gidx)[Q]) {
    __shared__ int wsum[SQ_THREADS / 64];
    __shared__ uint32_t s_nv;
    DevState* st = A.st;
    // list length and live entries: from the link pass of this launch, or
    // as the last one left them
    if (threadIdx.x == 0) s_nv = hs.valid ? (uint32_t)hs.v[7] : st->nvalid;
    const uint32_t nl0 = hs.valid ? hs.nl : st->nlinks;
    __syncthreads();
    const uint32_t nv = s_nv;
    uint32_t nl = nl0;
    const bool compact = nl > nv + nv / 4u + 1024u || (size_t)nl + (size_t)(Q - 1) * (size_t)K > (size_t)A.cap;
    if (!compact && K == 0) return;  // uniform
    if (compact) nl = link_compact_block(A.list, nl);
    const int NX = A.NX, NY = A.NY, NZ = A.NZ;
    const uint32_t uNX = (uint32_t)NX, uNXY = (uint32_t)NX * (uint32_t)NY;
    for (int w0 = 0; w0 < K * (Q - 1); w0 += SQ_THREADS) {
        const int w = w0 + (int)threadIdx.x;
        int emit = 0;
        uint2 e = make_uint2(0u, LINK_DEAD);
        if (w < K * (Q - 1)) {
            const int k = w / (Q - 1), i = 1 + w - k * (Q - 1);
            const uint32_t n = changed[k];
            const int x = (int)(n % uNX), y = (int)((n / uNX) % (uint32_t)NY), z = (int)(n / uNXY);
            if (!cover[k]) {  // uncovered, now FLUID: is n + c_i PARTICLE now?
                bool part;
                if (gfl) {
                    const int j = gidx[k][i];
                    part = j >= 0 ? cover[j] != 0 : gfl[k][i] == PARTICLE;
                } else {
                    const uint32_t nb = (uint32_t)step1(x, CX[i], NX) + uNX * (uint32_t)step1(y, CY[i], NY) +
                                        uNXY * (uint32_t)step1(z, CZ[i], NZ);
                    part = (A.flag[nb] & TYPE_MASK) == PARTICLE;
                }
                if (part) { emit = 1; e = make_uint2(n, (uint32_t)i); }
            } else {  // covered, now PARTICLE: is n - c_i an unchanged FLUID node?
                const uint32_t m = (uint32_t)step1(x, -CX[i], NX) + uNX * (uint32_t)step1(y, -CY[i], NY) +
                                   uNXY * (uint32_t)step1(z, -CZ[i], NZ);
                bool fluid;
                if (gfl) {
                    fluid = gidx[k][i] < 0 && gfl[k][i] == FLUID;
                } else {
                    const uint8_t fm = A.flag[m];
                    fluid = (fm & TYPE_MASK) == FLUID && !(fm & FLIPPED);
                }
                if (fluid) { emit = 1; e = make_uint2(m, (uint32_t)i); }
            }
        }
        int total;
        const int pos = block_exclusive_scan<SQ_THREADS>(emit, wsum, total);
        if (emit && nl + (uint32_t)pos < A.cap) A.list[nl + (uint32_t)pos] = e;
        nl += (uint32_t)total;
    }
    if (threadIdx.x == 0) {
        if (nl > A.cap) {
            atomicOr(&st->err, ERR_LINK_OVERFLOW);
            nl = A.cap;
        }
        st->nlinks = nl;
    }
}

// `part` / `nparts`: block partials of the link pass that ran just before
// (nullptr: none; Fh, Th are those published earlier).
template <class S>
__device__ void squirmer_block(const RemapArgs<S>& A, const double* part, int nparts) {
    __shared__ uint32_t list[REMAP_CAP];  // changed nodes in sweep order
    __shared__ uint8_t lcov[REMAP_CAP];   // 1: the node is covered, 0: uncovered
    __shared__ int lpos[REMAP_CAP];       // box position of each entry (ascending)
    __shared__ double lrr[REMAP_CAP], lub[REMAP_CAP][3];
    __shared__ double gval[GATHER][Q];    // cover: f_i(n); uncover: rho of neighbour i
    __shared__ uint8_t gfl[GATHER][Q];    // neighbour i's type before the remap
    __shared__ int16_t gidx[GATHER][Q];   // neighbour i's list index, or -1
    __shared__ double sprune[2];          // candidate band of squared distances
    __shared__ double srv[Q];             // refill: density of neighbour i (FLUID ones)
    __shared__ int sisf[Q];               // refill: neighbour i is FLUID
    __shared__ int sbox[6];  // x0, y0, z0, wx, wy, wz
    __shared__ int scount[SQ_THREADS / 64];
    __shared__ int sK, sOverflow;
    __shared__ double boxr[3][BOX_MAX];
    __shared__ int boxc[3][BOX_MAX];

    __shared__ DevSquirmer sq;  // working copy; written back at the end
    DevState* st = A.st;
    const int tid = threadIdx.x;
    const int NX = A.NX, NY = A.NY, NZ = A.NZ;
    const uint32_t SS = A.SS;

    // ---------------- A: rigid body (lbm.cpp:257-262) ----------------
    {
        // one coalesced copy of the squirmer state into LDS, with the loads
        // of the link partials
        const double* src = reinterpret_cast<const double*>(&st->sq);
        double* dst = reinterpret_cast<double*>(&sq);
        constexpr int nd = (int)(sizeof(DevSquirmer) / sizeof(double));
        static_assert(sizeof(DevSquirmer) % sizeof(double) == 0, "DevSquirmer must be a whole number of doubles");
        if (tid < nd) dst[tid] = src[tid];
    }
    LinkSums hs;
    hs.valid = part != nullptr;
    hs.nl = 0;
    if (part) {
        // the first wavefront adds them, thread 0 (which uses them) among
        // its lanes: partial b on lane b mod 64 in order, then a fixed tree
        hs.nl = st->nlinks;
        if (tid < 64) {
            double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            for (int b = tid; b < nparts; b += 64)
                for (int k = 0; k < 8; ++k) v[k] += part[(size_t)b * 8 + k];
            for (int k = 0; k < 8; ++k) {
                double x = v[k];
                for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
                hs.v[k] = x;
            }
        }
    }
    __syncthreads();
    if (tid == 0) {
        // sq.Fh / sq.Th: the link pass of this launch, or published by an
        // earlier one (lbm.cpp:220-223)
        if (hs.valid) {
            for (int d = 0; d < 3; ++d) { sq.Fh[d] = hs.v[d]; sq.Th[d] = hs.v[3 + d]; }
            st->link_dmass = hs.v[6];
            st->nvalid = (unsigned)hs.v[7];
        }
        // applyWallPotential (lbm.cpp:231-241)
        sq.Fw[0] = sq.Fw[1] = sq.Fw[2] = 0.0;
        if (A.walls_y) {
            sq.Fw[1] += soft_wall_force(sq.X[1] - sq.R - 0.5, sq.eps, sq.FS, sq.hc);
            sq.Fw[1] -= soft_wall_force((NY - 1.5) - sq.X[1] - sq.R, sq.eps, sq.FS, sq.hc);
        }
        if (A.walls_z) {
            sq.Fw[2] += soft_wall_force(sq.X[2] - sq.R - 0.5, sq.eps, sq.FS, sq.hc);
            sq.Fw[2] -= soft_wall_force((NZ - 1.5) - sq.X[2] - sq.R, sq.eps, sq.FS, sq.hc);
        }
        // Squirmer::integrate (squirmer.cpp:52-68)
        for (int d = 0; d < 3; ++d) {
            sq.U[d] += (sq.Fh[d] + sq.Fc[d] + sq.Fw[d]) / sq.M;
            sq.W[d] += (sq.Th[d] + sq.Tc[d]) / sq.I;
            sq.Fc[d] = sq.Tc[d] = 0.0;
        }
        if (sq.planar) {
            if (sq.plane_xz) { sq.U[1] = 0.0; sq.W[0] = 0.0; sq.W[2] = 0.0; }
            else { sq.U[2] = 0.0; sq.W[0] = 0.0; sq.W[1] = 0.0; }
        }
        for (int d = 0; d < 3; ++d) sq.X[d] += sq.U[d];
        double we[3];
        we[0] = sq.W[1] * sq.e[2] - sq.W[2] * sq.e[1];
        we[1] = sq.W[2] * sq.e[0] - sq.W[0] * sq.e[2];
        we[2] = sq.W[0] * sq.e[1] - sq.W[1] * sq.e[0];
        for (int d = 0; d < 3; ++d) sq.e[d] += we[d];
        if (sq.planar) {
            if (sq.plane_xz) sq.e[1] = 0.0;
            else sq.e[2] = 0.0;
        }
        const double nrm = sqrt(sq.e[0] * sq.e[0] + sq.e[1] * sq.e[1] + sq.e[2] * sq.e[2]);
        for (int d = 0; d < 3; ++d) sq.e[d] /= nrm;
        // enforceMinimumGap (lbm.cpp:245-253)
        auto clamp_axis = [&](int d, int Nd) {
            const double lo = 0.5 + sq.R + sq.hmin, hi = (Nd - 1.5) - sq.R - sq.hmin;
            if (sq.X[d] < lo) { sq.X[d] = lo; if (sq.U[d] < 0) sq.U[d] = 0; ++sq.contacts; }
            if (sq.X[d] > hi) { sq.X[d] = hi; if (sq.U[d] > 0) sq.U[d] = 0; ++sq.contacts; }
        };
        if (A.walls_y) clamp_axis(1, NY);
        if (A.walls_z) clamp_axis(2, NZ);

        // remap box (lbm.cpp:273-276); an axis longer than the domain is cut
        // to one period, since later periodic copies of a node are no-ops in
        // the reference's in-place sweep.
        const double R = sq.R;
        const int x0 = (int)floor(sq.X[0] - R) - 2, x1 = (int)ceil(sq.X[0] + R) + 2;
        int y0 = (int)floor(sq.X[1] - R) - 2, y1 = (int)ceil(sq.X[1] + R) + 2;
        const int z0 = (int)floor(sq.X[2] - R) - 2, z1 = (int)ceil(sq.X[2] + R) + 2;
        if (A.walls_y) { y0 = y0 < 0 ? 0 : y0; y1 = y1 > NY - 1 ? NY - 1 : y1; }
        int wx = x1 - x0 + 1, wy = y1 - y0 + 1, wz = z1 - z0 + 1;
        wx = wx > NX ? NX : wx; wy = wy > NY ? NY : wy; wz = wz > NZ ? NZ : wz;
        wy = wy < 0 ? 0 : wy;
        sbox[0] = x0; sbox[1] = y0; sbox[2] = z0; sbox[3] = wx; sbox[4] = wy; sbox[5] = wz;
        sOverflow = 0;
        // Only nodes whose distance from X is within |X - xf| of the sphere
        // surface can change (triangle inequality).  Applied when the box
        // holds no periodic image twice and the motion is below one node.
        double dl2 = 0.0;
        for (int d = 0; d < 3; ++d) { const double dd = sq.X[d] - st->xf[d]; dl2 += dd * dd; }
        const double dl = sqrt(dl2), rf = st->rf, margin = 1e-6;
        const bool cut = (wx > NX - 4) || (!A.walls_y && wy > NY - 4) || (!A.walls_z && wz > NZ - 4);
        if (dl < 1.0 && !cut) {
            const double lo = fmin(R, rf - dl) - margin, hi = fmax(R, rf + dl) + margin;
            sprune[0] = lo > 0.0 ? lo * lo : -1.0;
            sprune[1] = hi * hi;
        } else {
            sprune[0] = -1.0;
            sprune[1] = __builtin_inf();
        }
    }
    __syncthreads();

    // ---------------- B: classify the box in sweep order ----------------
    // Per-axis minimum-image offsets (lbm.cpp:100-104) and wrapped
    // coordinates of the box columns are computed once into LDS; a node's
    // inside test is then (rx^2 + ry^2) + rz^2 < R^2 with the same doubles
    // the reference forms.  Each thread owns a contiguous chunk of box
    // positions (sweep order), issues its flag loads in batches, and keeps
    // change / cover bits in registers for the compaction pass.
    const int bx0 = sbox[0], by0 = sbox[1], bz0 = sbox[2], bwx = sbox[3], bwy = sbox[4], bwz = sbox[5];
    if (bwx > BOX_MAX || bwy > BOX_MAX || bwz > BOX_MAX) {
        if (tid == 0) {
            atomicOr(&st->err, ERR_REMAP_OVERFLOW);
            st->sq = sq;
        }
        return;  // uniform
    }
    for (int k = tid; k < bwx + bwy + bwz; k += SQ_THREADS) {
        int d = 0, j = k;
        if (j >= bwx) { j -= bwx; d = 1; if (j >= bwy) { j -= bwy; d = 2; } }
        const int Nd = d == 0 ? NX : (d == 1 ? NY : NZ);
        const int c = (d == 0 ? bx0 : (d == 1 ? by0 : bz0)) + j;
        const int cw = wrapi(c, Nd);
        double r = cw - sq.X[d];
        const bool per = d == 0 || (d == 1 ? !A.walls_y : !A.walls_z);
        if (per) r -= Nd * round(r / Nd);
        boxr[d][j] = r;
        boxc[d][j] = cw;
    }
    __syncthreads();
    // One thread per (y, z) column of the box, columns in sweep order, in
    // rounds of SQ_THREADS columns.  With the candidate band [lo2, hi2] of
    // squared distances, a column's candidates are the x positions with
    // |rx| in [sqrt(lo2 - s), sqrt(hi2 - s)], s = ry^2 + rz^2: at most two
    // short intervals (rx = rx(0) + ix, the box not being cut by the period).
    // Without a band the whole column is a candidate.  Candidates load their
    // flag and take the exact inside test; a pass counts, a block scan
    // places, a second pass writes the list.
    const double R2 = sq.R * sq.R;
    const uint32_t uNX = (uint32_t)NX, uNXY = (uint32_t)NX * (uint32_t)NY;
    auto node_at = [&](int ix, int iy, int iz) -> uint32_t {
        return (uint32_t)boxc[0][ix] + uNX * (uint32_t)boxc[1][iy] + uNXY * (uint32_t)boxc[2][iz];
    };
    const double lo2 = sprune[0], hi2 = sprune[1];
    const bool banded = hi2 < __builtin_inf();
    const int lane = tid & 63, wave = tid >> 6;
    const int ncol = bwy * bwz;
    int K = 0;  // list length so far (uniform)
    for (int r0 = 0; r0 < ncol; r0 += SQ_THREADS) {
        const int col = r0 + tid;
        int iy = 0, iz = 0, a1 = 0, b1 = -1, a2 = 0, b2 = -1;
        if (col < ncol) {
            iy = col % bwy; iz = col / bwy;
            if (!banded) {
                b1 = bwx - 1;
            } else {
                const double ry = boxr[1][iy], rz = boxr[2][iz];
                const double sc = ry * ry + rz * rz;
                if (sc <= hi2) {
                    const double xb = sqrt(hi2 - sc), xa = lo2 > sc ? sqrt(lo2 - sc) : 0.0;
                    const double off = boxr[0][0];
                    const double eps = 1e-7;
                    a1 = (int)ceil(-xb - off - eps); b1 = (int)floor(-xa - off + eps);
                    a2 = (int)ceil(xa - off - eps);  b2 = (int)floor(xb - off + eps);
                    if (b1 >= a2) { b1 = b2; b2 = -1; }  // the two intervals meet
                    a1 = a1 < 0 ? 0 : a1; b1 = b1 > bwx - 1 ? bwx - 1 : b1;
                    a2 = a2 < 0 ? 0 : a2; b2 = b2 > bwx - 1 ? bwx - 1 : b2;
                    if (b1 >= a2 && b2 >= a2) a2 = b1 + 1;  // no position twice
                }
            }
        }
        // visit the candidates of this column in x order
        auto visit = [&](auto&& on_change) {
            const double ry = boxr[1][iy], rz = boxr[2][iz];
            for (int pass = 0; pass < 2; ++pass) {
                const int lo = pass ? a2 : a1, hi = pass ? b2 : b1;
                for (int i0 = lo; i0 <= hi; i0 += 4) {  // flag loads four at a time
                    uint8_t fv[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) fv[j] = i0 + j <= hi ? A.flag[node_at(i0 + j, iy, iz)] : (uint8_t)WALL;
#pragma unroll
                    for (int j = 0; j < 4; ++j) {
                        const uint8_t t = fv[j] & TYPE_MASK;
                        if (t == WALL) continue;
                        const double rx = boxr[0][i0 + j];
                        const bool inside = rx * rx + ry * ry + rz * rz < R2;  // lbm.cpp:287
                        if (inside != (t == PARTICLE)) on_change(i0 + j, inside);
                    }
                }
            }
        };
        int cnt = 0;
        visit([&](int, bool) { ++cnt; });
        // block exclusive scan of cnt
        int incl = cnt;
        for (int off = 1; off < 64; off <<= 1) {
            const int v = __shfl_up(incl, off, 64);
            if (lane >= off) incl += v;
        }
        if (lane == 63) scount[wave] = incl;
        __syncthreads();
        if (tid == 0) {
            int tot = K;
            for (int w = 0; w < SQ_THREADS / 64; ++w) { const int c = scount[w]; scount[w] = tot; tot += c; }
            sK = tot;
            if (tot > REMAP_CAP) { sOverflow = 1; atomicOr(&st->err, ERR_REMAP_OVERFLOW); }
        }
        __syncthreads();
        if (sOverflow) {  // uniform: the host reports the error
            if (tid == 0) st->sq = sq;
            return;
        }
        int pos = scount[wave] + incl - cnt;
        visit([&](int ix, bool inside) {
            lpos[pos] = ix + bwx * (iy + bwy * iz);
            lcov[pos] = inside ? 1 : 0;
            list[pos++] = node_at(ix, iy, iz);
        });
        K = sK;
        __syncthreads();  // scount / sK reused by the next round
    }

    // ---------------- C: cover / uncover in sweep order ----------------
    double Fc[3] = {0, 0, 0}, Tc[3] = {0, 0, 0}, dmass = 0.0;
    long long dnf = 0;
    const double dsh = st->shift[A.shift_slot];
    const bool wrapx = bwx == NX, wrapy = bwy == NY, wrapz = bwz == NZ;
    for (int c0 = 0; c0 < K; c0 += GATHER) {
        const int nc = (K - c0 < GATHER) ? K - c0 : GATHER;
        // C1: gather.  For direction i > 0 each work item looks at one
        // neighbour: the upstream n - c_i of a covered node (where its f_i
        // comes from), the downstream n + c_i of an uncovered one (a density
        // the refill may average): its type before the remap (flags changed
        // by an earlier pass carry FLIPPED, which gives it back) and its
        // index in the list if it changes too.  Every load of an item is
        // issued at once (both slots that may hold f_i, the density whatever
        // the type) and the type picks afterwards.
        for (int w = tid; w < nc * Q; w += SQ_THREADS) {
            const int kk = w / Q, i = w - kk * Q, k = c0 + kk;
            const bool cover = lcov[k] != 0;
            const int p = lpos[k];
            const int ix = p % bwx, iy = (p / bwx) % bwy, iz = p / (bwx * bwy);
            const int x = boxc[0][ix], y = boxc[1][iy], z = boxc[2][iz];
            if (i == 0) {
                if (cover) {  // rest population f_0(n), lbm.cpp:295-298
                    double fv = to_real<S>(A.f[pop_index(x, y, z, A.PX, NY)], 0);
                    if (dsh != 0.0) fv += W[0] * dsh;
                    gval[kk][0] = fv;
                }
                continue;
            }
            const int sg = cover ? -1 : 1;
            const int xm = step1(x, sg * CX[i], NX), ym = step1(y, sg * CY[i], NY), zm = step1(z, sg * CZ[i], NZ);
            const uint32_t m = (uint32_t)xm + (uint32_t)NX * ((uint32_t)ym + (uint32_t)NY * (uint32_t)zm);
            const uint8_t fm = A.flag[m];
            S own{}, far{};
            double rho_m = 0.0;
            if (cover) {
                own = A.f[(size_t)i * SS + pop_index(x, y, z, A.PX, NY)];
                far = A.f[(size_t)OPP[i] * SS + pop_index(xm, ym, zm, A.PX, NY)];
            } else {
                rho_m = load_macro<S>(A.rho, m);
            }
            uint8_t tm = fm & TYPE_MASK;
            if (fm & FLIPPED) tm = (tm == FLUID) ? PARTICLE : FLUID;
            gfl[kk][i] = tm;
            if (cover) {
                // population f_i(n) after the last step, lbm.cpp:295-298: after an
                // odd step in slot i at n when n - c_i is FLUID, else (and after
                // an even step) in slot opp(i) at n - c_i
                const bool up_fluid = !A.shifted && tm == FLUID;
                double fv = to_real<S>(up_fluid ? own : far, i);
                if (dsh != 0.0) fv += W[i] * dsh;
                gval[kk][i] = fv;
            } else {
                gval[kk][i] = tm == FLUID ? rho_m : 0.0;
            }
            int jx = ix + sg * CX[i], jy = iy + sg * CY[i], jz = iz + sg * CZ[i];
            if (wrapx) jx = jx < 0 ? bwx - 1 : (jx == bwx ? 0 : jx);
            if (wrapy) jy = jy < 0 ? bwy - 1 : (jy == bwy ? 0 : jy);
            if (wrapz) jz = jz < 0 ? bwz - 1 : (jz == bwz ? 0 : jz);
            int j = -1;
            if (jx >= 0 && jx < bwx && jy >= 0 && jy < bwy && jz >= 0 && jz < bwz) {
                const int q = jx + bwx * (jy + bwy * jz);
                int a = 0, bnd = K;  // lower bound of q in lpos[0, K)
                while (a < bnd) {
                    const int mid = (a + bnd) >> 1;
                    if (lpos[mid] < q) a = mid + 1; else bnd = mid;
                }
                if (a < K && lpos[a] == q) j = a;
            }
            gidx[kk][i] = (int16_t)j;
        }
        __syncthreads();
        // C2: the sweep, from LDS, by the first wavefront: every lane forms the
        // same sums in the reference's order from broadcast LDS reads
        if (wave == 0) {
            const int i = lane;  // direction handled by this lane (lanes >= Q idle)
            for (int kk = 0; kk < nc; ++kk) {
                const int k = c0 + kk;
                const uint32_t n = list[k];
                const bool cover = lcov[k] != 0;
                const int p = lpos[k];
                const int ix = p % bwx, iy = (p / bwx) % bwy, iz = p / (bwx * bwy);
                const double r[3] = {boxr[0][ix], boxr[1][iy], boxr[2][iz]};  // lbm.cpp:100-104
                if (cover) {
                    double pm[3] = {0, 0, 0}, msum = 0.0;
                    for (int j = 0; j < Q; ++j) {
                        const double v = gval[kk][j];
                        pm[0] += v * CX[j]; pm[1] += v * CY[j]; pm[2] += v * CZ[j];
                        msum += v;
                    }
                    const double t[3] = {r[1] * pm[2] - r[2] * pm[1], r[2] * pm[0] - r[0] * pm[2],
                                         r[0] * pm[1] - r[1] * pm[0]};
                    for (int d = 0; d < 3; ++d) { Fc[d] += pm[d]; Tc[d] += t[d]; }
                    dmass -= msum - 1.0;
                    --dnf;
                    if (lane == 0) A.flag[n] = (uint8_t)(PARTICLE | FLIPPED);
                } else {
                    // refill density from FLUID neighbours as of this point of the sweep
                    double rv = 0.0;
                    int isf = 0;
                    if (i >= 1 && i < Q) {
                        const int j = gidx[kk][i];
                        if (j >= 0 && j < k) {  // changed earlier in the sweep
                            isf = lcov[j] == 0;
                            if (isf) rv = macro_round<S>(lrr[j]);
                        } else {
                            isf = gfl[kk][i] == FLUID;
                            if (isf) rv = gval[kk][i];
                        }
                    }
                    if (i >= 1 && i < Q) { srv[i] = rv; sisf[i] = isf; }
                    __atomic_signal_fence(__ATOMIC_SEQ_CST);
                    __builtin_amdgcn_wave_barrier();
                    double rs = 0.0;
                    int c = 0;
                    for (int j = 1; j < Q; ++j)
                        if (sisf[j]) { rs += srv[j]; ++c; }
                    const double rr = c ? rs / c : 1.0;
                    double ub[3];  // Squirmer::surfaceVelocity(r, ub, false)
                    ub[0] = sq.U[0] + (sq.W[1] * r[2] - sq.W[2] * r[1]);
                    ub[1] = sq.U[1] + (sq.W[2] * r[0] - sq.W[0] * r[2]);
                    ub[2] = sq.U[2] + (sq.W[0] * r[1] - sq.W[1] * r[0]);
                    double fsum = 0.0;
                    for (int j = 0; j < Q; ++j) fsum += equilibrium(j, rr, ub[0], ub[1], ub[2]);
                    const double pm[3] = {rr * ub[0], rr * ub[1], rr * ub[2]};
                    const double t[3] = {r[1] * pm[2] - r[2] * pm[1], r[2] * pm[0] - r[0] * pm[2],
                                         r[0] * pm[1] - r[1] * pm[0]};
                    for (int d = 0; d < 3; ++d) { Fc[d] -= pm[d]; Tc[d] -= t[d]; }
                    dmass += fsum - 1.0;
                    ++dnf;
                    if (lane == 0) {
                        store_macro<S>(A.rho, n, rr);
                        store_macro<S>(A.ux, n, ub[0]);
                        store_macro<S>(A.uy, n, ub[1]);
                        store_macro<S>(A.uz, n, ub[2]);
                        A.flag[n] = (uint8_t)(FLUID | FLIPPED);
                        lrr[k] = rr; lub[k][0] = ub[0]; lub[k][1] = ub[1]; lub[k][2] = ub[2];
                    }
                }
                // lrr[k] is read by later iterations on other lanes of this
                // wave: keep the compiler from reordering across iterations
                __atomic_signal_fence(__ATOMIC_SEQ_CST);
                __builtin_amdgcn_wave_barrier();
            }
        }
        __syncthreads();
    }
    if (wave == 0) {
        if (lane == 0) {
            for (int d = 0; d < 3; ++d) { sq.Fc[d] += Fc[d]; sq.Tc[d] += Tc[d]; }
            st->remap_dmass = (A.accumulate ? st->remap_dmass : 0.0) + dmass;
            st->nfluid += dnf;
            st->changed = K;
        }
    }
    __syncthreads();
    {
        const double* src = reinterpret_cast<const double*>(&sq);
        double* dst = reinterpret_cast<double*>(&st->sq);
        for (int k = tid; k < (int)(sizeof(DevSquirmer) / sizeof(double)); k += SQ_THREADS) dst[k] = src[k];
    }
    __syncthreads();

    // ---------------- D: storage moves and equilibrium refill ----------------
    S* __restrict__ f = A.f;
    for (int w = tid; w < K * Q; w += SQ_THREADS) {
        const int k = w / Q, i = w - k * Q;
        const uint32_t n = list[k];
        const bool cover = lcov[k] != 0;
        const int x = (int)(n % (uint32_t)NX), y = (int)((n / (uint32_t)NX) % (uint32_t)NY),
                  z = (int)(n / ((uint32_t)NX * (uint32_t)NY));
        const uint32_t pn = pop_index(x, y, z, A.PX, NY);
        if (i == 0) {
            if (!cover) f[pn] = from_real<S>(equilibrium(0, lrr[k], lub[k][0], lub[k][1], lub[k][2]), 0);
            continue;
        }
        const int xm = step1(x, -CX[i], NX), ym = step1(y, -CY[i], NY), zm = step1(z, -CZ[i], NZ);
        const uint32_t m = (uint32_t)xm + (uint32_t)NX * ((uint32_t)ym + (uint32_t)NY * (uint32_t)zm);
        const size_t own = (size_t)i * SS + pn, far = (size_t)OPP[i] * SS + pop_index(xm, ym, zm, A.PX, NY);
        if (A.shifted) {  // loc(n, i) = slot opp(i) at n - c_i: nothing moves
            if (!cover) f[far] = from_real<S>(equilibrium(i, lrr[k], lub[k][0], lub[k][1], lub[k][2]), i);
            continue;
        }
        const uint8_t fm = A.flag[m];
        const bool mfluid = (fm & TYPE_MASK) == FLUID, mflip = (fm & FLIPPED) != 0;
        if (!cover) {
            const S fe = from_real<S>(equilibrium(i, lrr[k], lub[k][0], lub[k][1], lub[k][2]), i);
            if (mfluid && !mflip) { f[far] = f[own]; f[own] = fe; }  // m's bounced f_opp(i) back to m
            else if (mfluid) f[own] = fe;
            else f[far] = fe;
        } else {
            if (mfluid && !mflip) f[own] = f[far];  // m's f_opp(i) now bounces off n
        }
    }
    __syncthreads();
    if (A.maintain) append_links(A, list, lcov, K, hs, K <= GATHER ? gfl : nullptr, gidx);
    for (int k = tid; k < K; k += SQ_THREADS) {
        const uint32_t n = list[k];
        A.flag[n] = A.flag[n] & TYPE_MASK;
    }
    __syncthreads();
    // BOUNDARY bits of the changed nodes and their neighbours
    for (int w = tid; w < K * Q; w += SQ_THREADS) {
        const int k = w / Q, i = w - k * Q;
        const uint32_t n0 = list[k];
        const int x0 = (int)(n0 % (uint32_t)NX), y0 = (int)((n0 / (uint32_t)NX) % (uint32_t)NY),
                  z0 = (int)(n0 / ((uint32_t)NX * (uint32_t)NY));
        const int x = step1(x0, CX[i], NX), y = step1(y0, CY[i], NY), z = step1(z0, CZ[i], NZ);
        const uint32_t m = (uint32_t)x + (uint32_t)NX * ((uint32_t)y + (uint32_t)NY * (uint32_t)z);
        // the node's flag and its neighbours' in one round of loads (only the
        // type bits are read, which no thread changes here)
        uint8_t nb[Q];
#pragma unroll
        for (int j = 0; j < Q; ++j) {
            const uint32_t mm = (uint32_t)step1(x, CX[j], NX) +
                                (uint32_t)NX * ((uint32_t)step1(y, CY[j], NY) + (uint32_t)NY * (uint32_t)step1(z, CZ[j], NZ));
            nb[j] = A.flag[mm];
        }
        const uint8_t t = nb[0] & TYPE_MASK;
        bool b = false;
#pragma unroll
        for (int j = 1; j < Q; ++j) b |= (nb[j] & TYPE_MASK) != FLUID;
        A.flag[m] = (uint8_t)(t | (t == FLUID && b ? BOUNDARY : 0));
    }
    if (tid == 0) {
        // rho band the next step writes: the next remap box plus neighbours,
        // with a margin for one step of motion (|U| < 1)
        const double R = sq.R;
        int lo[3], wd[3];
        const int Nd[3] = {NX, NY, NZ};
        for (int d = 0; d < 3; ++d) {
            const int a = (int)floor(sq.X[d] - R) - 4, b = (int)ceil(sq.X[d] + R) + 4;
            int wdt = b - a + 1;
            wdt = wdt > Nd[d] ? Nd[d] : wdt;
            lo[d] = wrapi(a, Nd[d]);
            wd[d] = wdt;
        }
        for (int d = 0; d < 3; ++d) { st->band[d] = lo[d]; st->band[3 + d] = wd[d]; }
        for (int d = 0; d < 3; ++d) st->xf[d] = sq.X[d];
        st->rf = sq.R;
    }
}

// The squirmer block of a step: with `links`, first the momentum-exchange
// sums of the link pass that ran just before it (`nparts` block partials,
// added in a fixed order); then, with `remap`, the rigid-body update, remap
// and link-list upkeep.  Without `remap` it only publishes Fh, Th and the
// link mass change (the link pass of a step whose remap comes later or not
// at all).
template <class S>
__device__ __forceinline__ void couple_block(const LinkArgs<S>& L, const RemapArgs<S>& R, int nparts, int links,
                                             int remap) {
    if (remap) {
        squirmer_block<S>(R, links ? L.part : nullptr, nparts);
        return;
    }
    if (!links) return;
    __shared__ double lds[(CPL_THREADS / 64) * 8];
    double v[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = (int)threadIdx.x; b < nparts; b += CPL_THREADS)
        for (int k = 0; k < 8; ++k) v[k] += L.part[(size_t)b * 8 + k];
    block_reduce<8, CPL_THREADS>(v, lds);
    if (threadIdx.x == 0) {
        DevState* st = L.st;
        for (int d = 0; d < 3; ++d) { st->sq.Fh[d] = v[d]; st->sq.Th[d] = v[3 + d]; }
        st->link_dmass = v[6];
        st->nvalid = (unsigned)v[7];
    }
}

template <class S>
__global__ __launch_bounds__(CPL_THREADS) void couple_kernel(const LinkArgs<S> L, const RemapArgs<S> R, int nparts,
                                                             int links, int remap) {
    couple_block<S>(L, R, nparts, links, remap);
}

// ---------------------------------------------------------------------------
// Logical population f_i(n) for download / upload (either layout; for
// non-fluid nodes slot i at n), for the nodes n0 .. n0 + cnt - 1: value i of
// node n0 + k at out[i * cnt + k] (in[i * cnt + k]).
// ---------------------------------------------------------------------------
template <class S>
__global__ __launch_bounds__(256) void gather_kernel(const S* __restrict__ f, const uint8_t* __restrict__ flag,
                                                     double* __restrict__ out, int NX, int NY, int NZ, uint32_t n0,
                                                     uint32_t cnt, uint32_t SS, uint32_t PX, int shifted,
                                                     const DevState* st, int shift_slot) {
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k >= cnt) return;
    const uint32_t n = n0 + k;
    const int x = (int)(n % (uint32_t)NX), y = (int)((n / (uint32_t)NX) % (uint32_t)NY),
              z = (int)(n / ((uint32_t)NX * (uint32_t)NY));
    const bool fluid = (flag[n] & TYPE_MASK) == FLUID;
    const double dsh = st->shift[shift_slot];
    for (int i = 0; i < Q; ++i) {
        size_t slot = (size_t)i * SS + pop_index(x, y, z, PX, NY);
        if (fluid && i > 0) {
            const int xm = wrapi(x - CX[i], NX), ym = wrapi(y - CY[i], NY), zm = wrapi(z - CZ[i], NZ);
            const uint32_t m = (uint32_t)xm + (uint32_t)NX * ((uint32_t)ym + (uint32_t)NY * (uint32_t)zm);
            if (shifted || (flag[m] & TYPE_MASK) != FLUID) slot = (size_t)OPP[i] * SS + pop_index(xm, ym, zm, PX, NY);
        }
        double v = to_real<S>(f[slot], i);
        if (fluid && dsh != 0.0) v += W[i] * dsh;
        out[(size_t)i * cnt + k] = v;
    }
}

template <class S>
__global__ __launch_bounds__(256) void scatter_kernel(S* __restrict__ f, const uint8_t* __restrict__ flag,
                                                      const double* __restrict__ in, int NX, int NY, int NZ, uint32_t n0,
                                                      uint32_t cnt, uint32_t SS, uint32_t PX, int shifted) {
    const uint32_t k = blockIdx.x * 256u + threadIdx.x;
    if (k >= cnt) return;
    const uint32_t n = n0 + k;
    const int x = (int)(n % (uint32_t)NX), y = (int)((n / (uint32_t)NX) % (uint32_t)NY),
              z = (int)(n / ((uint32_t)NX * (uint32_t)NY));
    // only FLUID nodes own populations: the slots of solid nodes may hold the
    // bounced populations of their fluid neighbours
    const bool fluid = (flag[n] & TYPE_MASK) == FLUID;
    if (!fluid) return;
    for (int i = 0; i < Q; ++i) {
        size_t slot = (size_t)i * SS + pop_index(x, y, z, PX, NY);
        if (fluid && i > 0) {
            const int xm = wrapi(x - CX[i], NX), ym = wrapi(y - CY[i], NY), zm = wrapi(z - CZ[i], NZ);
            const uint32_t m = (uint32_t)xm + (uint32_t)NX * ((uint32_t)ym + (uint32_t)NY * (uint32_t)zm);
            if (shifted || (flag[m] & TYPE_MASK) != FLUID) slot = (size_t)OPP[i] * SS + pop_index(xm, ym, zm, PX, NY);
        }
        f[slot] = from_real<S>(in[(size_t)i * cnt + k], i);
    }
}

// Excess mass sum_i f_i - 1 of every FLUID node from the logical populations,
// one partial per block (used when the last step emitted no mass partials).
template <class S>
__global__ __launch_bounds__(256) void mass_kernel(const S* __restrict__ f, const uint8_t* __restrict__ flag,
                                                   MassPart* __restrict__ part, int NX, int NY, int NZ, uint32_t N,
                                                   uint32_t SS, uint32_t PX, int shifted, const DevState* st, int shift_slot) {
    static_assert(BLOCK == 256, "one mass partial per block of 256 nodes");
    const uint32_t n = blockIdx.x * 256u + threadIdx.x;
    double ex[1] = {0.0};
    if (n < N && (flag[n] & TYPE_MASK) == FLUID) {
        const int x = (int)(n % (uint32_t)NX), y = (int)((n / (uint32_t)NX) % (uint32_t)NY),
                  z = (int)(n / ((uint32_t)NX * (uint32_t)NY));
        const double dsh = st->shift[shift_slot];
        double s = 0.0;
        for (int i = 0; i < Q; ++i) {
            size_t slot = (size_t)i * SS + pop_index(x, y, z, PX, NY);
            if (i > 0) {
                const int xm = wrapi(x - CX[i], NX), ym = wrapi(y - CY[i], NY), zm = wrapi(z - CZ[i], NZ);
                const uint32_t m = (uint32_t)xm + (uint32_t)NX * ((uint32_t)ym + (uint32_t)NY * (uint32_t)zm);
                if (shifted || (flag[m] & TYPE_MASK) != FLUID) slot = (size_t)OPP[i] * SS + pop_index(xm, ym, zm, PX, NY);
            }
            double v = to_real<S>(f[slot], i);
            if (dsh != 0.0) v += W[i] * dsh;
            s += v;
        }
        ex[0] = s - 1.0;
    }
    block_mass_sum(ex[0], &part[blockIdx.x]);
}

}  // namespace lbmhip
