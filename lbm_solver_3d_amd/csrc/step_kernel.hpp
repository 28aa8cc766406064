// step_kernel.hpp -- the fused D3Q19 TRT collide + stream kernel (gfx950).
//
// One thread per lattice node, one population array updated in place with
// the AA access pattern (P. Bailey et al., ICPP 2009):
//
//   even step: node n reads f_i from slot i at n, collides and writes f*_i
//              into slot opp(i) at n                  (all accesses local)
//   odd step:  node n reads f_i from slot opp(i) at n - c_i, collides and
//              writes f*_i into slot i at n + c_i     (all accesses remote)
//
// Each slot is read and written by one thread only, so the update is race
// free in place.  The post-collision values are the reference's
// (src/utils/lbm.cpp:128-224); only the places they are kept differ.
//
// Halfway bounce-back.  The odd step stays uniform: a link (n, i) whose
// neighbour n + c_i is solid still writes f*_i into slot i at n + c_i, i.e.
// into the solid node's own slot, which no fluid node uses otherwise.  The
// even step then reads a population whose upstream node is solid from that
// place (slot opp(j) at n - c_j) and, for a solid downstream node, stores the
// bounced value into slot i at n + c_i, where the next odd step's node n
// reads f_opp(i).  So the bounced value of link (n, i) always sits in slot i
// of the solid node n + c_i after either parity; the wall bounce-back is that
// plain copy (lbm.cpp:197-199 with u_b = 0), and the moving-boundary
// correction and the momentum exchange of PARTICLE links are applied to
// those slots afterwards by particle_links_kernel.  Only even-step lanes
// whose flag byte carries BOUNDARY touch neighbour flags, and they do so in
// addition to the uniform accesses, which every lane performs.
//
// Storage slot of the population f_j(n) that node n collides next
// ("loc(n, j)"): after an even step ("shifted" layout) slot opp(j) at
// n - c_j, whatever the flags; after an odd step ("natural" layout) slot j at
// n if n - c_j is FLUID, else slot opp(j) at n - c_j.
#pragma once
#include <type_traits>

#include "device_state.hpp"
#include "fixed_point.hpp"
#include "storage.hpp"

namespace lbmhip {

constexpr int STEP_MACRO = 1;
constexpr int STEP_WALLF = 2;
constexpr int STEP_MASS = 4;
// Kernel-template bit only (not a caller flag): the lattice has a body
// force, so the FP32 collision carries the forcing terms.  Chosen by the
// host from the force, so lattices without one run no force arithmetic and
// no branch on it.
constexpr int STEP_BODY_FORCE = 0x100;

constexpr int BLOCK = 256;

template <int B, int E, class Fn>
__device__ __forceinline__ void static_for(Fn&& fn) {
    if constexpr (B < E) {
        fn(std::integral_constant<int, B>{});
        static_for<B + 1, E>(fn);
    }
}

// ---------------------------------------------------------------------------
// FP64 collision in the reference's operation order (lbm.cpp:143-172).
// g holds f_i; on return g holds f*_i.
// ---------------------------------------------------------------------------
__device__ __forceinline__ void collide_exact(double (&g)[Q], double omP, double omM, double Fx, double Fy,
                                              double Fz, double& r, double& u, double& v, double& wz) {
    r = 0.0;
    double mx = 0.0, my = 0.0, mz = 0.0;
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        r += g[i];
        if constexpr (CX[i] != 0) mx += (double)CX[i] * g[i];
        if constexpr (CY[i] != 0) my += (double)CY[i] * g[i];
        if constexpr (CZ[i] != 0) mz += (double)CZ[i] * g[i];
    });
    u = (mx + 0.5 * Fx) / r;
    v = (my + 0.5 * Fy) / r;
    wz = (mz + 0.5 * Fz) / r;
    const double usq = u * u + v * v + wz * wz, uF = u * Fx + v * Fy + wz * Fz;
    const double gP = 1.0 - 0.5 * omP, gM = 1.0 - 0.5 * omM;

    {  // rest population: c_0 = 0
        const double cu = 0.0, cF = 0.0;
        const double feqP = W[0] * r * (1.0 + 4.5 * cu * cu - 1.5 * usq);
        const double feqM = W[0] * r * 3.0 * cu;
        const double fP = 0.5 * (g[0] + g[0]), fM = 0.5 * (g[0] - g[0]);
        const double SP = W[0] * (9.0 * cu * cF - 3.0 * uF), SM = W[0] * 3.0 * cF;
        g[0] = g[0] - omP * (fP - feqP) - omM * (fM - feqM) + gP * SP + gM * SM;
    }
    // Opposite pairs (i, i+1): every term of direction i+1 is the exact
    // negation or copy of the one of direction i, so the four products are
    // formed once and combined in the reference's order for both.
    static_for<0, 9>([&](auto P) {
        constexpr int i = 2 * decltype(P)::value + 1, io = i + 1;
        double cu, cF;
        {
            bool first = true;
            auto addc = [&](int c, double a, double b) {
                const double ta = c > 0 ? a : -a, tb = c > 0 ? b : -b;
                if (first) { cu = ta; cF = tb; first = false; }
                else { cu = cu + ta; cF = cF + tb; }
            };
            if (CX[i] != 0) addc(CX[i], u, Fx);
            if (CY[i] != 0) addc(CY[i], v, Fy);
            if (CZ[i] != 0) addc(CZ[i], wz, Fz);
        }
        const double feqP = W[i] * r * (1.0 + 4.5 * cu * cu - 1.5 * usq);
        const double feqM = W[i] * r * 3.0 * cu;
        const double fP = 0.5 * (g[i] + g[io]), fM = 0.5 * (g[i] - g[io]);
        const double SP = W[i] * (9.0 * cu * cF - 3.0 * uF), SM = W[i] * 3.0 * cF;
        const double A = omP * (fP - feqP), B = omM * (fM - feqM), C = gP * SP, D = gM * SM;
        const double gi = g[i], gio = g[io];
        g[i] = gi - A - B + C + D;
        g[io] = gio - A + B + C - D;
    });
}

// ---------------------------------------------------------------------------
// FP32 collision on shifted, scaled populations gs_i = s (f_i - w_i).
// Same TRT + Guo update; constants fold the scale.  On return gs holds
// s (f*_i - w_i); rho is the density and (jx, jy, jz) = s * rho * u.
// ---------------------------------------------------------------------------
struct FastCoef {
    float inv_s, s, a, b, one_m_omp;  // a = omP/2, b = omM/2
    float kp_axis, kp_diag, kp_rest;  // omP * w_i
    float km_axis, km_diag;           // 3 * omM * w_i
    float hFx, hFy, hFz;              // s * F / 2
    float gPw9[3], gPw3[3];           // 9 gP w (axis/diag/rest), 3 gP w
    float sm_axis, sm_diag;           // 3 s gM w
    float Fx, Fy, Fz;
    float one;                        // 1.0f (see Store<_Float16>::load_add)
};

template <bool FORCE>
__device__ __forceinline__ void collide_fast(float (&g)[Q], const FastCoef& K, float& rho, float& jx, float& jy,
                                             float& jz, float& drho_s) {
    // Every multiply-add is an explicit fmaf and the unit is compiled
    // without contraction, so each kernel that inlines this (bulk, boundary,
    // ensemble) rounds identically.
    // pair sums and differences feed both the moments and the collision
    float sp[9], df[9];
    static_for<0, 9>([&](auto P) {
        constexpr int p = decltype(P)::value, i = 2 * p + 1;
        sp[p] = g[i] + g[i + 1];
        df[p] = g[i] - g[i + 1];
    });
    // directions: 1:+x 3:+y 5:+z 7:(+x+y) 9:(+x-y) 11:(+x+z) 13:(+x-z) 15:(+y+z) 17:(+y-z)
    float d = g[0];
    static_for<0, 9>([&](auto P) { d += sp[decltype(P)::value]; });
    drho_s = d;
    jx = ((df[0] + df[3]) + (df[4] + df[5])) + df[6];
    jy = ((df[1] + df[3]) - (df[4] - df[7])) + df[8];
    jz = ((df[2] + df[5]) - (df[6] - df[7])) - df[8];
    if constexpr (FORCE) { jx += K.hFx; jy += K.hFy; jz += K.hFz; }
    rho = fmaf(d, K.inv_s, 1.0f);
    const float isr = K.inv_s / rho;  // 1 / (s rho)
    const float q = 4.5f * isr;
    const float jj = fmaf(jz, jz, fmaf(jy, jy, jx * jx));
    const float t0 = fmaf(-1.5f * isr, jj, d);
    float uF = 0.f, irho = 0.f;
    if constexpr (FORCE) {
        irho = 1.0f / rho;
        uF = fmaf(jz, K.Fz, fmaf(jy, K.Fy, jx * K.Fx));  // = s rho (u . F)
    }
    // rest population
    {
        float r0 = fmaf(K.one_m_omp, g[0], K.kp_rest * t0);
        if constexpr (FORCE) r0 = fmaf(-K.gPw3[2] * uF, irho, r0);
        g[0] = r0;
    }
    static_for<0, 9>([&](auto P) {
        constexpr int p = decltype(P)::value, i = 2 * p + 1, io = i + 1;
        constexpr bool diag = p >= 3;
        float pcu;  // s rho (c_i . u)
        if constexpr (p == 0) pcu = jx;
        else if constexpr (p == 1) pcu = jy;
        else if constexpr (p == 2) pcu = jz;
        else if constexpr (p == 3) pcu = jx + jy;
        else if constexpr (p == 4) pcu = jx - jy;
        else if constexpr (p == 5) pcu = jx + jz;
        else if constexpr (p == 6) pcu = jx - jz;
        else if constexpr (p == 7) pcu = jy + jz;
        else pcu = jy - jz;
        const float kp = diag ? K.kp_diag : K.kp_axis;
        const float km = diag ? K.km_diag : K.km_axis;
        const float e = fmaf(q * pcu, pcu, t0);
        float Pp = fmaf(kp, e, -K.a * sp[p]);
        float Mm = fmaf(km, pcu, -K.b * df[p]);
        if constexpr (FORCE) {
            // c_i . F for this direction
            float cF;
            if constexpr (p == 0) cF = K.Fx;
            else if constexpr (p == 1) cF = K.Fy;
            else if constexpr (p == 2) cF = K.Fz;
            else if constexpr (p == 3) cF = K.Fx + K.Fy;
            else if constexpr (p == 4) cF = K.Fx - K.Fy;
            else if constexpr (p == 5) cF = K.Fx + K.Fz;
            else if constexpr (p == 6) cF = K.Fx - K.Fz;
            else if constexpr (p == 7) cF = K.Fy + K.Fz;
            else cF = K.Fy - K.Fz;
            const int k = diag ? 1 : 0;
            Pp = fmaf(fmaf(K.gPw9[k] * cF, pcu, -K.gPw3[k] * uF), irho, Pp);
            Mm = fmaf(diag ? K.sm_diag : K.sm_axis, cF, Mm);
        }
        const float gi = g[i], gio = g[io];
        g[i] = gi + Pp + Mm;
        g[io] = gio + Pp - Mm;
    });
}

// ---------------------------------------------------------------------------
// Squirmer surface velocity at a link midpoint (lbm.cpp:100-104 minimum
// image, squirmer.cpp:35-46 rigid-body motion + two-mode slip), FP64.
// ---------------------------------------------------------------------------
// r - N round(r / N) (lbm.cpp:100-104) without the division where it cannot
// matter: for |r| < N / 4 the quotient rounds to zero and r comes back as is.
__device__ __forceinline__ double min_image(double r, int N) {
    return fabs(r) < 0.25 * N ? r : r - N * round(r / N);
}
__device__ __forceinline__ void link_velocity(const DevSquirmer& sq, int NX, int NY, int NZ, int wallsY, int wallsZ,
                                              double px, double py, double pz, double rb[3], double ub[3]) {
    rb[0] = min_image(px - sq.X[0], NX);
    rb[1] = py - sq.X[1]; if (!wallsY) rb[1] = min_image(rb[1], NY);
    rb[2] = pz - sq.X[2]; if (!wallsZ) rb[2] = min_image(rb[2], NZ);
    ub[0] = sq.U[0] + (sq.W[1] * rb[2] - sq.W[2] * rb[1]);
    ub[1] = sq.U[1] + (sq.W[2] * rb[0] - sq.W[0] * rb[2]);
    ub[2] = sq.U[2] + (sq.W[0] * rb[1] - sq.W[1] * rb[0]);
    const double rn = sqrt(rb[0] * rb[0] + rb[1] * rb[1] + rb[2] * rb[2]);
    if (rn < 1e-12) return;
    const double rh[3] = {rb[0] / rn, rb[1] / rn, rb[2] / rn};
    const double ct = sq.e[0] * rh[0] + sq.e[1] * rh[1] + sq.e[2] * rh[2];
    const double amp = sq.B1 * (1.0 + sq.beta * ct);
    ub[0] += amp * (ct * rh[0] - sq.e[0]);
    ub[1] += amp * (ct * rh[1] - sq.e[1]);
    ub[2] += amp * (ct * rh[2] - sq.e[2]);
}

// Rows of the squirmer's link box (nodes within R + 1 of X) and the range of
// 64-node x chunks that cover it.
struct LinkBoxDims {
    int y0, z0, wy, wz, c0, nck;
};
__device__ inline int wrap_coord(int a, int n) {
    a %= n;
    return a < 0 ? a + n : a;
}
__device__ inline LinkBoxDims link_box_dims(const double X[3], double R, int NX, int NY, int NZ) {
    LinkBoxDims b;
    const int x0 = (int)floor(X[0] - R) - 1, x1 = (int)ceil(X[0] + R) + 1;
    const int y0 = (int)floor(X[1] - R) - 1, y1 = (int)ceil(X[1] + R) + 1;
    const int z0 = (int)floor(X[2] - R) - 1, z1 = (int)ceil(X[2] + R) + 1;
    b.y0 = y0; b.z0 = z0;
    b.wy = (y1 - y0 + 1) < NY ? (y1 - y0 + 1) : NY;
    b.wz = (z1 - z0 + 1) < NZ ? (z1 - z0 + 1) : NZ;
    const int nchx = (NX + 63) / 64;
    const int wx = (x1 - x0 + 1) < NX ? (x1 - x0 + 1) : NX;
    const int xs = wrap_coord(x0, NX);
    b.c0 = xs / 64;
    if (xs + wx > NX)  // wraps around x: chunks c0 .. last, then 0 .. the one holding the tail
        b.nck = (nchx - b.c0) + (xs + wx - 1 - NX) / 64 + 1;
    else
        b.nck = (xs + wx - 1) / 64 - b.c0 + 1;
    if (b.nck > nchx) b.nck = nchx;
    return b;
}
__device__ inline LinkBoxDims link_box_dims(const DevSquirmer& sq, int NX, int NY, int NZ) {
    return link_box_dims(sq.X, sq.R, NX, NY, NZ);
}
// A sphere (centre c, radius r) covering the squirmer at its centre X and
// at the centre xf its flags were classified at: the two are the same after
// every remap, and a host update moves X by less than a node.  The link and
// boundary boxes are built around it, so they hold every link of the flags.
__device__ inline double covering_sphere(const DevState* st, int NX, int NY, int NZ, double c[3]) {
    const DevSquirmer& sq = st->sq;
    const int Nd[3] = {NX, NY, NZ};
    const double r0 = sq.R > st->rf ? sq.R : st->rf;
    double r = r0;
    for (int d = 0; d < 3; ++d) {
        double dx = st->xf[d] - sq.X[d];
        dx -= Nd[d] * round(dx / Nd[d]);  // nearest image (a walled axis never wraps this far)
        c[d] = sq.X[d] + 0.5 * dx;
        r = fmax(r, r0 + 0.5 * fabs(dx));
    }
    return r;
}
// Rows of the even-step boundary chunks near the squirmer.
__device__ inline LinkBoxDims boundary_box_dims(const DevState* st, int NX, int NY, int NZ) {
    double c[3];
    const double r = covering_sphere(st, NX, NY, NZ, c);
    return link_box_dims(c, r, NX, NY, NZ);
}
// Whether chunk c of row (y, z) lies in box b.
__device__ inline bool in_link_box(const LinkBoxDims& b, int c, int y, int z, int nchx, int NY, int NZ) {
    int dy = y - wrap_coord(b.y0, NY); dy += dy < 0 ? NY : 0;
    int dz = z - wrap_coord(b.z0, NZ); dz += dz < 0 ? NZ : 0;
    int dc = c - b.c0; dc += dc < 0 ? nchx : 0;
    return dy < b.wy && dz < b.wz && dc < b.nck;
}

// Block sum (blocks of `waves` wavefronts, at most 16) of the nodes' excess
// mass in 128-bit fixed point (fixed_point.hpp); thread 0 stores it.  `lds`
// holds 2 * waves words.
__device__ __forceinline__ void block_mass_sum(double excess, MassPart* out, int waves, unsigned long long* lds) {
    u128 v = fixed_from_double(excess);
    for (int off = 32; off > 0; off >>= 1) {
        const unsigned long long h = __shfl_xor((unsigned long long)fixed_hi(v), off, 64);
        const unsigned long long l = __shfl_xor((unsigned long long)fixed_lo(v), off, 64);
        v += fixed_join(h, l);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        lds[2 * w] = (unsigned long long)fixed_hi(v);
        lds[2 * w + 1] = fixed_lo(v);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        u128 s = 0;
        for (int k = 0; k < waves; ++k) s += fixed_join(lds[2 * k], lds[2 * k + 1]);
        out->hi = fixed_hi(s);
        out->lo = fixed_lo(s);
    }
}
__device__ __forceinline__ void block_mass_sum(double excess, MassPart* out) {
    __shared__ unsigned long long lds[2 * (BLOCK / 64)];
    block_mass_sum(excess, out, BLOCK / 64, lds);
}

// Wavefront + block reduction of K doubles over a block of NT threads; thread
// 0 of the block gets the sums.  `lds` holds (NT / 64) * K doubles.
template <int K, int NT = BLOCK>
__device__ __forceinline__ void block_reduce(double (&v)[K], double* lds) {
    static_for<0, K>([&](auto J) {
        constexpr int j = decltype(J)::value;
        double x = v[j];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        v[j] = x;
    });
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0)
        static_for<0, K>([&](auto J) { lds[wave * K + decltype(J)::value] = v[decltype(J)::value]; });
    __syncthreads();
    if (threadIdx.x == 0) {
        static_for<0, K>([&](auto J) {
            constexpr int j = decltype(J)::value;
            double s = lds[j];
            for (int w = 1; w < NT / 64; ++w) s += lds[w * K + j];
            v[j] = s;
        });
    }
    __syncthreads();
}

// ---------------------------------------------------------------------------
// The kernels.
//
// Addressing: a wavefront is 64 consecutive x of one (y, z) row (chunks of
// 64 aligned to multiples of 64), so the row offsets of the 9 (dy, dz)
// neighbour rows are wavefront-uniform and live in scalar registers together
// with the 19 slot bases; a lane only holds its three x byte offsets (x-1,
// x, x+1, periodic).  Every uniform load and store is one global access with
// a scalar base and a 32-bit lane offset.
//
// Even steps run in two kernels: aa_step_kernel skips every 64-node chunk
// that contains a BOUNDARY node, and aa_boundary_kernel updates exactly those
// chunks (same test), so the bulk kernel carries no bounce-back code at all.
// Odd steps need none (see the layout notes above) and run in aa_step_kernel
// alone.
// ---------------------------------------------------------------------------
#ifndef LBM_STEP_WAVES_PER_EU
#define LBM_STEP_WAVES_PER_EU 4
#endif
#ifndef LBM_ODD_NT
#define LBM_ODD_NT 1
#endif
constexpr bool ODD_NT = LBM_ODD_NT != 0;

// Neighbour offsets of node (x, y, z): node-array rows (yo, zo, pitch NX)
// and population rows (pyo, pzo, pitch PX); x offsets in nodes and in bytes.
struct RowGeom {
    uint32_t xo[3], xb[3], yo[3], zo[3], pyo[3], pzo[3], n;
};

template <class S>
__device__ __forceinline__ RowGeom row_geom(uint32_t x, uint32_t y, uint32_t z, uint32_t NX, uint32_t NY,
                                            uint32_t NZ, uint32_t PX) {
    RowGeom g;
    const uint32_t ym = y == 0 ? NY - 1 : y - 1, yp = y + 1 == NY ? 0u : y + 1;
    const uint32_t zm = z == 0 ? NZ - 1 : z - 1, zp = z + 1 == NZ ? 0u : z + 1;
    const uint32_t NXY = NX * NY, PXY = PX * NY;
    g.yo[0] = ym * NX; g.yo[1] = y * NX; g.yo[2] = yp * NX;
    g.zo[0] = zm * NXY; g.zo[1] = z * NXY; g.zo[2] = zp * NXY;
    g.pyo[0] = ym * PX; g.pyo[1] = y * PX; g.pyo[2] = yp * PX;
    g.pzo[0] = zm * PXY; g.pzo[1] = z * PXY; g.pzo[2] = zp * PXY;
    g.xo[0] = x == 0 ? NX - 1 : x - 1; g.xo[1] = x; g.xo[2] = x + 1 == NX ? 0u : x + 1;
    for (int k = 0; k < 3; ++k) g.xb[k] = g.xo[k] * (uint32_t)sizeof(S);
    g.n = x + g.yo[1] + g.zo[1];
    return g;
}

// Population index of node (x, y, z) within a slot.
__device__ __forceinline__ uint32_t pop_index(int x, int y, int z, uint32_t PX, int NY) {
    return (uint32_t)x + PX * ((uint32_t)y + (uint32_t)NY * (uint32_t)z);
}

// Loads and stores that bypass the caches' retention.  The even step's bulk
// reads each slot once and writes each slot once, with no reuse by other
// wavefronts, where these hints let the in-place traffic stream faster; the
// odd step's shifted reads share lines between neighbouring wavefronts and
// keep the default policy.
template <class S>
__device__ __forceinline__ S load_stream(const S* p) {
    if constexpr (std::is_same<S, bf16_t>::value) {
        bf16_t r;
        r.bits = __builtin_nontemporal_load(reinterpret_cast<const uint16_t*>(p));
        return r;
    } else {
        return __builtin_nontemporal_load(p);
    }
}
template <class S>
__device__ __forceinline__ void store_stream(S* p, S v) {
    if constexpr (std::is_same<S, bf16_t>::value) __builtin_nontemporal_store(v.bits, reinterpret_cast<uint16_t*>(p));
    else __builtin_nontemporal_store(v, p);
}

// Moments of a collided node (real units) and its excess mass sum_i f*_i - 1.
template <class T>
struct NodeOut {
    T rho, ux, uy, uz;
    double excess;
};

// Collision of one node whose 19 populations (mass shift included) are in g
// (all update paths share it); the excess mass only when MASS, the velocity
// only when MACRO.  The FP32 collision carries the body-force terms only
// when MODE has STEP_BODY_FORCE (the FP64 one always does).
template <class S, int MODE>
__device__ __forceinline__ NodeOut<typename Store<S>::C> collide_node(const StepArgs<S>& A, const FastCoef& K,
                                                                      typename Store<S>::C (&g)[Q]) {
    using T = typename Store<S>::C;
    constexpr bool EXACT = std::is_same<S, double>::value;
    constexpr bool MACRO = (MODE & STEP_MACRO) != 0;
    constexpr bool MASS = (MODE & STEP_MASS) != 0;
    constexpr bool FORCE = (MODE & STEP_BODY_FORCE) != 0;

    // ---- collide ----
    double excess = 0.0;
    T rho_t, m0, m1, m2;  // macroscopic outputs (real units)
    if constexpr (EXACT) {
        double r, u, v, wz;
        collide_exact(g, A.omPlus, A.omMinus, A.Fx, A.Fy, A.Fz, r, u, v, wz);
        rho_t = r; m0 = u; m1 = v; m2 = wz;
        if constexpr (MASS) {
            double s = 0.0;
            static_for<0, Q>([&](auto I) { s += g[decltype(I)::value]; });
            excess = s - 1.0;
        }
    } else {
        float rho, jx, jy, jz, drs;
        collide_fast<FORCE>(g, K, rho, jx, jy, jz, drs);
        rho_t = rho;
        if constexpr (MACRO) {
            const float isr = K.inv_s / rho;
            m0 = jx * isr; m1 = jy * isr; m2 = jz * isr;
        } else {
            m0 = m1 = m2 = 0.f;
        }
        if constexpr (MASS) excess = (double)drs * (double)K.inv_s;
    }
    return NodeOut<T>{rho_t, m0, m1, m2, excess};
}

// Macroscopic fields of a collided node: all fluid nodes on request, else
// the band around the squirmer that the remap and the link kernel read
// (lbm.cpp:152, 198, 307).  Written after the node's populations, so the
// stores of the populations sit in the block of the collision (scalar row
// bases with 32-bit lane offsets).
template <class S, int MODE>
__device__ __forceinline__ void write_moments(const StepArgs<S>& A, const NodeOut<typename Store<S>::C>& o, uint32_t n,
                                              int x, int y, int z) {
    using T = typename Store<S>::C;
    constexpr bool MACRO = (MODE & STEP_MACRO) != 0;
    const T rho_t = o.rho;
    if constexpr (MACRO) {
        static_cast<T*>(A.rho)[n] = rho_t;
        static_cast<T*>(A.ux)[n] = o.ux;
        static_cast<T*>(A.uy)[n] = o.uy;
        static_cast<T*>(A.uz)[n] = o.uz;
    } else {
        if (A.has_squirmer) {
            const int* b = A.st->band;
            int dx = x - b[0]; dx += dx < 0 ? A.NX : 0;
            int dy = y - b[1]; dy += dy < 0 ? A.NY : 0;
            int dz = z - b[2]; dz += dz < 0 ? A.NZ : 0;
            if (dx < b[3] && dy < b[4] && dz < b[5]) static_cast<T*>(A.rho)[n] = rho_t;
        }
    }
}

// Update one fluid node; returns its excess mass sum_i f*_i - 1 when MASS.
// BPATH: even-step node of a boundary chunk (bounce-back redirections).
// `row_solid` (BPATH): the solid directions, the same for every node of the
// wavefront's row chunk (a row next to a wall, away from the squirmer), or
// -1 to read them from the neighbours' flags.  `dsh`: the pending
// mass-correction shift, read by the caller at the start of the kernel (its
// scalar load then overlaps the index arithmetic instead of delaying the
// node's loads).
template <class S, bool ODD, int MODE, bool BPATH>
__device__ __forceinline__ double node_update(const StepArgs<S>& A, const FastCoef& K, const RowGeom& rg, int x,
                                              int y, int z, uint8_t fl, double dsh, int row_solid = -1) {
    using St = Store<S>;
    using T = typename St::C;
    const uint32_t SS = A.SS, n = rg.n;
    S* __restrict__ f = A.f;
    auto base = [&](int s, uint32_t row) -> char* { return (char*)(f + ((size_t)s * SS + row)); };
    auto at = [](char* b, uint32_t off) -> S* { return (S*)(b + off); };
    const uint32_t* xb = rg.xb;
    const uint32_t* yo = rg.pyo;  // population rows
    const uint32_t* zo = rg.pzo;

    // Pending mass-correction shift (lbm.cpp:337): W_i dsh is added to every
    // population as it is loaded (zero on all steps but the one after a mass
    // correction).  Adding it unconditionally keeps a node's loads and
    // stores in one basic block, where the compiler addresses both with the
    // scalar slot-row bases and 32-bit lane offsets; a branch between them
    // costs 64-bit vector addresses for the stores and a third more vector
    // registers.  FP16 converts and adds in one fused instruction.
    T sh[Q];
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        if constexpr (St::shifted) sh[i] = (T)(W[i] * dsh * St::scale);
        else sh[i] = W[i] * dsh;
    });

    // ---- gather the 19 populations node n collides (stored form) ----
    S gs[Q];
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        if constexpr (!ODD && !BPATH) gs[i] = load_stream(at(base(i, yo[1] + zo[1]), xb[1]));
        else if constexpr (!ODD) gs[i] = *at(base(i, yo[1] + zo[1]), xb[1]);
        else if constexpr (CX[i] == 0 && ODD_NT) gs[i] = load_stream(at(base(OPP[i], yo[1 - CY[i]] + zo[1 - CZ[i]]), xb[1]));
        else gs[i] = *at(base(OPP[i], yo[1 - CY[i]] + zo[1 - CZ[i]]), xb[1 - CX[i]]);
    });
    // even step, boundary node: populations whose upstream node is solid
    // come back from that node's slot opp(j) (bounce-back)
    uint32_t solid = 0;
    if constexpr (BPATH) {
        if (row_solid >= 0) {  // wavefront-uniform: no flag loads, uniform branches
            solid = (uint32_t)row_solid;
            static_for<1, Q>([&](auto I) {
                constexpr int j = decltype(I)::value;
                if ((solid >> OPP[j]) & 1u) gs[j] = *at(base(OPP[j], yo[1 - CY[j]] + zo[1 - CZ[j]]), xb[1 - CX[j]]);
            });
        } else if (fl & BOUNDARY) {
            static_for<1, Q>([&](auto I) {
                constexpr int i = decltype(I)::value;
                const uint32_t m = rg.xo[1 + CX[i]] + rg.yo[1 + CY[i]] + rg.zo[1 + CZ[i]];
                solid |= (uint32_t)((A.flag[m] & TYPE_MASK) != FLUID) << i;
            });
            static_for<1, Q>([&](auto I) {
                constexpr int j = decltype(I)::value;
                if ((solid >> OPP[j]) & 1u) gs[j] = *at(base(OPP[j], yo[1 - CY[j]] + zo[1 - CZ[j]]), xb[1 - CX[j]]);
            });
        }
    }
    // to the collision's form, after the (possibly divergent) redirections,
    // so no redirected load waits inside its branch
    T g[Q];
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        g[i] = St::load_add(gs[i], sh[i], K.one);
    });

    const NodeOut<T> out = collide_node<S, MODE>(A, K, g);

    // ---- stream ----
    static_for<0, Q>([&](auto I) {
        constexpr int i = decltype(I)::value;
        const S val = St::store(g[i]);
        if constexpr (!ODD && !BPATH) store_stream(at(base(OPP[i], yo[1] + zo[1]), xb[1]), val);
        else if constexpr (!ODD) *at(base(OPP[i], yo[1] + zo[1]), xb[1]) = val;
        else if constexpr (CX[i] == 0 && ODD_NT) store_stream(at(base(i, yo[1 + CY[i]] + zo[1 + CZ[i]]), xb[1]), val);
        else *at(base(i, yo[1 + CY[i]] + zo[1 + CZ[i]]), xb[1 + CX[i]]) = val;
    });
    // even step, boundary node: bounce-back into the solid neighbour's slot i
    if constexpr (BPATH) {
        if (solid) {
            static_for<1, Q>([&](auto I) {
                constexpr int i = decltype(I)::value;
                if ((solid >> i) & 1u) *at(base(i, yo[1 + CY[i]] + zo[1 + CZ[i]]), xb[1 + CX[i]]) = St::store(g[i]);
            });
        }
    }
    write_moments<S, MODE>(A, out, n, x, y, z);
    return out.excess;
}

// The hardware deals blocks to the 8 XCDs round robin; this gives each XCD
// a contiguous range of the G tiles, so neighbouring rows share an L2.
__device__ __forceinline__ int xcd_tile(int lin, int G) {
    const int q8 = G >> 3, r8 = G & 7, xcd = lin & 7;
    return xcd * q8 + (xcd < r8 ? xcd : r8) + (lin >> 3);
}

// The tiles of one tile row (gx tiles across x; `rows` tile rows) on one
// XCD, consecutive tile rows dealt to the XCDs round robin: the lines a
// row's tiles share at their x edges (the odd step's shifted accesses) stay
// in one L2.  The rows past the last multiple of 8 keep the identity order.
__device__ __forceinline__ int xcd_tile_rows(int lin, int rows, int gx, const FastDiv& div_gx) {
    const int rows8 = rows & ~7;
    if (lin >= rows8 * gx) return lin;
    const uint32_t xcd = (uint32_t)lin & 7u, k = (uint32_t)lin >> 3;
    const uint32_t kr = quot(k, div_gx);
    return (int)(((kr * 8u + xcd) * (uint32_t)gx) + (k - kr * (uint32_t)gx));
}

// One tile (one block of BLOCK threads) of the collide-stream update: either
// 2^tx_shift x ty nodes of a z plane, or (tx_shift == 0, a "row run")
// BLOCK / 64 consecutive 64-node chunks of the rows in row-major order,
// which may end one row and start the next (no idle lanes on rows narrower
// than a tile).  Either way a wavefront is one 64-node chunk of one row.
template <class S, bool ODD, int MODE>
__device__ __forceinline__ void step_tile(const StepArgs<S>& A, const FastCoef& K, int tile) {
    constexpr bool MASS = (MODE & STEP_MASS) != 0;
    const double dsh = A.st->shift[A.t & 1];  // pending mass-correction shift (lbm.cpp:337)
    int x, y, z;
    if (A.tx_shift == 0) {
        const uint32_t chunk = __builtin_amdgcn_readfirstlane(tile * (BLOCK / 64) + (int)(threadIdx.x >> 6));
        const uint32_t row = quot(chunk, A.div_nchx);
        x = (int)(chunk - row * (uint32_t)A.nchx) * 64 + (int)(threadIdx.x & 63);
        z = (int)quot(row, A.div_ny);
        y = (int)row - z * A.NY;
        if (z >= A.NZ) { y = -1; z = 0; }  // past the last row of the lattice
    } else {
        const uint32_t trest = quot((uint32_t)tile, A.div_gx);
        const uint32_t tgz_u = quot(trest, A.div_gy);
        const int tgx = tile - (int)trest * A.gx, tgy = (int)(trest - tgz_u * (uint32_t)A.gy), tgz = (int)tgz_u;
        const int lx = threadIdx.x & ((1 << A.tx_shift) - 1);
        const int ly = __builtin_amdgcn_readfirstlane(threadIdx.x >> A.tx_shift);  // one row per wavefront
        x = (tgx << A.tx_shift) + lx; y = tgy * A.ty + ly; z = tgz;
    }

    const uint32_t NX = (uint32_t)A.NX, NY = (uint32_t)A.NY, NZ = (uint32_t)A.NZ;
    const bool row_ok = (uint32_t)y < NY;  // wavefront-uniform
    const bool inb = row_ok && (uint32_t)x < NX;
    const RowGeom rg = row_geom<S>(inb ? (uint32_t)x : 0u, row_ok ? (uint32_t)y : 0u, (uint32_t)z, NX, NY, NZ, A.PX);
    const uint8_t fl = inb ? A.flag[rg.n] : (uint8_t)WALL;
    bool active = (fl & TYPE_MASK) == FLUID;
    if constexpr (!ODD) {
        // chunks with bounce-back links belong to aa_boundary_kernel
        if (__builtin_amdgcn_ballot_w64(active && (fl & BOUNDARY)) != 0) active = false;
    }

    // clear the slot step t+1 reads its mass shift from
    if (tile == 0 && threadIdx.x == 0) A.st->shift[(A.t + 1) & 1] = 0.0;

    double excess = 0.0;
    if (active) excess = node_update<S, ODD, MODE, false>(A, K, rg, x, y, z, fl, dsh);

    if constexpr (MASS) block_mass_sum(excess, &A.mass_part[tile]);
}

template <class S, bool ODD, int MODE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(LBM_STEP_WAVES_PER_EU)))
void aa_step_kernel(const StepArgs<S> A, const FastCoef K) {
    const int G = A.gx * A.gy * A.gz;
    const int lin = (int)blockIdx.x;
    step_tile<S, ODD, MODE>(A, K,
                            A.xcd_remap == 1 ? xcd_tile(lin, G)
                                             : (A.xcd_remap == 2 ? xcd_tile_rows(lin, A.gy * A.gz, A.gx, A.div_gx) : lin));
}

// ---------------------------------------------------------------------------
// Odd step in row blocks (stored rows of whole wavefronts: the row pitch PX a
// multiple of 64, at most 1024).
//
// In the odd step node n stores f*_i into slot i at n + c_i.  For the ten
// populations with c_x != 0 a wavefront's 64 stores then straddle two lines
// of the slot, each line shared with a neighbouring wavefront.  Here a block
// updates whole rows (PX / 64 wavefronts each) and those ten stores are made
// line-aligned: each value moves one lane up (c_x = +1) or down
// (c_x = -1) by a wavefront shift (DPP), the value that crosses into the
// neighbouring wavefront's line comes from that wavefront through LDS (the
// row is periodic in x, so the last wavefront's neighbour is the first), and
// lane l stores at its own x what node x - c_x computed, if that node is
// fluid.  Every slot address and value is the one the per-node form stores;
// only the lane that issues the store changes.  The loads keep the per-node
// form (straddling reads cost nothing measurable once the stores are whole
// lines).  One block barrier separates the collisions from those stores.
constexpr int ROW_MAX_WAVES = 16;
constexpr int NXS = 5;                             // populations per sign of c_x
// The row kernel's static LDS, the same for every storage and mode (the
// occupancy cap's LDS reservation counts on it): edge values, fluid masks,
// mass sums.
constexpr int ROW_LDS_WORDS = 256;
static_assert(2 * ROW_MAX_WAVES * NXS + ROW_MAX_WAVES + 2 * ROW_MAX_WAVES <= ROW_LDS_WORDS, "row kernel LDS");  // edges, fluid masks, mass
constexpr int XPLUS[NXS] = {1, 7, 9, 11, 13};      // c_x = +1
constexpr int XMINUS[NXS] = {2, 8, 10, 12, 14};    // c_x = -1

// Wavefront shift of a stored value by one lane: UP gives lane l the value
// of lane l - 1 and lane 0 `edge`; otherwise lane l gets lane l + 1 and lane
// 63 `edge` (DPP wave_shr:1 / wave_shl:1, all lanes enabled).
template <bool UP, class S>
__device__ __forceinline__ S lane_shift(S v, S edge) {
    constexpr int CTRL = UP ? 0x138 : 0x130;
    if constexpr (sizeof(S) == 8) {
        const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
        const unsigned long long e = __builtin_bit_cast(unsigned long long, edge);
        const int lo = __builtin_amdgcn_update_dpp((int)(uint32_t)e, (int)(uint32_t)u, CTRL, 0xF, 0xF, false);
        const int hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(e >> 32), (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF,
                                                   false);
        return __builtin_bit_cast(S, ((unsigned long long)(uint32_t)hi << 32) | (uint32_t)lo);
    } else if constexpr (sizeof(S) == 4) {
        return __builtin_bit_cast(S, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, edge),
                                                                 __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false));
    } else {
        const int r = __builtin_amdgcn_update_dpp((int)__builtin_bit_cast(uint16_t, edge),
                                                  (int)__builtin_bit_cast(uint16_t, v), CTRL, 0xF, 0xF, false);
        return __builtin_bit_cast(S, (uint16_t)r);
    }
}

// One block of the odd-step row kernel: `rpb` consecutive rows from
// `first_row`, each of A.row_waves wavefronts (PX / 64; waves past
// rpb * row_waves, and rows past the last, idle but take part in the
// barrier).  Rows may be padded (NX < PX): the row's last node is then lane
// (NX - 1) % 64 of its last wavefront, and the lanes past it neither load
// nor store.  `mass_out`: the block's excess-mass partial (MASS).  GENERAL
// false: one unpadded row per block (NX = PX, rpb = 1), where the bounds and
// padding logic compile away (measured: with it the FP16 odd step takes
// 220 instead of 208 us on 512 x 256 x 128).
template <class S, int MODE, bool GENERAL>
__device__ __forceinline__ void odd_rows_block(const StepArgs<S>& A, const FastCoef& K, int first_row, int rpb,
                                               MassPart* mass_out) {
    using St = Store<S>;
    using T = typename St::C;
    constexpr bool MASS = (MODE & STEP_MASS) != 0;
    __shared__ unsigned long long row_lds[ROW_LDS_WORDS];
    S* edge_up = reinterpret_cast<S*>(row_lds);                        // [wave][k]: c_x = +1 values of the wave's last node
    S* edge_dn = reinterpret_cast<S*>(row_lds + ROW_MAX_WAVES * NXS);  // [wave][k]: c_x = -1 values of lane 0
    unsigned long long* fluid = row_lds + 2 * ROW_MAX_WAVES * NXS;     // [wave]: its fluid lanes
    unsigned long long* mass_lds = fluid + ROW_MAX_WAVES;
    const double dsh = A.st->shift[A.t & 1];  // pending mass-correction shift (lbm.cpp:337)
    const uint32_t NX = (uint32_t)A.NX, NY = (uint32_t)A.NY, NZ = (uint32_t)A.NZ;
    const int nw = A.row_waves;
    const int wb = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));  // wave in the block
    const int lane = threadIdx.x & 63;
    const int rib = GENERAL ? wb / nw : 0, w = wb - rib * nw;  // row in the block, wave in the row
    const int row = first_row + rib;
    const bool row_ok = !GENERAL || (rib < rpb && row < A.NY * A.NZ);  // wavefront-uniform
    const int rb0 = rib * nw;                                           // the row's first wave in the block
    const uint32_t z = row_ok ? quot((uint32_t)row, A.div_ny) : 0u, y = row_ok ? (uint32_t)row - z * NY : 0u;
    const int x = w * 64 + lane;
    const bool valid = !GENERAL || (row_ok && (uint32_t)x < NX);
    const RowGeom rg = row_geom<S>(valid ? (uint32_t)x : 0u, y, z, NX, NY, NZ, A.PX);
    const uint8_t fl = valid ? A.flag[rg.n] : (uint8_t)WALL;
    const bool active = (fl & TYPE_MASK) == FLUID;

    // clear the slot step t+1 reads its mass shift from
    if (first_row == 0 && threadIdx.x == 0) A.st->shift[(A.t + 1) & 1] = 0.0;

    S* __restrict__ f = A.f;
    const uint32_t SS = A.SS;
    auto base = [&](int s, uint32_t r) -> char* { return (char*)(f + ((size_t)s * SS + r)); };
    auto at = [](char* b, uint32_t off) -> S* { return (S*)(b + off); };
    const uint32_t* xb = rg.xb;
    const uint32_t* yo = rg.pyo;
    const uint32_t* zo = rg.pzo;

    S up[NXS], dn[NXS];
    const S zero = St::store((T)0);
    static_for<0, NXS>([&](auto K_) { up[decltype(K_)::value] = zero; dn[decltype(K_)::value] = zero; });
    double excess = 0.0;
    if (active) {
        T sh[Q];
        static_for<0, Q>([&](auto I) {
            constexpr int i = decltype(I)::value;
            if constexpr (St::shifted) sh[i] = (T)(W[i] * dsh * St::scale);
            else sh[i] = W[i] * dsh;
        });
        S gs[Q];
        static_for<0, Q>([&](auto I) {
            constexpr int i = decltype(I)::value;
            if constexpr (CX[i] == 0 && ODD_NT) gs[i] = load_stream(at(base(OPP[i], yo[1 - CY[i]] + zo[1 - CZ[i]]), xb[1]));
            else gs[i] = *at(base(OPP[i], yo[1 - CY[i]] + zo[1 - CZ[i]]), xb[1 - CX[i]]);
        });
        T g[Q];
        static_for<0, Q>([&](auto I) {
            constexpr int i = decltype(I)::value;
            g[i] = St::load_add(gs[i], sh[i], K.one);
        });
        const NodeOut<T> out = collide_node<S, MODE>(A, K, g);
        static_for<0, Q>([&](auto I) {
            constexpr int i = decltype(I)::value;
            if constexpr (CX[i] == 0) {
                const S val = St::store(g[i]);
                if constexpr (ODD_NT) store_stream(at(base(i, yo[1 + CY[i]] + zo[1 + CZ[i]]), xb[1]), val);
                else *at(base(i, yo[1 + CY[i]] + zo[1 + CZ[i]]), xb[1]) = val;
            }
        });
        static_for<0, NXS>([&](auto K_) {
            constexpr int k = decltype(K_)::value;
            up[k] = St::store(g[XPLUS[k]]);
            dn[k] = St::store(g[XMINUS[k]]);
        });
        write_moments<S, MODE>(A, out, rg.n, x, (int)y, (int)z);
        excess = out.excess;
    }

    const int last = GENERAL ? (int)((NX - 1) & 63u) : 63;  // lane of the row's last node in its last wave
    const int lu = w == nw - 1 ? last : 63;                  // this wave's last node
    const unsigned long long act = __builtin_amdgcn_ballot_w64(active);
    const unsigned long long vm = GENERAL ? __builtin_amdgcn_ballot_w64(valid) : ~0ull;
    if (lane == lu) static_for<0, NXS>([&](auto K_) { edge_up[wb * NXS + decltype(K_)::value] = up[decltype(K_)::value]; });
    if (lane == 0) {
        static_for<0, NXS>([&](auto K_) { edge_dn[wb * NXS + decltype(K_)::value] = dn[decltype(K_)::value]; });
        fluid[wb] = act;
    }
    __syncthreads();
    // neighbours in the row (x periodic: the last wave's next is the first)
    const int wp = row_ok ? rb0 + (w == 0 ? nw - 1 : w - 1) : wb, wn = row_ok ? rb0 + (w + 1 == nw ? 0 : w + 1) : wb;
    // node x - 1 (x + 1) fluid: lane l stores for it at x.  The last node's
    // x + 1 is node 0 of the row, lane 0 of wave wn of the last wave.
    unsigned long long src_up = 0, src_dn = 0;
    if (row_ok) {
        const int lu_p = w == 0 ? last : 63;  // the last node of wave wp
        src_up = ((act << 1) | ((fluid[wp] >> lu_p) & 1ull)) & vm;
        const unsigned long long d = fluid[wn] & 1ull;
        src_dn = ((act >> 1) | (d << 63) | (GENERAL && w == nw - 1 ? d << last : 0ull)) & vm;
    }
    static_for<0, NXS>([&](auto K_) {
        constexpr int k = decltype(K_)::value;
        constexpr int i = XPLUS[k];
        const S v = lane_shift<true>(up[k], edge_up[wp * NXS + k]);
        if ((src_up >> lane) & 1ull) store_stream(at(base(i, yo[1 + CY[i]] + zo[1 + CZ[i]]), xb[1]), v);
    });
    static_for<0, NXS>([&](auto K_) {
        constexpr int k = decltype(K_)::value;
        constexpr int i = XMINUS[k];
        S v = lane_shift<false>(dn[k], edge_dn[wn * NXS + k]);
        if (GENERAL && (uint32_t)x == NX - 1) v = edge_dn[wn * NXS + k];  // a padded row's last node: from node 0
        if ((src_dn >> lane) & 1ull) store_stream(at(base(i, yo[1 + CY[i]] + zo[1 + CZ[i]]), xb[1]), v);
    });
    if constexpr (MASS) block_mass_sum(excess, mass_out, (int)(blockDim.x >> 6), mass_lds);
}

// Odd step, A.rows_per_block rows of A.row_waves wavefronts per block;
// `odd_rows` 2 gives each XCD a contiguous range of blocks (else consecutive
// blocks go to consecutive XCDs).
template <class S, int MODE, bool GENERAL>
__global__ __launch_bounds__(64 * ROW_MAX_WAVES) __attribute__((amdgpu_waves_per_eu(LBM_STEP_WAVES_PER_EU)))
void aa_odd_row_kernel(const StepArgs<S> A, const FastCoef K) {
    const int nb = (int)gridDim.x;
    const int lin = (int)blockIdx.x;
    const int b = A.odd_rows == 2 ? xcd_tile(lin, nb) : lin;
    odd_rows_block<S, MODE, GENERAL>(A, K, b * A.rows_per_block, A.rows_per_block, &A.mass_part[b]);
}

// Boundary chunks of an even step.  Work list: every 64-node chunk of the
// rows next to a wall (static, `wall_rows`), then every chunk of the rows of
// the squirmer's link box that are not wall rows (from the device state).
struct BoundaryArgs {
    const uint32_t* wall_rows;  // y + NY * z of each row next to a wall
    int n_wall_rows;
    int nchx;                   // 64-node chunks per row
    int box_rows_max, box_chunks;  // bound of box rows and chunks per box row
    int wy_lo, wy_hi, wz_lo, wz_hi;  // wall-adjacent coordinates (-1: none)
    MassPart* mass_part;        // partials after the step kernel's
};

template <class S, int MODE>
__device__ __forceinline__ void boundary_block(const StepArgs<S>& A, const FastCoef& K, const BoundaryArgs& B,
                                               int block) {
    constexpr bool MASS = (MODE & STEP_MASS) != 0;
    const double dsh = A.st->shift[A.t & 1];  // pending mass-correction shift (lbm.cpp:337)
    const uint32_t NX = (uint32_t)A.NX, NY = (uint32_t)A.NY, NZ = (uint32_t)A.NZ;
    const int wave = block * (BLOCK / 64) + (int)(threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    int y = -1, z = -1, chunk = 0;
    int row_solid = -1;
    const int nwall = B.n_wall_rows * B.nchx;
    if (wave < nwall) {
        const uint32_t r = B.wall_rows[wave / B.nchx];
        y = (int)(r % NY); z = (int)(r / NY);
        chunk = wave % B.nchx;
        // A wall-row chunk outside the box of the squirmer's links has the
        // walls as its only solid neighbours: the same directions for every
        // node.
        bool near_sphere = false;
        if (A.has_squirmer)
            near_sphere = in_link_box(boundary_box_dims(A.st, A.NX, A.NY, A.NZ), chunk, y, z, B.nchx, A.NY, A.NZ);
        if (!near_sphere) {
            uint32_t m = 0;
            for (int i = 1; i < Q; ++i) {
                const int yn = y + CY[i], zn = z + CZ[i];
                const bool wall = (A.walls_y && (yn == 0 || yn == A.NY - 1)) || (A.walls_z && (zn == 0 || zn == A.NZ - 1));
                m |= (uint32_t)wall << i;
            }
            row_solid = (int)m;
        }
    } else if (A.has_squirmer) {
        const int w = wave - nwall;
        const LinkBoxDims bd = boundary_box_dims(A.st, A.NX, A.NY, A.NZ);
        const int row = w / B.box_chunks, k = w % B.box_chunks;
        if (row < bd.wy * bd.wz && k < bd.nck) {
            const int iy = row % bd.wy, iz = row / bd.wy;
            y = wrap_coord(bd.y0 + iy, A.NY);
            z = wrap_coord(bd.z0 + iz, A.NZ);
            if (y == B.wy_lo || y == B.wy_hi || z == B.wz_lo || z == B.wz_hi) y = -1;  // a wall row: done above
            chunk = (bd.c0 + k) % B.nchx;
        }
    }
    y = __builtin_amdgcn_readfirstlane(y);
    z = __builtin_amdgcn_readfirstlane(z);
    chunk = __builtin_amdgcn_readfirstlane(chunk);
    row_solid = __builtin_amdgcn_readfirstlane(row_solid);
    const int x = chunk * 64 + lane;
    const bool inb = y >= 0 && (uint32_t)x < NX;
    const RowGeom rg = row_geom<S>(inb ? (uint32_t)x : 0u, y >= 0 ? (uint32_t)y : 0u, y >= 0 ? (uint32_t)z : 0u,
                                   NX, NY, NZ, A.PX);
    const uint8_t fl = inb ? A.flag[rg.n] : (uint8_t)WALL;
    bool active = (fl & TYPE_MASK) == FLUID;
    if (__builtin_amdgcn_ballot_w64(active && (fl & BOUNDARY)) == 0) active = false;  // not a boundary chunk
    double excess = 0.0;
    if (active) excess = node_update<S, false, MODE, true>(A, K, rg, x, y, z, fl, dsh, row_solid);
    if constexpr (MASS) block_mass_sum(excess, &B.mass_part[block]);
}

template <class S, int MODE>
__global__ __launch_bounds__(BLOCK) void aa_boundary_kernel(const StepArgs<S> A, const FastCoef K,
                                                            const BoundaryArgs B) {
    boundary_block<S, MODE>(A, K, B, (int)blockIdx.x);
}

}  // namespace lbmhip
