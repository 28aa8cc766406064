// mlp.hip -- the learned velocity surrogate (include/lbm_mlp.h): an ensemble
// of 5 -> 128 -> 128 -> 128 -> 6 tanh MLPs, forward, backward and Adam in
// hand-written HIP.
//
// BF16 mode (mlp_block_kernel): a block of 4 wavefronts takes BS = 32
// samples of one net through forward and backward; one launch covers every
// net and every block of a minibatch.  Every product is built from 16 x 16
// output tiles accumulated by the gfx950 matrix instruction
// v_mfma_f32_16x16x32_bf16 (A 16 x 32, B 32 x 16 in bfloat16, FP32
// accumulator).  Activations and backpropagated errors are kept in LDS as
// bfloat16 twice, sample-major ([sample][unit]: the A operand of the layer
// products) and unit-major ([unit][sample]: both operands of the weight
// gradients, which contract over the samples), so each lane fetches its 8
// operand values as one 16-byte LDS read.  The weights come as bfloat16
// copies kept beside the FP32 master weights, row-major for the forward
// products and transposed for the backward ones, each fragment one 16-byte
// load (they stay in L2; every block reads them).  The tanh derivatives
// stay in registers in the layout of the tiles that produced them.  Each
// block writes one gradient partial; a second kernel adds the blocks'
// partials in block order (deterministic), a third applies Adam to the FP32
// master weights and refreshes the bfloat16 copies.
//
// FP32 mode (mlp_tile_kernel): one wavefront per 16 samples, the same tile
// structure with FP32 multiply-adds, for exact gradient checks.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <algorithm>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "lbm_mlp.h"

namespace {

constexpr int IN = LBM_MLP_IN, HID = LBM_MLP_HIDDEN, OUT = LBM_MLP_OUT, P = LBM_MLP_PARAMS;
constexpr int TILE = 16;
constexpr int HS = HID + 4;  // LDS row stride of a 128-wide activation tile (no bank conflicts)
constexpr int SS16 = 16 + 4; // LDS row stride of a 16-wide tile

// parameter offsets
constexpr int OW1 = 0, OB1 = OW1 + HID * IN, OW2 = OB1 + HID, OB2 = OW2 + HID * HID, OW3 = OB2 + HID,
              OB3 = OW3 + HID * HID, OW4 = OB3 + HID, OB4 = OW4 + OUT * HID;
static_assert(OB4 + OUT == P, "parameter layout");

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// BF16 copies of the parameters: row-major as the FP32 master ([nets][PP],
// per-net stride rounded up to 8 values so every 16-byte fragment is
// aligned), and the transposes of W2, W3, W4 for the backward products
// ([nets][PT]: W2^T and W3^T [128][128], W4^T [128][8] with outputs 6, 7 zero).
constexpr int PP = (P + 7) / 8 * 8;
constexpr int T2 = 0, T3 = HID * HID, T4 = 2 * HID * HID, PT = 2 * HID * HID + HID * 8;

__device__ __forceinline__ uint16_t to_bf16(float f) {
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);  // round to nearest even (finite values)
}
__device__ __forceinline__ float from_bf16(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// FP32 path.  C (M x N, in 16 x 16 blocks) = A (M x K) * B (K x N); M, N,
// K multiples of 16 (the fetchers return 0 outside the real sizes).  Lane l
// of the block (mt, nt) receives C[mt*16 + 4*(l/16) + j][nt*16 + l%16],
// j = 0..3, and hands each to `emit`.
template <class FA, class FB, class FE>
__device__ __forceinline__ void wave_gemm(int M, int N, int K, FA&& a, FB&& b, FE&& emit) {
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, q = lane >> 4;
    for (int mt = 0; mt < M / 16; ++mt)
        for (int nt = 0; nt < N / 16; ++nt) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            for (int k = 0; k < K; ++k) {
                const float bk = b(k, nt * 16 + r16);
                for (int j = 0; j < 4; ++j) acc[j] = fmaf(a(mt * 16 + 4 * q + j, k), bk, acc[j]);
            }
            for (int j = 0; j < 4; ++j) emit(mt * 16 + 4 * q + j, nt * 16 + r16, acc[j]);
        }
}

struct TileArgs {
    const float* x;      // [nsrc][5]
    const float* t;      // [nsrc][6] targets, or null (forward only)
    const int* idx;      // sample of row s of net (net0 + blockIdx.y) is idx[blockIdx.y * idx_stride + s], or s if null
    int idx_stride;
    int n;               // samples of this pass
    int net0;            // first net of the launch (grid.y counts nets from here)
    float* y;            // [nets][n][6] outputs, or null
    const float* wf;     // [nets][P] FP32 parameters
    const uint16_t* wb;  // [nets][PP] BF16 copy (row-major layers)
    const uint16_t* wt;  // [nets][PT] BF16 transposed W2, W3, W4
    float* part;         // [nets][units][P] gradient partials, or null
    float* loss_part;    // [nets][units] loss partials, or null
    float scale;         // dL/dy = scale * (y - t)
    int units;           // blocks (BF16) or tiles (FP32) per net in this pass
};

__device__ __forceinline__ int sample_of(const TileArgs& A, int s) {
    return A.idx ? A.idx[(size_t)blockIdx.y * A.idx_stride + s] : s;
}

__global__ __launch_bounds__(64) void mlp_tile_kernel(const TileArgs A) {
    __shared__ float xin[TILE][SS16];
    __shared__ float h[3][TILE][HS];
    __shared__ float yo[TILE][SS16];
    __shared__ float dz[2][TILE][HS];
    __shared__ float dy[TILE][SS16];
    const int lane = threadIdx.x;
    const int tile = blockIdx.x, net = A.net0 + (int)blockIdx.y;
    const int row0 = tile * TILE;
    const int nrows = A.n - row0 < TILE ? A.n - row0 : TILE;
    const float* wf = A.wf + (size_t)net * P;
    auto W = [&](int off) -> float { return wf[off]; };

    // ---- inputs ----
    for (int e = lane; e < TILE * 16; e += 64) {
        const int r = e >> 4, c = e & 15;
        xin[r][c] = (r < nrows && c < IN) ? A.x[(size_t)sample_of(A, row0 + r) * IN + c] : 0.f;
    }
    __syncthreads();

    // ---- forward ----
    auto dense_tanh = [&](auto&& act_in, int kin, int wo, int bo, float (*out)[HS]) {
        wave_gemm(TILE, HID, kin, act_in, [&](int k, int o) -> float { return k < (kin == 16 ? IN : HID) ? W(wo + o * (kin == 16 ? IN : HID) + k) : 0.f; },
                  [&](int r, int o, float v) { out[r][o] = tanhf(v + wf[bo + o]); });
        __syncthreads();
    };
    dense_tanh([&](int r, int k) -> float { return xin[r][k]; }, 16, OW1, OB1, h[0]);
    dense_tanh([&](int r, int k) -> float { return h[0][r][k]; }, HID, OW2, OB2, h[1]);
    dense_tanh([&](int r, int k) -> float { return h[1][r][k]; }, HID, OW3, OB3, h[2]);
    wave_gemm(TILE, 16, HID, [&](int r, int k) -> float { return h[2][r][k]; },
              [&](int k, int o) -> float { return o < OUT ? W(OW4 + o * HID + k) : 0.f; },
              [&](int r, int o, float v) { yo[r][o] = o < OUT ? v + wf[OB4 + o] : 0.f; });
    __syncthreads();
    if (A.y)
        for (int e = lane; e < TILE * OUT; e += 64) {
            const int r = e / OUT, c = e % OUT;
            if (r < nrows) A.y[((size_t)blockIdx.y * A.n + row0 + r) * OUT + c] = yo[r][c];
        }
    if (!A.t) return;

    // ---- loss and dL/dy ----
    float l = 0.f;
    for (int e = lane; e < TILE * 16; e += 64) {
        const int r = e >> 4, c = e & 15;
        float d = 0.f;
        if (r < nrows && c < OUT) {
            d = yo[r][c] - A.t[(size_t)sample_of(A, row0 + r) * OUT + c];
            l += d * d;
        }
        dy[r][c] = A.scale * d;
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
    if (A.loss_part && lane == 0) A.loss_part[(size_t)blockIdx.y * A.units + tile] = l;
    __syncthreads();
    if (!A.part) return;
    float* g = A.part + ((size_t)blockIdx.y * A.units + tile) * P;

    // ---- backward ----
    // layer 4: dW4 = dy^T h3, db4 = sum dy, dh3 = dy W4 -> dz[0]
    wave_gemm(16, HID, TILE, [&](int o, int r) -> float { return dy[r][o]; },
              [&](int r, int i) -> float { return h[2][r][i]; },
              [&](int o, int i, float v) { if (o < OUT) g[OW4 + o * HID + i] = v; });
    if (lane < OUT) {
        float s = 0.f;
        for (int r = 0; r < TILE; ++r) s += dy[r][lane];
        g[OB4 + lane] = s;
    }
    wave_gemm(TILE, HID, 16, [&](int r, int o) -> float { return dy[r][o]; },
              [&](int o, int i) -> float { return o < OUT ? W(OW4 + o * HID + i) : 0.f; },
              [&](int r, int i, float v) { dz[0][r][i] = v * (1.f - h[2][r][i] * h[2][r][i]); });
    __syncthreads();
    // layers 3 and 2: dW = dz^T h_prev, db = sum dz, dh_prev = dz W -> through tanh
    auto hidden_back = [&](int wo, int bo, int hl, int cur) {
        float(*z)[HS] = dz[cur];
        float(*hp)[HS] = h[hl - 1];
        wave_gemm(HID, HID, TILE, [&](int o, int r) -> float { return z[r][o]; },
                  [&](int r, int i) -> float { return hp[r][i]; },
                  [&](int o, int i, float v) { g[wo + o * HID + i] = v; });
        for (int o = lane; o < HID; o += 64) {
            float s = 0.f;
            for (int r = 0; r < TILE; ++r) s += z[r][o];
            g[bo + o] = s;
        }
        float(*zn)[HS] = dz[cur ^ 1];
        wave_gemm(TILE, HID, HID, [&](int r, int o) -> float { return z[r][o]; },
                  [&](int o, int i) -> float { return W(wo + o * HID + i); },
                  [&](int r, int i, float v) { zn[r][i] = v * (1.f - hp[r][i] * hp[r][i]); });
        __syncthreads();
    };
    hidden_back(OW3, OB3, 2, 0);  // dz[0] (layer 3) -> dz[1] (layer 2)
    hidden_back(OW2, OB2, 1, 1);  // dz[1] (layer 2) -> dz[0] (layer 1)
    // layer 1: dW1 = dz1^T x, db1
    wave_gemm(HID, 16, TILE, [&](int o, int r) -> float { return dz[0][r][o]; },
              [&](int r, int i) -> float { return xin[r][i]; },
              [&](int o, int i, float v) { if (i < IN) g[OW1 + o * IN + i] = v; });
    for (int o = lane; o < HID; o += 64) {
        float s = 0.f;
        for (int r = 0; r < TILE; ++r) s += dz[0][r][o];
        g[OB1 + o] = s;
    }
}

// ---------------------------------------------------------------------------
// BF16 path: blocks of NW wavefronts, BS samples of one net.
// ---------------------------------------------------------------------------
constexpr int BS = 32;         // samples per block
constexpr int HP = HID + 8;    // bf16 row stride of [sample][unit] tiles: 272 B, 16-byte aligned rows
constexpr int SP = BS + 8;     // bf16 row stride of [unit][sample] tiles: 80 B
#ifndef LBM_MLP_BLOCK_WAVES
#define LBM_MLP_BLOCK_WAVES 16
#endif
constexpr int BLOCK_WAVES = LBM_MLP_BLOCK_WAVES;  // wavefronts per block (8 or 16)

__device__ __forceinline__ bf16x8 lds8(const uint16_t* p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(p));
}
__device__ __forceinline__ bf16x8 gld8(const uint16_t* p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const u32x4*>(p));
}
__device__ __forceinline__ bf16x8 zero8() {
    const u32x4 z = {0u, 0u, 0u, 0u};
    return __builtin_bit_cast(bf16x8, z);
}
__device__ __forceinline__ f32x4 mma(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// Lane l of a 16 x 16 x 32 product holds A[l % 16][8 (l / 16) + 0..7] and
// B[8 (l / 16) + 0..7][l % 16] and receives C[4 (l / 16) + 0..3][l % 16].
//
// NW wavefronts share the tiles of each product:
//   a layer output (32 samples x 128 units, 16 tiles): wave w owns tiles
//     f = w * TF + q, q < TF = 16 / NW: sample rows 16 (f / 8), units 16 (f % 8);
//   a 128 x 128 weight gradient (64 tiles): wave w owns d = w * TD + u,
//     u < TD = 64 / NW: output units 16 (d / 8), input units 16 (d % 8);
//   the 16 x 128 gradient of W4 and the 128 x 16 gradient of W1 (8 tiles
//     each) and the 32 x 16 outputs (2 tiles) go round the waves.
// The block is a chain of layers separated by barriers; every global load
// is issued as early as its registers allow, so the chain waits on memory
// twice instead of once per layer: the forward weights, biases, inputs and
// targets at the start, the backward (transposed) weights while the last
// forward layers run.
template <int NW>
__global__ __launch_bounds__(64 * NW) void mlp_block_kernel(const TileArgs A) {
    static_assert(NW == 8 || NW == 16, "wavefronts per block");
    constexpr int NT = 64 * NW, TF = 16 / NW, TD = 64 / NW, KT = HID / 32;
    static_assert(NT >= BS * 16, "one thread per (sample, output) pair of the loss");
    __shared__ __attribute__((aligned(16))) uint16_t xs[BS][8];    // x as bf16, inputs 5..7 zero
    __shared__ __attribute__((aligned(16))) uint16_t xT[8][SP];    // x^T
    __shared__ __attribute__((aligned(16))) uint16_t hs[3][BS][HP];
    __shared__ __attribute__((aligned(16))) uint16_t hT[3][HID][SP];
    // backpropagated errors, sample-major, in the space of h1 and h2 (read
    // only by the forward pass, which has ended when the first is written)
    uint16_t (*const zs)[BS][HP] = hs;
    __shared__ __attribute__((aligned(16))) uint16_t zT[2][HID][SP];
    __shared__ __attribute__((aligned(16))) uint16_t dys[BS][16];  // dL/dy as bf16, outputs 6..15 zero
    __shared__ __attribute__((aligned(16))) uint16_t dyT[16][SP];
    __shared__ float dyf[BS][8];      // dL/dy in FP32 (bias gradient)
    __shared__ float yo[BS][8];
    __shared__ float dbp[3][2][HID];  // per layer, per 16 samples: bias gradient partials
    __shared__ float lsum[NW];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, r16 = lane & 15, g = lane >> 4;
    const int blk = blockIdx.x, net = A.net0 + (int)blockIdx.y;
    const int row0 = blk * BS;
    const int nrows = A.n - row0 < BS ? A.n - row0 : BS;
    const float* wf = A.wf + (size_t)net * P;
    const uint16_t* wb = A.wb + (size_t)net * PP;
    const uint16_t* wt = A.wt + (size_t)net * PT;
    const f32x4 zero4 = {0.f, 0.f, 0.f, 0.f};
    const int mt = (w * TF) >> 3;  // sample-row tile of this wave's layer-output tiles
    // Unit column of this lane in layer-output tile q of this wave.
    auto unit = [&](int q) { return ((w * TF + q) & 7) * 16 + r16; };

    // ---- loads of the forward pass ----
    // the input of thread tid < 256 and the target of thread tid < 512
    const int xr = tid >> 3, xc = tid & 7, tr = tid >> 4, tc = tid & 15;
    float xv = 0.f, tv = 0.f;
    if (tid < BS * 8 && xr < nrows && xc < IN) xv = A.x[(size_t)sample_of(A, row0 + xr) * IN + xc];
    if (A.t && tid < BS * 16 && tr < nrows && tc < OUT) tv = A.t[(size_t)sample_of(A, row0 + tr) * OUT + tc];
    // weight fragments: B[k][n] = W[n][k], lane (r16, g) holds W[unit][32 kt + 8 g + 0..7]
    bf16x8 f1[TF], f2[KT][TF], f3[KT][TF], f4[KT];
    float bias[3][TF], b4 = 0.f;
#pragma unroll
    for (int q = 0; q < TF; ++q) {
        const int o = unit(q);
        uint16_t t[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        if (g == 0)
            for (int k = 0; k < IN; ++k) t[k] = wb[OW1 + o * IN + k];
        const u32x4 uv = {(uint32_t)t[0] | ((uint32_t)t[1] << 16), (uint32_t)t[2] | ((uint32_t)t[3] << 16),
                          (uint32_t)t[4] | ((uint32_t)t[5] << 16), (uint32_t)t[6] | ((uint32_t)t[7] << 16)};
        f1[q] = __builtin_bit_cast(bf16x8, uv);
#pragma unroll
        for (int kt = 0; kt < KT; ++kt) {
            f2[kt][q] = gld8(wb + OW2 + o * HID + kt * 32 + 8 * g);
            f3[kt][q] = gld8(wb + OW3 + o * HID + kt * 32 + 8 * g);
        }
        bias[0][q] = wf[OB1 + o];
        bias[1][q] = wf[OB2 + o];
        bias[2][q] = wf[OB3 + o];
    }
#pragma unroll
    for (int kt = 0; kt < KT; ++kt) f4[kt] = (w < 2 && r16 < OUT) ? gld8(wb + OW4 + r16 * HID + kt * 32 + 8 * g) : zero8();
    if (w < 2 && r16 < OUT) b4 = wf[OB4 + r16];

    if (tid < BS * 8) {
        const uint16_t b = to_bf16(xv);
        xs[xr][xc] = b;
        xT[xc][xr] = b;
    }
    __syncthreads();

    // Writes this wave's layer-output tiles (rows 16 mt + 4 g + j) into the
    // [sample][unit] and [unit][sample] bf16 arrays.
    auto store_tiles = [&](uint16_t (*rows)[HP], uint16_t (*cols)[SP], const f32x4 (&v)[TF]) {
#pragma unroll
        for (int q = 0; q < TF; ++q) {
            const int o = unit(q), s0 = mt * 16 + 4 * g;
            uint16_t b[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                b[j] = to_bf16(v[q][j]);
                rows[s0 + j][o] = b[j];
            }
            // four consecutive samples of one unit: one 8-byte write
            *reinterpret_cast<uint2*>(&cols[o][s0]) =
                make_uint2((uint32_t)b[0] | ((uint32_t)b[1] << 16), (uint32_t)b[2] | ((uint32_t)b[3] << 16));
        }
    };
    // Column sums over this wave's 16 samples, per unit, into dbp[layer][mt].
    auto bias_partial = [&](int layer, const f32x4 (&v)[TF]) {
#pragma unroll
        for (int q = 0; q < TF; ++q) {
            float sum = (v[q][0] + v[q][1]) + (v[q][2] + v[q][3]);
            sum += __shfl_xor(sum, 16, 64);
            sum += __shfl_xor(sum, 32, 64);
            if (g == 0) dbp[layer][mt][unit(q)] = sum;
        }
    };
    // acc[q] = rows(16 mt.., K = 128) x the fragments of tile q
    auto dense128 = [&](f32x4 (&acc)[TF], const uint16_t (*rows)[HP], const bf16x8 (&f)[KT][TF]) {
#pragma unroll
        for (int q = 0; q < TF; ++q) acc[q] = zero4;
#pragma unroll
        for (int kt = 0; kt < KT; ++kt) {
            const bf16x8 a = lds8(&rows[mt * 16 + r16][kt * 32 + 8 * g]);
#pragma unroll
            for (int q = 0; q < TF; ++q) acc[q] = mma(a, f[kt][q], acc[q]);
        }
    };
    // h = tanh(acc + bias), dd = 1 - h^2, stored as layer `layer`'s output
    f32x4 dd[3][TF];  // tanh derivatives of this wave's tiles, layers 1..3
    auto activate = [&](int layer, f32x4 (&acc)[TF]) {
#pragma unroll
        for (int q = 0; q < TF; ++q)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const float h = tanhf(acc[q][j] + bias[layer][q]);
                acc[q][j] = h;
                dd[layer][q][j] = 1.f - h * h;
            }
        store_tiles(hs[layer], hT[layer], acc);
    };

    // ---- forward ----
    {
        // layer 1: K = 5 inputs, in the first 8 of the 32
        const bf16x8 a = g == 0 ? lds8(&xs[mt * 16 + r16][0]) : zero8();
        f32x4 acc[TF];
#pragma unroll
        for (int q = 0; q < TF; ++q) acc[q] = mma(a, f1[q], zero4);
        activate(0, acc);
    }
    __syncthreads();
    {
        f32x4 acc[TF];
        dense128(acc, hs[0], f2);
        activate(1, acc);
    }
    __syncthreads();
    {
        f32x4 acc[TF];
        dense128(acc, hs[1], f3);
        activate(2, acc);
    }
    // the backward weights (W2, W3, W4 transposed), in flight while layer 4
    // and the loss run
    bf16x8 g4[TF], g3[KT][TF], g2[KT][TF];
    if (A.part) {
#pragma unroll
        for (int q = 0; q < TF; ++q) {
            g4[q] = g == 0 ? gld8(wt + T4 + unit(q) * 8) : zero8();
#pragma unroll
            for (int kt = 0; kt < KT; ++kt) {
                g3[kt][q] = gld8(wt + T3 + unit(q) * HID + kt * 32 + 8 * g);
                g2[kt][q] = gld8(wt + T2 + unit(q) * HID + kt * 32 + 8 * g);
            }
        }
    }
    __syncthreads();
    // layer 4: 6 outputs, one 16-wide tile per 16 samples
    if (w < 2) {
        f32x4 acc = zero4;
#pragma unroll
        for (int kt = 0; kt < KT; ++kt) acc = mma(lds8(&hs[2][w * 16 + r16][kt * 32 + 8 * g]), f4[kt], acc);
        if (r16 < 8)
#pragma unroll
            for (int j = 0; j < 4; ++j) yo[w * 16 + 4 * g + j][r16] = r16 < OUT ? acc[j] + b4 : 0.f;
    }
    __syncthreads();
    if (A.y)
        for (int e = tid; e < BS * OUT; e += NT) {
            const int r = e / OUT, c = e % OUT;
            if (r < nrows) A.y[((size_t)blockIdx.y * A.n + row0 + r) * OUT + c] = yo[r][c];
        }
    if (!A.t) return;

    // ---- loss and dL/dy ----
    float l = 0.f;
    if (tid < BS * 16) {
        float d = 0.f;
        if (tr < nrows && tc < OUT) {
            d = yo[tr][tc] - tv;
            l = d * d;
        }
        const float dv = A.scale * d;
        const uint16_t b = to_bf16(dv);
        dys[tr][tc] = b;
        dyT[tc][tr] = b;
        if (tc < 8) dyf[tr][tc] = dv;
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
    if (lane == 0) lsum[w] = l;
    __syncthreads();
    if (A.loss_part && tid == 0) {
        float sum = 0.f;
        for (int k = 0; k < NW; ++k) sum += lsum[k];
        A.loss_part[(size_t)blockIdx.y * A.units + blk] = sum;
    }
    if (!A.part) return;
    float* gp = A.part + ((size_t)blockIdx.y * A.units + blk) * P;

    // ---- backward ----
    // dW4 = dy^T h3 (16 x 128, K = the 32 samples)
    for (int nt = w; nt < HID / 16; nt += NW) {
        const f32x4 acc = mma(lds8(&dyT[r16][8 * g]), lds8(&hT[2][nt * 16 + r16][8 * g]), zero4);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int o = 4 * g + j;
            if (o < OUT) gp[OW4 + o * HID + nt * 16 + r16] = acc[j];
        }
    }
    if (tid < OUT) {
        float sum = 0.f;
        for (int r = 0; r < BS; ++r) sum += dyf[r][tid];
        gp[OB4 + tid] = sum;
    }
    // dz3 = (dy W4) .* (1 - h3^2): K = the 6 outputs, in the first 8 of 32
    {
        const bf16x8 a = g == 0 ? lds8(&dys[mt * 16 + r16][0]) : zero8();
        f32x4 acc[TF];
#pragma unroll
        for (int q = 0; q < TF; ++q) {
            acc[q] = mma(a, g4[q], zero4);
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[q][j] *= dd[2][q][j];
        }
        store_tiles(zs[0], zT[0], acc);
        bias_partial(2, acc);
    }
    __syncthreads();
    // layers 3 and 2: dW = dz^T h_prev (128 x 128, K = 32), db = the column
    // sums, dz_prev = (dz W) .* (1 - h_prev^2)
    auto hidden_back = [&](int c, int ow, int ob, const bf16x8 (&f)[KT][TF]) {
        const int hl = 1 - c;  // index of h_prev
#pragma unroll
        for (int u = 0; u < TD; ++u) {
            const int d = w * TD + u, mo = d >> 3, nt = d & 7;
            const f32x4 acc = mma(lds8(&zT[c][mo * 16 + r16][8 * g]), lds8(&hT[hl][nt * 16 + r16][8 * g]), zero4);
#pragma unroll
            for (int j = 0; j < 4; ++j) gp[ow + (mo * 16 + 4 * g + j) * HID + nt * 16 + r16] = acc[j];
        }
        if (tid < HID) gp[ob + tid] = dbp[2 - c][0][tid] + dbp[2 - c][1][tid];
        f32x4 acc[TF];
        dense128(acc, zs[c], f);
#pragma unroll
        for (int q = 0; q < TF; ++q)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[q][j] *= dd[hl][q][j];
        store_tiles(zs[c ^ 1], zT[c ^ 1], acc);
        bias_partial(hl, acc);
    };
    hidden_back(0, OW3, OB3, g3);  // dz3 -> dW3, db3, dz2
    __syncthreads();
    hidden_back(1, OW2, OB2, g2);  // dz2 -> dW2, db2, dz1
    __syncthreads();
    // dW1 = dz1^T x (128 x 5, K = 32); db1
    for (int mo = w; mo < HID / 16; mo += NW) {
        const f32x4 acc = mma(lds8(&zT[0][mo * 16 + r16][8 * g]), r16 < 8 ? lds8(&xT[r16][8 * g]) : zero8(), zero4);
        if (r16 < IN)
#pragma unroll
            for (int j = 0; j < 4; ++j) gp[OW1 + (mo * 16 + 4 * g + j) * IN + r16] = acc[j];
    }
    if (tid < HID) gp[OB1 + tid] = dbp[0][0][tid] + dbp[0][1][tid];
}

// grad[net][p] = sum over tiles, in tile order
__global__ __launch_bounds__(256) void mlp_reduce_kernel(const float* part, int tiles, float* grad) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= P) return;
    const float* src = part + (size_t)blockIdx.y * tiles * P + p;
    float s = 0.f;
    for (int k = 0; k < tiles; ++k) s += src[(size_t)k * P];
    grad[(size_t)blockIdx.y * P + p] = s;
}

// The BF16 copies of parameter p of net k (row-major and, for W2-W4, transposed).
__device__ __forceinline__ void publish_bf16(uint16_t* wb, uint16_t* wt, int k, int p, float v) {
    const uint16_t b = to_bf16(v);
    wb[(size_t)k * PP + p] = b;
    uint16_t* t = wt + (size_t)k * PT;
    if (p >= OW2 && p < OW2 + HID * HID) t[T2 + ((p - OW2) % HID) * HID + (p - OW2) / HID] = b;
    else if (p >= OW3 && p < OW3 + HID * HID) t[T3 + ((p - OW3) % HID) * HID + (p - OW3) / HID] = b;
    else if (p >= OW4 && p < OW4 + OUT * HID) t[T4 + ((p - OW4) % HID) * 8 + (p - OW4) / HID] = b;
}

// One training step of every net: the gradient of parameter p of net
// blockIdx.y is the sum of the block partials in block order (as
// mlp_reduce_kernel), then Adam (bias-corrected moments, c1 = 1 - b1^t,
// c2 = 1 - b2^t) updates the FP32 master and its BF16 copies.
__global__ __launch_bounds__(256) void mlp_reduce_adam_kernel(float* w, uint16_t* wb, uint16_t* wt, float* m, float* v,
                                                              const float* part, int units, float lr, float b1,
                                                              float b2, float eps, float c1, float c2) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= P) return;
    const int k = blockIdx.y;
    const float* src = part + (size_t)k * units * P + p;
    float gr = 0.f;
    for (int u = 0; u < units; ++u) gr += src[(size_t)u * P];
    const size_t e = (size_t)k * P + p;
    const float mm = b1 * m[e] + (1.f - b1) * gr;
    const float vv = b2 * v[e] + (1.f - b2) * gr * gr;
    m[e] = mm;
    v[e] = vv;
    const float nw = w[e] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    w[e] = nw;
    publish_bf16(wb, wt, k, p, nw);
}

// BF16 copies of parameters [first, first + total) of the flat [nets][P] array.
__global__ __launch_bounds__(256) void mlp_cast_kernel(const float* w, uint16_t* wb, uint16_t* wt, int first, int total) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e < total) publish_bf16(wb, wt, (first + e) / P, (first + e) % P, w[first + e]);
}

// Ensemble statistics of y[nets][n][6]: mean over the nets -> mean[n][6];
// spread[s] = the largest standard deviation over the nets among the six
// outputs / the largest |mean| among them.  In FP64 without contraction,
// one thread per sample.
__global__ __launch_bounds__(256) void mlp_stats_kernel(const float* y, int nets, int n, float* mean, float* spread) {
#pragma clang fp contract(off)
    const int s = blockIdx.x * 256 + threadIdx.x;
    if (s >= n) return;
    double big_mean = 0.0, big_sd = 0.0;
    for (int c = 0; c < OUT; ++c) {
        double mu = 0.0;
        for (int k = 0; k < nets; ++k) mu += (double)y[((size_t)k * n + s) * OUT + c];
        mu /= nets;
        double var = 0.0;
        for (int k = 0; k < nets; ++k) {
            const double d = (double)y[((size_t)k * n + s) * OUT + c] - mu;
            var += d * d;
        }
        const double sd = nets > 1 ? sqrt(var / (nets - 1)) : 0.0;
        mean[(size_t)s * OUT + c] = (float)mu;
        big_mean = fmax(big_mean, fabs(mu));
        big_sd = fmax(big_sd, sd);
    }
    spread[s] = (float)(big_mean > 0.0 ? big_sd / big_mean : (big_sd > 0.0 ? INFINITY : 0.0));
}

// Layout of the device memory of lbm_mlp_predict_noalloc (`scratch` below):
// inputs, per-net outputs, mean and spread of up to LBM_MLP_NOALLOC_MAX_SAMPLES
// samples for up to 64 nets (6.5 MB).
constexpr int NA_MAX = LBM_MLP_NOALLOC_MAX_SAMPLES;
constexpr int MAX_NETS = 64;
constexpr size_t NA_X = 0, NA_Y = NA_X + (size_t)NA_MAX * IN, NA_MEAN = NA_Y + (size_t)MAX_NETS * NA_MAX * OUT,
                 NA_SPREAD = NA_MEAN + (size_t)NA_MAX * OUT, NA_FLOATS = NA_SPREAD + NA_MAX;
static_assert(NA_FLOATS * sizeof(float) == LBM_MLP_NOALLOC_SCRATCH_BYTES, "scratch size in lbm_mlp.h");

}  // namespace

// One copy per GPU, for the life of the process.  Host code uses the address
// hipGetSymbolAddress returns for the current device (noalloc_scratch()),
// never this name.
static __device__ __align__(256) float scratch[NA_FLOATS];

// The error text is the one lbm_hip_last_error returns (lbm_hip.hip).
void lbmhip_set_error_text(const char* msg);

struct lbm_mlp {
    int nets = 0, mode = 0, device = 0;
    unsigned long long seed = 0;
    hipStream_t stream = nullptr;
    float* w = nullptr;      // [nets][P]
    uint16_t* wb = nullptr;  // [nets][PP] BF16 copy
    uint16_t* wt = nullptr;  // [nets][PT] BF16 transposes of W2, W3, W4
    float* m = nullptr;      // Adam moments
    float* v = nullptr;
    float* grad = nullptr;   // [nets][P]
    long long adam_t = 0;
    // scratch sized for the largest pass so far
    float* part = nullptr;
    size_t part_cap = 0;  // floats
    float* loss_part = nullptr;
    size_t loss_cap = 0;
    float* x = nullptr;
    float* t = nullptr;
    float* y = nullptr;
    float* mean = nullptr;    // [n][6] and [n] of lbm_mlp_predict
    float* spread = nullptr;
    size_t xcap = 0, tcap = 0, ycap = 0, mcap = 0, scap = 0;
    float* noalloc = nullptr;  // this device's copy of `scratch` (device address)
};

namespace {

int mfail(int code, const std::string& m) {
    lbmhip_set_error_text(m.c_str());
    return code;
}

#define MLP_TRY(expr)                                                                               \
    do {                                                                                            \
        hipError_t e_ = (expr);                                                                     \
        if (e_ != hipSuccess) return mfail(LBM_HIP_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
int ensure(T*& p, size_t& cap, size_t need) {
    if (need <= cap) return LBM_HIP_OK;
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
    if (hipMalloc((void**)&p, need * sizeof(T)) != hipSuccess) return mfail(LBM_HIP_ENOMEM, "MLP scratch allocation failed");
    cap = need;
    return LBM_HIP_OK;
}

// Device address of `scratch` on `device` (current), resolved once per device.
float* noalloc_scratch(int device, int* rc) {
    static std::mutex mu;
    static float* addr[256] = {};
    std::lock_guard<std::mutex> lock(mu);
    if (device < 0 || device >= 256) { *rc = mfail(LBM_HIP_EINVAL, "device ordinal out of range"); return nullptr; }
    if (!addr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            *rc = mfail(LBM_HIP_EDEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return nullptr;
        }
        addr[device] = static_cast<float*>(p);
    }
    *rc = LBM_HIP_OK;
    return addr[device];
}

int launch_stats(lbm_mlp* M, const float* y, int n, float* mean, float* spread) {
    hipLaunchKernelGGL(mlp_stats_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, M->stream, y, M->nets, n,
                       mean, spread);
    MLP_TRY(hipGetLastError());
    return LBM_HIP_OK;
}

// Blocks (BF16) or wavefront tiles (FP32) that cover n samples of one net.
int units_of(const lbm_mlp* M, int n) {
    const int per = M->mode == LBM_MLP_BF16 ? BS : TILE;
    return (n + per - 1) / per;
}

// Forward (and, with targets, backward) of a.n samples for nets
// a.net0 .. a.net0 + nets - 1, all in one launch.
int launch_tiles(lbm_mlp* M, TileArgs a, int nets) {
    a.wf = M->w;
    a.wb = M->wb;
    a.wt = M->wt;
    a.units = units_of(M, a.n);
    if (a.units == 0) return LBM_HIP_OK;
    if (a.net0 < 0 || nets < 1 || a.net0 + nets > M->nets) return mfail(LBM_HIP_EINVAL, "net range out of bounds");
    if (M->mode == LBM_MLP_BF16)
        hipLaunchKernelGGL(mlp_block_kernel<BLOCK_WAVES>, dim3(a.units, nets), dim3(64 * BLOCK_WAVES), 0, M->stream, a);
    else
        hipLaunchKernelGGL(mlp_tile_kernel, dim3(a.units, nets), dim3(64), 0, M->stream, a);
    MLP_TRY(hipGetLastError());
    return LBM_HIP_OK;
}

int upload_xt(lbm_mlp* M, const float* x, const float* t, int n) {
    int rc;
    if ((rc = ensure(M->x, M->xcap, (size_t)n * IN + 1))) return rc;
    MLP_TRY(hipMemcpyAsync(M->x, x, sizeof(float) * n * IN, hipMemcpyHostToDevice, M->stream));
    if (t) {
        if ((rc = ensure(M->t, M->tcap, (size_t)n * OUT + 1))) return rc;
        MLP_TRY(hipMemcpyAsync(M->t, t, sizeof(float) * n * OUT, hipMemcpyHostToDevice, M->stream));
    }
    return LBM_HIP_OK;
}

}  // namespace

extern "C" {

int lbm_mlp_create(int nets, int mode, unsigned long long seed, int device, lbm_mlp** out) {
    if (!out) return mfail(LBM_HIP_EINVAL, "null output pointer");
    *out = nullptr;
    if (nets < 1 || nets > 64) return mfail(LBM_HIP_EINVAL, "nets must be in [1, 64]");
    if (mode != LBM_MLP_FP32 && mode != LBM_MLP_BF16) return mfail(LBM_HIP_EINVAL, "unknown MLP mode");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return mfail(LBM_HIP_EDEVICE, "no HIP device available (liblbm_hip needs an AMD GPU)");
    if (device < 0 || device >= ndev) return mfail(LBM_HIP_EINVAL, "device ordinal out of range");
    auto* M = new lbm_mlp();
    M->nets = nets; M->mode = mode; M->device = device; M->seed = seed;
    auto bail = [&](int code, const char* msg) {
        lbm_mlp_destroy(M);
        return mfail(code, msg);
    };
    if (hipSetDevice(device) != hipSuccess) return bail(LBM_HIP_EDEVICE, "hipSetDevice failed");
    if (hipStreamCreateWithFlags(&M->stream, hipStreamNonBlocking) != hipSuccess)
        return bail(LBM_HIP_EDEVICE, "hipStreamCreate failed");
    {
        int src = LBM_HIP_OK;
        M->noalloc = noalloc_scratch(device, &src);
        if (src) {
            const std::string msg = lbm_hip_last_error();
            lbm_mlp_destroy(M);
            return mfail(src, msg);
        }
    }
    const size_t tot = (size_t)nets * P;
    if (hipMalloc((void**)&M->w, tot * 4) != hipSuccess || hipMalloc((void**)&M->wb, (size_t)nets * PP * 2) != hipSuccess ||
        hipMalloc((void**)&M->wt, (size_t)nets * PT * 2) != hipSuccess ||
        hipMalloc((void**)&M->m, tot * 4) != hipSuccess || hipMalloc((void**)&M->v, tot * 4) != hipSuccess ||
        hipMalloc((void**)&M->grad, tot * 4) != hipSuccess)
        return bail(LBM_HIP_ENOMEM, "MLP allocation failed");
    // Glorot-uniform weights (tanh), zero biases; one generator per net
    std::vector<float> w(tot, 0.f);
    const int fan[4][2] = {{IN, HID}, {HID, HID}, {HID, HID}, {HID, OUT}};
    const int off[4] = {OW1, OW2, OW3, OW4};
    for (int k = 0; k < nets; ++k) {
        std::mt19937_64 gen(seed * 1000003ull + (unsigned long long)k);
        for (int l = 0; l < 4; ++l) {
            const double a = std::sqrt(6.0 / (fan[l][0] + fan[l][1]));
            std::uniform_real_distribution<double> U(-a, a);
            for (int e = 0; e < fan[l][0] * fan[l][1]; ++e) w[(size_t)k * P + off[l] + e] = (float)U(gen);
        }
    }
    // Every step on the handle's stream (it does not synchronise with the
    // null stream), so the clears are ordered before the cast that fills the
    // BF16 copies; their padding (entries no parameter maps to) stays zero.
    if (hipMemcpyAsync(M->w, w.data(), tot * 4, hipMemcpyHostToDevice, M->stream) != hipSuccess ||
        hipMemsetAsync(M->m, 0, tot * 4, M->stream) != hipSuccess ||
        hipMemsetAsync(M->v, 0, tot * 4, M->stream) != hipSuccess ||
        hipMemsetAsync(M->wb, 0, (size_t)nets * PP * 2, M->stream) != hipSuccess ||
        hipMemsetAsync(M->wt, 0, (size_t)nets * PT * 2, M->stream) != hipSuccess)
        return bail(LBM_HIP_EDEVICE, "MLP initialisation failed");
    hipLaunchKernelGGL(mlp_cast_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, M->stream, M->w, M->wb,
                       M->wt, 0, (int)tot);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(M->stream) != hipSuccess)
        return bail(LBM_HIP_EDEVICE, "MLP initialisation failed");
    *out = M;
    return LBM_HIP_OK;
}

void lbm_mlp_destroy(lbm_mlp* M) {
    if (!M) return;
    (void)hipSetDevice(M->device);
    if (M->stream) hipStreamSynchronize(M->stream);
    for (void* q : {(void*)M->w, (void*)M->wb, (void*)M->wt, (void*)M->m, (void*)M->v, (void*)M->grad, (void*)M->part,
                    (void*)M->loss_part, (void*)M->x, (void*)M->t, (void*)M->y, (void*)M->mean, (void*)M->spread})
        if (q) hipFree(q);
    if (M->stream) hipStreamDestroy(M->stream);
    delete M;
}

int lbm_mlp_get_params(lbm_mlp* M, int k, float* params) {
    if (!M || !params) return mfail(LBM_HIP_EINVAL, "null argument");
    if (k < 0 || k >= M->nets) return mfail(LBM_HIP_EINVAL, "net index out of range");
    MLP_TRY(hipMemcpyAsync(params, M->w + (size_t)k * P, P * 4, hipMemcpyDeviceToHost, M->stream));
    MLP_TRY(hipStreamSynchronize(M->stream));
    return LBM_HIP_OK;
}

int lbm_mlp_set_params(lbm_mlp* M, int k, const float* params) {
    if (!M || !params) return mfail(LBM_HIP_EINVAL, "null argument");
    if (k < 0 || k >= M->nets) return mfail(LBM_HIP_EINVAL, "net index out of range");
    MLP_TRY(hipMemcpyAsync(M->w + (size_t)k * P, params, P * 4, hipMemcpyHostToDevice, M->stream));
    hipLaunchKernelGGL(mlp_cast_kernel, dim3((P + 255) / 256), dim3(256), 0, M->stream, M->w, M->wb, M->wt, k * P, P);
    MLP_TRY(hipGetLastError());
    MLP_TRY(hipStreamSynchronize(M->stream));
    return LBM_HIP_OK;
}

int lbm_mlp_forward(lbm_mlp* M, const float* x, int n, float* y) {
    if (!M || !x || !y) return mfail(LBM_HIP_EINVAL, "null argument");
    if (n < 0) return mfail(LBM_HIP_EINVAL, "n must be >= 0");
    if (n == 0) return LBM_HIP_OK;
    int rc;
    if ((rc = upload_xt(M, x, nullptr, n))) return rc;
    if ((rc = ensure(M->y, M->ycap, (size_t)M->nets * n * OUT))) return rc;
    TileArgs a{};
    a.x = M->x; a.n = n; a.y = M->y;
    if ((rc = launch_tiles(M, a, M->nets))) return rc;
    MLP_TRY(hipMemcpyAsync(y, M->y, sizeof(float) * M->nets * n * OUT, hipMemcpyDeviceToHost, M->stream));
    MLP_TRY(hipStreamSynchronize(M->stream));
    return LBM_HIP_OK;
}

int lbm_mlp_predict(lbm_mlp* M, const float* x, int n, float* mean, float* spread) {
    if (!M || !x || !mean) return mfail(LBM_HIP_EINVAL, "null argument");
    if (n < 0) return mfail(LBM_HIP_EINVAL, "n must be >= 0");
    if (n == 0) return LBM_HIP_OK;
    int rc;
    if ((rc = upload_xt(M, x, nullptr, n))) return rc;
    if ((rc = ensure(M->y, M->ycap, (size_t)M->nets * n * OUT))) return rc;
    if ((rc = ensure(M->mean, M->mcap, (size_t)n * OUT))) return rc;
    if ((rc = ensure(M->spread, M->scap, (size_t)n))) return rc;
    TileArgs a{};
    a.x = M->x; a.n = n; a.y = M->y;
    if ((rc = launch_tiles(M, a, M->nets))) return rc;
    if ((rc = launch_stats(M, M->y, n, M->mean, M->spread))) return rc;
    MLP_TRY(hipMemcpyAsync(mean, M->mean, sizeof(float) * n * OUT, hipMemcpyDeviceToHost, M->stream));
    if (spread) MLP_TRY(hipMemcpyAsync(spread, M->spread, sizeof(float) * n, hipMemcpyDeviceToHost, M->stream));
    MLP_TRY(hipStreamSynchronize(M->stream));
    return LBM_HIP_OK;
}

int lbm_mlp_predict_noalloc(lbm_mlp* M, const float* x, int n, float* mean, float* spread) {
    if (!M || !x || !mean) return mfail(LBM_HIP_EINVAL, "null argument");
    if (n < 0) return mfail(LBM_HIP_EINVAL, "n must be >= 0");
    if (n > NA_MAX)
        return mfail(LBM_HIP_EINVAL, "lbm_mlp_predict_noalloc takes at most " + std::to_string(NA_MAX) +
                                         " samples per call (LBM_MLP_NOALLOC_MAX_SAMPLES); got " + std::to_string(n));
    if (n == 0) return LBM_HIP_OK;
    MLP_TRY(hipSetDevice(M->device));
    float* sx = M->noalloc + NA_X;
    float* sy = M->noalloc + NA_Y;
    float* sm = M->noalloc + NA_MEAN;
    float* ss = M->noalloc + NA_SPREAD;
    MLP_TRY(hipMemcpyAsync(sx, x, sizeof(float) * n * IN, hipMemcpyHostToDevice, M->stream));
    TileArgs a{};
    a.x = sx; a.n = n; a.y = sy;
    int rc;
    if ((rc = launch_tiles(M, a, M->nets))) return rc;
    if ((rc = launch_stats(M, sy, n, sm, ss))) return rc;
    MLP_TRY(hipMemcpyAsync(mean, sm, sizeof(float) * n * OUT, hipMemcpyDeviceToHost, M->stream));
    if (spread) MLP_TRY(hipMemcpyAsync(spread, ss, sizeof(float) * n, hipMemcpyDeviceToHost, M->stream));
    MLP_TRY(hipStreamSynchronize(M->stream));
    return LBM_HIP_OK;
}

int lbm_mlp_loss_grad(lbm_mlp* M, int k, const float* x, const float* t, int n, float* loss, float* grad) {
    if (!M || !x || !t) return mfail(LBM_HIP_EINVAL, "null argument");
    if (k < 0 || k >= M->nets) return mfail(LBM_HIP_EINVAL, "net index out of range");
    if (n < 1) return mfail(LBM_HIP_EINVAL, "n must be >= 1");
    int rc;
    if ((rc = upload_xt(M, x, t, n))) return rc;
    const int tiles = units_of(M, n);
    if ((rc = ensure(M->part, M->part_cap, (size_t)tiles * P))) return rc;
    if ((rc = ensure(M->loss_part, M->loss_cap, (size_t)tiles))) return rc;
    TileArgs a{};
    a.x = M->x; a.t = M->t; a.n = n; a.net0 = k;
    a.part = M->part; a.loss_part = M->loss_part;
    a.scale = 2.0f / ((float)n * OUT);
    if ((rc = launch_tiles(M, a, 1))) return rc;
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3((P + 255) / 256, 1), dim3(256), 0, M->stream, M->part, tiles,
                       M->grad + (size_t)k * P);
    MLP_TRY(hipGetLastError());
    std::vector<float> lp((size_t)tiles);
    MLP_TRY(hipMemcpyAsync(lp.data(), M->loss_part, sizeof(float) * tiles, hipMemcpyDeviceToHost, M->stream));
    if (grad) MLP_TRY(hipMemcpyAsync(grad, M->grad + (size_t)k * P, P * 4, hipMemcpyDeviceToHost, M->stream));
    MLP_TRY(hipStreamSynchronize(M->stream));
    double s = 0.0;
    for (float v : lp) s += v;
    if (loss) *loss = (float)(s / ((double)n * OUT));
    return LBM_HIP_OK;
}

int lbm_mlp_train(lbm_mlp* M, const float* x, const float* t, int n, int epochs, int batch, float lr,
                  float* loss_out) {
    if (!M || !x || !t) return mfail(LBM_HIP_EINVAL, "null argument");
    if (n < 1 || epochs < 0 || batch < 1) return mfail(LBM_HIP_EINVAL, "n >= 1, epochs >= 0, batch >= 1 required");
    int rc;
    if ((rc = upload_xt(M, x, t, n))) return rc;
    const int nb = batch < n ? batch : n;
    const int tiles = units_of(M, nb);
    if ((rc = ensure(M->part, M->part_cap, (size_t)M->nets * tiles * P))) return rc;
    if ((rc = ensure(M->loss_part, M->loss_cap, (size_t)M->nets * units_of(M, n)))) return rc;
    int* idx = nullptr;
    if (hipMalloc((void**)&idx, sizeof(int) * (size_t)M->nets * n) != hipSuccess)
        return mfail(LBM_HIP_ENOMEM, "MLP index allocation failed");
    std::vector<int> order((size_t)M->nets * n);
    std::vector<std::mt19937_64> gens;
    for (int k = 0; k < M->nets; ++k) gens.emplace_back(M->seed * 7919ull + 17ull * (unsigned long long)k + 1ull);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    rc = LBM_HIP_OK;
    for (int ep = 0; ep < epochs && rc == LBM_HIP_OK; ++ep) {
        for (int k = 0; k < M->nets; ++k) {
            int* o = order.data() + (size_t)k * n;
            for (int s = 0; s < n; ++s) o[s] = s;
            std::shuffle(o, o + n, gens[(size_t)k]);
        }
        if (hipMemcpyAsync(idx, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice, M->stream) !=
            hipSuccess) { rc = mfail(LBM_HIP_EDEVICE, "MLP index upload failed"); break; }
        // the host order vector is rewritten next epoch: wait for the copy
        if (hipStreamSynchronize(M->stream) != hipSuccess) { rc = mfail(LBM_HIP_EDEVICE, "MLP sync failed"); break; }
        for (int s0 = 0; s0 + nb <= n && rc == LBM_HIP_OK; s0 += nb) {
            // every net in one launch: net k's minibatch is idx[k][s0 .. s0 + nb)
            TileArgs a{};
            a.x = M->x; a.t = M->t; a.n = nb; a.scale = 2.0f / ((float)nb * OUT);
            a.idx = idx + s0;
            a.idx_stride = n;
            a.part = M->part;
            if ((rc = launch_tiles(M, a, M->nets))) break;
            ++M->adam_t;
            const float c1 = 1.f - std::pow(b1, (float)M->adam_t), c2 = 1.f - std::pow(b2, (float)M->adam_t);
            hipLaunchKernelGGL(mlp_reduce_adam_kernel, dim3((P + 255) / 256, M->nets), dim3(256), 0, M->stream, M->w,
                               M->wb, M->wt, M->m, M->v, M->part, tiles, lr, b1, b2, eps, c1, c2);
            if (hipGetLastError() != hipSuccess) rc = mfail(LBM_HIP_EDEVICE, "MLP update launch failed");
        }
    }
    if (rc == LBM_HIP_OK && hipStreamSynchronize(M->stream) != hipSuccess) rc = mfail(LBM_HIP_EDEVICE, "MLP sync failed");
    hipFree(idx);
    if (rc) return rc;
    if (loss_out) {
        const int tl = units_of(M, n);
        TileArgs a{};
        a.x = M->x; a.t = M->t; a.n = n; a.loss_part = M->loss_part; a.scale = 0.f;
        if ((rc = launch_tiles(M, a, M->nets))) return rc;
        std::vector<float> lp((size_t)M->nets * tl);
        MLP_TRY(hipMemcpyAsync(lp.data(), M->loss_part, sizeof(float) * lp.size(), hipMemcpyDeviceToHost, M->stream));
        MLP_TRY(hipStreamSynchronize(M->stream));
        for (int k = 0; k < M->nets; ++k) {
            double s = 0.0;
            for (int j = 0; j < tl; ++j) s += lp[(size_t)k * tl + j];
            loss_out[k] = (float)(s / ((double)n * OUT));
        }
    }
    return LBM_HIP_OK;
}

}  // extern "C"
