// mlp.hip -- the learned velocity surrogate (include/lbm_mlp.h): an ensemble
// of 5 -> 128 -> 128 -> 128 -> 6 tanh MLPs, forward, backward and Adam in
// hand-written HIP.
//
// One wavefront (one block of 64 threads) takes a tile of 16 samples of one
// net through the whole network: activations stay in LDS, every layer is a
// product of a 16-row activation tile with a weight matrix, and the backward
// pass forms the weight gradients of its 16 samples (another such product,
// contracted over the samples) into a per-tile partial.  A second kernel adds
// the partials of all tiles in tile order (deterministic), a third applies
// Adam to the FP32 master weights and refreshes their BF16 copy.
//
// Products: in the BF16 mode each 16 x 16 output block is accumulated by
// v_mfma_f32_16x16x16_bf16 (A 16 x 16 and B 16 x 16 in BF16, FP32
// accumulator); in the FP32 mode the same blocks, with the same lane layout,
// are accumulated with FP32 multiply-adds.
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

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ uint16_t to_bf16(float f) {
    const uint32_t u = __float_as_uint(f);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);  // round to nearest even (finite values)
}
__device__ __forceinline__ float from_bf16(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// C (M x N, in 16 x 16 blocks) = A (M x K) * B (K x N); M, N, K multiples of
// 16 (the fetchers return 0 outside the real sizes).  Lane l of the block
// (mt, nt) receives C[mt*16 + 4*(l/16) + j][nt*16 + l%16], j = 0..3, and
// hands each to `emit`.
template <bool BF16, class FA, class FB, class FE>
__device__ __forceinline__ void wave_gemm(int M, int N, int K, FA&& a, FB&& b, FE&& emit) {
    const int lane = threadIdx.x & 63;
    const int r16 = lane & 15, q = lane >> 4;
    for (int mt = 0; mt < M / 16; ++mt)
        for (int nt = 0; nt < N / 16; ++nt) {
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            for (int kt = 0; kt < K / 16; ++kt) {
                if constexpr (BF16) {
                    bf16x4 av, bv;
                    for (int j = 0; j < 4; ++j) {
                        av[j] = (short)to_bf16(a(mt * 16 + r16, kt * 16 + 4 * q + j));
                        bv[j] = (short)to_bf16(b(kt * 16 + 4 * q + j, nt * 16 + r16));
                    }
                    acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(av, bv, acc, 0, 0, 0);
                } else {
                    for (int k = kt * 16; k < kt * 16 + 16; ++k) {
                        const float bk = b(k, nt * 16 + r16);
                        for (int j = 0; j < 4; ++j) acc[j] = fmaf(a(mt * 16 + 4 * q + j, k), bk, acc[j]);
                    }
                }
            }
            for (int j = 0; j < 4; ++j) emit(mt * 16 + 4 * q + j, nt * 16 + r16, acc[j]);
        }
}

struct TileArgs {
    const float* x;      // [nsrc][5]
    const float* t;      // [nsrc][6] targets, or null (forward only)
    const int* idx;      // sample of row s of this pass is idx[s] (single-net launches), or s if null
    int n;               // samples of this pass
    int net0;            // first net of the launch (grid.y counts nets from here)
    float* y;            // [nets][n][6] outputs, or null
    const float* wf;     // [nets][P] FP32 parameters
    const uint16_t* wb;  // [nets][P] BF16 copy
    float* part;         // [nets][tiles][P] gradient partials, or null
    float* loss_part;    // [nets][tiles] loss partials, or null
    float scale;         // dL/dy = scale * (y - t)
    int tiles;
};

template <bool BF16>
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
    const uint16_t* wb = A.wb + (size_t)net * P;
    auto W = [&](int off) -> float { return BF16 ? from_bf16(wb[off]) : wf[off]; };
    auto sample = [&](int r) -> int {
        const int s = row0 + r;
        return A.idx ? A.idx[s] : s;
    };

    // ---- inputs ----
    for (int e = lane; e < TILE * 16; e += 64) {
        const int r = e >> 4, c = e & 15;
        xin[r][c] = (r < nrows && c < IN) ? A.x[(size_t)sample(r) * IN + c] : 0.f;
    }
    __syncthreads();

    // ---- forward ----
    auto dense_tanh = [&](auto&& act_in, int kin, int wo, int bo, float (*out)[HS]) {
        wave_gemm<BF16>(TILE, HID, kin, act_in, [&](int k, int o) -> float { return k < (kin == 16 ? IN : HID) ? W(wo + o * (kin == 16 ? IN : HID) + k) : 0.f; },
                        [&](int r, int o, float v) { out[r][o] = tanhf(v + wf[bo + o]); });
        __syncthreads();
    };
    dense_tanh([&](int r, int k) -> float { return xin[r][k]; }, 16, OW1, OB1, h[0]);
    dense_tanh([&](int r, int k) -> float { return h[0][r][k]; }, HID, OW2, OB2, h[1]);
    dense_tanh([&](int r, int k) -> float { return h[1][r][k]; }, HID, OW3, OB3, h[2]);
    wave_gemm<BF16>(TILE, 16, HID, [&](int r, int k) -> float { return h[2][r][k]; },
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
            d = yo[r][c] - A.t[(size_t)sample(r) * OUT + c];
            l += d * d;
        }
        dy[r][c] = A.scale * d;
    }
    for (int off = 32; off > 0; off >>= 1) l += __shfl_xor(l, off, 64);
    if (A.loss_part && lane == 0) A.loss_part[(size_t)blockIdx.y * A.tiles + tile] = l;
    __syncthreads();
    if (!A.part) return;
    float* g = A.part + ((size_t)blockIdx.y * A.tiles + tile) * P;

    // ---- backward ----
    // layer 4: dW4 = dy^T h3, db4 = sum dy, dh3 = dy W4 -> dz[0]
    wave_gemm<BF16>(16, HID, TILE, [&](int o, int r) -> float { return dy[r][o]; },
                    [&](int r, int i) -> float { return h[2][r][i]; },
                    [&](int o, int i, float v) { if (o < OUT) g[OW4 + o * HID + i] = v; });
    if (lane < OUT) {
        float s = 0.f;
        for (int r = 0; r < TILE; ++r) s += dy[r][lane];
        g[OB4 + lane] = s;
    }
    wave_gemm<BF16>(TILE, HID, 16, [&](int r, int o) -> float { return dy[r][o]; },
                    [&](int o, int i) -> float { return o < OUT ? W(OW4 + o * HID + i) : 0.f; },
                    [&](int r, int i, float v) { dz[0][r][i] = v * (1.f - h[2][r][i] * h[2][r][i]); });
    __syncthreads();
    // layers 3 and 2: dW = dz^T h_prev, db = sum dz, dh_prev = dz W -> through tanh
    auto hidden_back = [&](int wo, int bo, int hl, int cur) {
        float(*z)[HS] = dz[cur];
        float(*hp)[HS] = h[hl - 1];
        wave_gemm<BF16>(HID, HID, TILE, [&](int o, int r) -> float { return z[r][o]; },
                        [&](int r, int i) -> float { return hp[r][i]; },
                        [&](int o, int i, float v) { g[wo + o * HID + i] = v; });
        for (int o = lane; o < HID; o += 64) {
            float s = 0.f;
            for (int r = 0; r < TILE; ++r) s += z[r][o];
            g[bo + o] = s;
        }
        float(*zn)[HS] = dz[cur ^ 1];
        wave_gemm<BF16>(TILE, HID, HID, [&](int r, int o) -> float { return z[r][o]; },
                        [&](int o, int i) -> float { return W(wo + o * HID + i); },
                        [&](int r, int i, float v) { zn[r][i] = v * (1.f - hp[r][i] * hp[r][i]); });
        __syncthreads();
    };
    hidden_back(OW3, OB3, 2, 0);  // dz[0] (layer 3) -> dz[1] (layer 2)
    hidden_back(OW2, OB2, 1, 1);  // dz[1] (layer 2) -> dz[0] (layer 1)
    // layer 1: dW1 = dz1^T x, db1
    wave_gemm<BF16>(HID, 16, TILE, [&](int o, int r) -> float { return dz[0][r][o]; },
                    [&](int r, int i) -> float { return xin[r][i]; },
                    [&](int o, int i, float v) { if (i < IN) g[OW1 + o * IN + i] = v; });
    for (int o = lane; o < HID; o += 64) {
        float s = 0.f;
        for (int r = 0; r < TILE; ++r) s += dz[0][r][o];
        g[OB1 + o] = s;
    }
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

__global__ __launch_bounds__(256) void mlp_adam_kernel(float* w, uint16_t* wb, float* m, float* v, const float* grad,
                                                       float lr, float b1, float b2, float eps, float c1, float c2,
                                                       int total) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= total) return;
    const float gr = grad[e];
    const float mm = b1 * m[e] + (1.f - b1) * gr;
    const float vv = b2 * v[e] + (1.f - b2) * gr * gr;
    m[e] = mm;
    v[e] = vv;
    const float nw = w[e] - lr * (mm / c1) / (sqrtf(vv / c2) + eps);
    w[e] = nw;
    wb[e] = to_bf16(nw);
}

__global__ __launch_bounds__(256) void mlp_cast_kernel(const float* w, uint16_t* wb, int total) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e < total) wb[e] = to_bf16(w[e]);
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
    uint16_t* wb = nullptr;  // [nets][P]
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

int launch_tiles(lbm_mlp* M, TileArgs a, int nets) {
    a.wf = M->w;
    a.wb = M->wb;
    a.tiles = (a.n + TILE - 1) / TILE;
    if (a.tiles == 0) return LBM_HIP_OK;
    if (M->mode == LBM_MLP_BF16)
        hipLaunchKernelGGL(mlp_tile_kernel<true>, dim3(a.tiles, nets), dim3(64), 0, M->stream, a);
    else
        hipLaunchKernelGGL(mlp_tile_kernel<false>, dim3(a.tiles, nets), dim3(64), 0, M->stream, a);
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
    if (hipMalloc((void**)&M->w, tot * 4) != hipSuccess || hipMalloc((void**)&M->wb, tot * 2) != hipSuccess ||
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
    if (hipMemcpy(M->w, w.data(), tot * 4, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(M->m, 0, tot * 4) != hipSuccess || hipMemset(M->v, 0, tot * 4) != hipSuccess)
        return bail(LBM_HIP_EDEVICE, "MLP initialisation failed");
    hipLaunchKernelGGL(mlp_cast_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, M->stream, M->w, M->wb,
                       (int)tot);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(M->stream) != hipSuccess)
        return bail(LBM_HIP_EDEVICE, "MLP initialisation failed");
    *out = M;
    return LBM_HIP_OK;
}

void lbm_mlp_destroy(lbm_mlp* M) {
    if (!M) return;
    hipSetDevice(M->device);
    if (M->stream) hipStreamSynchronize(M->stream);
    for (void* q : {(void*)M->w, (void*)M->wb, (void*)M->m, (void*)M->v, (void*)M->grad, (void*)M->part,
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
    hipLaunchKernelGGL(mlp_cast_kernel, dim3((P + 255) / 256), dim3(256), 0, M->stream, M->w + (size_t)k * P,
                       M->wb + (size_t)k * P, P);
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
    const int tiles = (n + TILE - 1) / TILE;
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
    const int tiles = (nb + TILE - 1) / TILE;
    if ((rc = ensure(M->part, M->part_cap, (size_t)M->nets * tiles * P))) return rc;
    if ((rc = ensure(M->loss_part, M->loss_cap, (size_t)M->nets * ((n + TILE - 1) / TILE)))) return rc;
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
            TileArgs a{};
            a.x = M->x; a.t = M->t; a.n = nb; a.scale = 2.0f / ((float)nb * OUT);
            // one launch per net: its minibatch is idx[k][s0 .. s0 + nb)
            for (int k = 0; k < M->nets && rc == LBM_HIP_OK; ++k) {
                a.net0 = k;
                a.part = M->part + (size_t)k * tiles * P;
                a.idx = idx + (size_t)k * n + s0;
                rc = launch_tiles(M, a, 1);
                if (rc) break;
                hipLaunchKernelGGL(mlp_reduce_kernel, dim3((P + 255) / 256, 1), dim3(256), 0, M->stream, a.part, tiles,
                                   M->grad + (size_t)k * P);
                if (hipGetLastError() != hipSuccess) rc = mfail(LBM_HIP_EDEVICE, "MLP reduce launch failed");
            }
            if (rc) break;
            ++M->adam_t;
            const float c1 = 1.f - std::pow(b1, (float)M->adam_t), c2 = 1.f - std::pow(b2, (float)M->adam_t);
            const int tot = M->nets * P;
            hipLaunchKernelGGL(mlp_adam_kernel, dim3((tot + 255) / 256), dim3(256), 0, M->stream, M->w, M->wb, M->m,
                               M->v, M->grad, lr, b1, b2, eps, c1, c2, tot);
            if (hipGetLastError() != hipSuccess) rc = mfail(LBM_HIP_EDEVICE, "MLP Adam launch failed");
        }
    }
    if (rc == LBM_HIP_OK && hipStreamSynchronize(M->stream) != hipSuccess) rc = mfail(LBM_HIP_EDEVICE, "MLP sync failed");
    hipFree(idx);
    if (rc) return rc;
    if (loss_out) {
        const int tl = (n + TILE - 1) / TILE;
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
