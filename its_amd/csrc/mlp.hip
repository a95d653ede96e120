// Warm-start MLP for gfx950: BF16 MFMA forward and backward kernels with
// binary32 accumulation, and an Adam step on binary32 master weights.
//
// No reference counterpart: the reference always starts its least-squares
// solve from zero (ThinShells.cpp:615-617).
//
// Network: 16 inputs -> 64 -> 64 -> 64 -> 1, GELU (erf form) after each of the
// three hidden layers.  Parameters, binary32, in this order:
//   W1 [16][64], b1 [64], W2 [64][64], b2 [64], W3 [64][64], b3 [64],
//   W4 [64], b4                                            (9473 values)
// with W[k][n] the weight from input k to output n.  The MFMA operands are
// the weights and the layer inputs rounded to bfloat16; products accumulate
// in binary32, and biases, activations and the last layer stay binary32.
//
// Work decomposition: every wavefront owns tiles of 32 rows.  A layer is
// computed transposed, Z^T (64 x 32) = W^T (64 x K) . H^T (K x 32), with
// v_mfma_f32_32x32x16_bf16, so that each lane ends up holding 32 of the 64
// outputs of one row (row = lane % 32) and applies bias, GELU and GELU' in
// registers.  Layer inputs pass through LDS (row-major, padded rows) to
// become the next layer's operand.  The backward pass uses the same layouts:
// dH^T = W . dZ^T gives the lane's row again, and the weight gradients
// dW = H^T dZ (reduction over the 32 rows of the tile) accumulate in MFMA
// registers across all tiles of the wavefront.  Each wavefront writes its
// gradient once; a second kernel sums the wavefront partials in a fixed
// order, so gradients are deterministic.
#include <algorithm>
#include <cmath>
#include <random>
#include <vector>

#include "internal.hpp"

namespace its {

namespace {
constexpr int IN = ITS_MLP_IN, H = ITS_MLP_HIDDEN, NP = ITS_MLP_NPARAM;
constexpr int oW1 = 0, oB1 = oW1 + IN * H, oW2 = oB1 + H, oB2 = oW2 + H * H, oW3 = oB2 + H, oB3 = oW3 + H * H,
              oW4 = oB3 + H, oB4 = oW4 + H;
static_assert(oB4 + 1 == NP, "parameter layout");

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kRows = 32;      // rows per tile
constexpr int XS = IN + 8;     // padded row strides (bf16 elements); multiples of 8 keep 16-byte alignment
constexpr int HS = H + 8;
constexpr int W1S = IN + 8;

// Inference: 4 wavefronts per workgroup, forward operands only.  Training:
// 2 wavefronts, with the transposed weights for the backward pass, the dZ
// tile, a binary32 tile for row reductions and the pre-activations of the
// first two layers (the third stays in registers).
template <bool TRAIN> constexpr int waves() { return TRAIN ? 2 : 4; }
template <bool TRAIN>
struct Tile {
    __bf16 x[kRows * XS];
    __bf16 h1[kRows * HS];
    __bf16 h2[kRows * HS];
};
template <>
struct Tile<true> {
    __bf16 x[kRows * XS];
    __bf16 h1[kRows * HS];
    __bf16 h2[kRows * HS];
    __bf16 dz[kRows * HS];
    float fs[kRows * H];
    float dy[kRows];
    float z1[kRows * H], z2[kRows * H];
};
template <bool TRAIN>
struct Shared {
    __bf16 w1t[H * W1S];                           // W1^T [n][k]
    __bf16 w2t[H * HS], w3t[H * HS];               // W^T [n][k]: forward operands
    __bf16 w2[TRAIN ? H * HS : 8], w3[TRAIN ? H * HS : 8];   // W [k][n]: backward operands
    float b1[H], b2[H], b3[H], w4[H];
    float b4;
    Tile<TRAIN> wv[waves<TRAIN>()];
};

__device__ inline f32x16 mfma32(bf16x8 a, bf16x8 b, f32x16 c) {
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}
__device__ inline f32x4 mfma16(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ inline bf16x8 ld8(const __bf16* p) { return *reinterpret_cast<const bf16x8*>(p); }
__device__ inline bf16x8 col8(const __bf16* p, int stride) {
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 8; ++j) v[j] = p[j * stride];
    return v;
}
__device__ inline float gelu(float z) { return 0.5f * z * (1.f + erff(z * 0.70710678118654752f)); }
__device__ inline float gelu_d(float z) {
    return 0.5f * (1.f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}
__device__ inline void wave_sync() {
    // LDS written by some lanes is read by other lanes of the same wavefront
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}
// output n of the 32 held by a lane: m = 32-block, q, t: n = 32m + 8q + 4(lane/32) + t
__device__ inline int out_n(int m, int q, int t, int half) { return 32 * m + 8 * q + 4 * half + t; }

template <bool TRAIN>
__device__ void load_weights(Shared<TRAIN>& S, const float* P) {
    for (int i = threadIdx.x; i < IN * H; i += blockDim.x) {
        const int k = i / H, n = i % H;
        S.w1t[n * W1S + k] = (__bf16)P[oW1 + i];
    }
    for (int i = threadIdx.x; i < H * H; i += blockDim.x) {
        const int k = i / H, n = i % H;
        const __bf16 a = (__bf16)P[oW2 + i], b = (__bf16)P[oW3 + i];
        S.w2t[n * HS + k] = a;
        S.w3t[n * HS + k] = b;
        if (TRAIN) {
            S.w2[k * HS + n] = a;
            S.w3[k * HS + n] = b;
        }
    }
    for (int i = threadIdx.x; i < H; i += blockDim.x) {
        S.b1[i] = P[oB1 + i];
        S.b2[i] = P[oB2 + i];
        S.b3[i] = P[oB3 + i];
        S.w4[i] = (float)(__bf16)P[oW4 + i];
    }
    if (threadIdx.x == 0) S.b4 = P[oB4];
    __syncthreads();
}

// layer with K = 64 inputs from LDS tile `hin` (row-major), weights W^T `wt`
__device__ inline void layer64(const __bf16* wt, const __bf16* hin, int lane, f32x16 z[2]) {
    const int r = lane & 31, half = lane >> 5;
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        f32x16 acc = {};
#pragma unroll
        for (int s = 0; s < 4; ++s)
            acc = mfma32(ld8(wt + (32 * m + r) * HS + 16 * s + 8 * half), ld8(hin + r * HS + 16 * s + 8 * half), acc);
        z[m] = acc;
    }
}
// add bias; store GELU(z) to the row-major tile `hout` (bf16); z keeps z + b
__device__ inline void bias_gelu_store(f32x16 z[2], const float* b, __bf16* hout, int lane) {
    const int r = lane & 31, half = lane >> 5;
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int n = out_n(m, q, t, half);
                const float v = z[m][4 * q + t] + b[n];
                z[m][4 * q + t] = v;
                if (hout) hout[r * HS + n] = (__bf16)gelu(v);
            }
}
// keep / reload a lane's 32 pre-activations (binary32 tile, row-major)
__device__ inline void z_store(const f32x16 z[2], float* zs, int lane) {
    const int r = lane & 31, half = lane >> 5;
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int q = 0; q < 4; ++q)
            *reinterpret_cast<f32x4*>(zs + r * H + out_n(m, q, 0, half)) =
                f32x4{z[m][4 * q], z[m][4 * q + 1], z[m][4 * q + 2], z[m][4 * q + 3]};
}
__device__ inline void z_load(f32x16 z[2], const float* zs, int lane) {
    const int r = lane & 31, half = lane >> 5;
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const f32x4 v = *reinterpret_cast<const f32x4*>(zs + r * H + out_n(m, q, 0, half));
#pragma unroll
            for (int t = 0; t < 4; ++t) z[m][4 * q + t] = v[t];
        }
}
// dZ = dH * GELU'(Z) in place on dh; store dZ as bf16 (MFMA operand) and
// binary32 (bias gradient)
__device__ inline void gelu_back_store(f32x16 dh[2], const f32x16 z[2], __bf16* dz, float* fs, int lane) {
    const int r = lane & 31, half = lane >> 5;
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const float v = dh[m][4 * q + t] * gelu_d(z[m][4 * q + t]);
                const int n = out_n(m, q, t, half);
                dh[m][4 * q + t] = v;
                dz[r * HS + n] = (__bf16)v;
                fs[r * H + n] = v;
            }
}
// sum over the 32 rows of column `lane` of a binary32 tile
__device__ inline float col_sum(const float* fs, int lane) {
    float s = 0.f;
#pragma unroll 8
    for (int row = 0; row < kRows; ++row) s += fs[row * H + lane];
    return s;
}
// dW (64 x 64) += Hin^T dZ over the 32 rows of the tile
__device__ inline void grad_w64(const __bf16* hin, const __bf16* dz, int lane, f32x16 gw[2][2]) {
    const int r = lane & 31, half = lane >> 5;
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int row0 = 16 * s + 8 * half;
        bf16x8 a[2], b[2];
#pragma unroll
        for (int m = 0; m < 2; ++m) a[m] = col8(hin + row0 * HS + 32 * m + r, HS);
#pragma unroll
        for (int c = 0; c < 2; ++c) b[c] = col8(dz + row0 * HS + 32 * c + r, HS);
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int c = 0; c < 2; ++c) gw[m][c] = mfma32(a[m], b[c], gw[m][c]);
    }
}
// dH^T (64 x 32) = W (64 x 64, [k][n]) . dZ^T
__device__ inline void back64(const __bf16* w, const __bf16* dz, int lane, f32x16 dh[2]) {
    const int r = lane & 31, half = lane >> 5;
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        f32x16 acc = {};
#pragma unroll
        for (int s = 0; s < 4; ++s)
            acc = mfma32(ld8(w + (32 * m + r) * HS + 16 * s + 8 * half), ld8(dz + r * HS + 16 * s + 8 * half), acc);
        dh[m] = acc;
    }
}
__device__ inline float half_sum(float v) {
    // sum over the 32 lanes of this half-wavefront
#pragma unroll
    for (int o = 1; o < 32; o <<= 1) v += __shfl_xor(v, o);
    return v;
}

// TRAIN = false: y[row] = net(x[row]) for rows < n.
// TRAIN = true: also the squared-error loss against t and its gradient with
// respect to every parameter, scaled by gscale (2/B for a mean over B rows),
// written as one partial per wavefront: part[wave][NP + 1] (last: loss sum).
template <bool TRAIN>
__global__ __launch_bounds__(64 * waves<TRAIN>()) void k_mlp(const float* P, const uint16_t* X, const float* T,
                                                             int64_t n, float gscale, float* Y, float* part) {
    __shared__ Shared<TRAIN> S;
    load_weights<TRAIN>(S, P);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, r = lane & 31, half = lane >> 5;
    Tile<TRAIN>& Wt = S.wv[w];
    const int64_t wid = (int64_t)blockIdx.x * waves<TRAIN>() + w, nw = (int64_t)gridDim.x * waves<TRAIN>();
    const int64_t tiles = (n + kRows - 1) / kRows;

    f32x16 g2[2][2], g3[2][2];
    f32x4 g1[4];
    // per-lane sums for output n = lane: bias gradients and the W4 gradient
    float gb1 = 0.f, gb2 = 0.f, gb3 = 0.f, gw4 = 0.f;
    float gb4 = 0.f, loss = 0.f;
    if (TRAIN) {
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int c = 0; c < 2; ++c) g2[m][c] = g3[m][c] = f32x16{};
#pragma unroll
        for (int c = 0; c < 4; ++c) g1[c] = f32x4{};
    }

    for (int64_t tile = wid; tile < tiles; tile += nw) {
        const int64_t row = tile * kRows + r;
        const bool valid = row < n;
        // input tile: each lane copies 8 of the 16 features of its row
        {
            bf16x8 v = {};
            if (valid) v = *reinterpret_cast<const bf16x8*>(X + row * IN + 8 * half);
            *reinterpret_cast<bf16x8*>(Wt.x + r * XS + 8 * half) = v;
        }
        wave_sync();
        // layer 1 (K = 16)
        f32x16 z1[2], z2[2], z3[2];
#pragma unroll
        for (int m = 0; m < 2; ++m)
            z1[m] = mfma32(ld8(S.w1t + (32 * m + r) * W1S + 8 * half), ld8(Wt.x + r * XS + 8 * half), f32x16{});
        bias_gelu_store(z1, S.b1, Wt.h1, lane);
        if constexpr (TRAIN) z_store(z1, Wt.z1, lane);
        wave_sync();
        layer64(S.w2t, Wt.h1, lane, z2);
        bias_gelu_store(z2, S.b2, Wt.h2, lane);
        if constexpr (TRAIN) z_store(z2, Wt.z2, lane);
        wave_sync();
        layer64(S.w3t, Wt.h2, lane, z3);
        bias_gelu_store(z3, S.b3, nullptr, lane);
        // output layer in binary32; training keeps h3 in the row-reduction tile
        float yp = 0.f;
#pragma unroll
        for (int m = 0; m < 2; ++m)
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int n = out_n(m, q, t, half);
                    const float h = gelu(z3[m][4 * q + t]);
                    yp += h * S.w4[n];
                    if constexpr (TRAIN) Wt.fs[r * H + n] = h;
                }
        const float y = yp + __shfl_xor(yp, 32) + S.b4;
        if (!TRAIN) {
            if (valid && half == 0) Y[row] = y;
            continue;
        }
        if constexpr (TRAIN) {
            const float e = valid ? y - T[row] : 0.f;
            const float dy = gscale * e;
            if (half == 0) {
                loss += e * e;
                gb4 += dy;
                Wt.dy[r] = dy;
            }
            wave_sync();
            {
                // dW4[n] = sum_rows h3[row][n] dy[row]
                float s4 = 0.f;
#pragma unroll 8
                for (int rr = 0; rr < kRows; ++rr) s4 += Wt.fs[rr * H + lane] * Wt.dy[rr];
                gw4 += s4;
            }
            f32x16 dh[2];
#pragma unroll
            for (int m = 0; m < 2; ++m)
#pragma unroll
                for (int q = 0; q < 4; ++q)
#pragma unroll
                    for (int t = 0; t < 4; ++t) dh[m][4 * q + t] = dy * S.w4[out_n(m, q, t, half)];
            wave_sync();
            // layer 3
            gelu_back_store(dh, z3, Wt.dz, Wt.fs, lane);
            wave_sync();
            gb3 += col_sum(Wt.fs, lane);
            grad_w64(Wt.h2, Wt.dz, lane, g3);
            back64(S.w3, Wt.dz, lane, dh);
            wave_sync();
            // layer 2
            z_load(z2, Wt.z2, lane);
            gelu_back_store(dh, z2, Wt.dz, Wt.fs, lane);
            wave_sync();
            gb2 += col_sum(Wt.fs, lane);
            grad_w64(Wt.h1, Wt.dz, lane, g2);
            back64(S.w2, Wt.dz, lane, dh);
            wave_sync();
            // layer 1
            z_load(z1, Wt.z1, lane);
            gelu_back_store(dh, z1, Wt.dz, Wt.fs, lane);
            wave_sync();
            gb1 += col_sum(Wt.fs, lane);
            {
                // dW1 (16 x 64) = X^T dZ1 with 16x16x32 tiles: A[i][k] = X[row k][feature i]
                const int c16 = lane & 15, g = lane >> 4;
                const bf16x8 a = col8(Wt.x + (8 * g) * XS + c16, XS);
#pragma unroll
                for (int c = 0; c < 4; ++c) g1[c] = mfma16(a, col8(Wt.dz + (8 * g) * HS + 16 * c + c16, HS), g1[c]);
            }
            wave_sync();
        }
    }
    if constexpr (!TRAIN) return;
    float* out = part + wid * (NP + 1);
    {
        const int c16 = lane & 15, g = lane >> 4;
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int i = 0; i < 4; ++i) out[oW1 + (4 * g + i) * H + 16 * c + c16] = g1[c][i];
    }
#pragma unroll
    for (int m = 0; m < 2; ++m)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int i = 0; i < 16; ++i) {
                const int k = 32 * m + 8 * (i / 4) + 4 * half + (i % 4), nn = 32 * c + r;
                out[oW2 + k * H + nn] = g2[m][c][i];
                out[oW3 + k * H + nn] = g3[m][c][i];
            }
    out[oB1 + lane] = gb1;
    out[oB2 + lane] = gb2;
    out[oB3 + lane] = gb3;
    out[oW4 + lane] = gw4;
    const float sb4 = half_sum(gb4), sl = half_sum(loss);
    if (lane == 0) {
        out[oB4] = sb4;
        out[NP] = sl;
    }
}

// g[p] = sum over wavefronts of part[w][p], fixed order
__global__ void k_sum_parts(const float* part, int64_t nw, float* g) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p > NP) return;
    float s = 0.f;
    for (int64_t w = 0; w < nw; ++w) s += part[w * (NP + 1) + p];
    g[p] = s;
}

// Adam (Kingma and Ba 2015) with bias correction
__global__ void k_adam(float* P, float* M, float* V, const float* g, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NP) return;
    const float gi = g[p];
    const float m = b1 * M[p] + (1.f - b1) * gi;
    const float v = b2 * V[p] + (1.f - b2) * gi * gi;
    M[p] = m;
    V[p] = v;
    P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

unsigned mlp_grid(int64_t n, int wpb, int64_t cap) {
    const int64_t tiles = (n + kRows - 1) / kRows;
    const int64_t want = std::min<int64_t>((tiles + wpb - 1) / wpb, cap);
    return (unsigned)std::max<int64_t>(want, 1);
}
}  // namespace

struct Mlp {
    its_ctx ctx = nullptr;
    DBuf<float> P, M, V, g, part;
    int64_t step = 0;
};
}  // namespace its

struct its_mlp_s {
    its::Mlp m;
};

namespace its {

// Glorot-uniform weights, zero biases, from a 64-bit seed (host RNG)
void mlp_init(Mlp& m, its_ctx ctx, uint64_t seed) {
    m.ctx = ctx;
    m.P.alloc(NP);
    m.M.alloc(NP);
    m.V.alloc(NP);
    m.g.alloc(NP + 1);
    std::vector<float> h(NP, 0.f);
    std::mt19937_64 rng(seed);
    auto fill = [&](int off, int fan_in, int fan_out, int count) {
        const double a = std::sqrt(6.0 / (fan_in + fan_out));
        std::uniform_real_distribution<double> u(-a, a);
        for (int i = 0; i < count; ++i) h[off + i] = (float)u(rng);
    };
    fill(oW1, IN, H, IN * H);
    fill(oW2, H, H, H * H);
    fill(oW3, H, H, H * H);
    fill(oW4, H, 1, H);
    hipStream_t st = ctx->stream;
    ITS_HIP(hipMemcpyAsync(m.P.p, h.data(), NP * sizeof(float), hipMemcpyHostToDevice, st));
    ITS_HIP(hipMemsetAsync(m.M.p, 0, NP * sizeof(float), st));
    ITS_HIP(hipMemsetAsync(m.V.p, 0, NP * sizeof(float), st));
    ITS_HIP(hipStreamSynchronize(st));
    m.step = 0;
}

void mlp_forward(Mlp& m, const uint16_t* d_x, int64_t n, float* d_y, hipStream_t st) {
    if (n <= 0) return;
    const unsigned grid = mlp_grid(n, waves<false>(), (int64_t)m.ctx->prop.multiProcessorCount * 4);
    hipLaunchKernelGGL(k_mlp<false>, dim3(grid), dim3(64 * waves<false>()), 0, st, (const float*)m.P.p, d_x, (const float*)nullptr, n,
                       0.f, d_y, (float*)nullptr);
    ITS_LAUNCH_CHECK();
}

// gradient of mean((y - t)^2) into m.g (NP values, then the loss sum)
void mlp_grad(Mlp& m, const uint16_t* d_x, const float* d_t, int64_t n, hipStream_t st) {
    ITS_REQUIRE(n > 0, ITS_E_INVALID, "empty batch");
    const unsigned grid = mlp_grid(n, waves<true>(), 64);
    const int64_t nw = (int64_t)grid * waves<true>();
    m.part.ensure(nw * (NP + 1));
    hipLaunchKernelGGL(k_mlp<true>, dim3(grid), dim3(64 * waves<true>()), 0, st, (const float*)m.P.p, d_x, d_t, n, 2.f / (float)n,
                       (float*)nullptr, m.part.p);
    ITS_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_sum_parts, dim3((NP + 1 + 255) / 256), dim3(256), 0, st, (const float*)m.part.p, nw, m.g.p);
    ITS_LAUNCH_CHECK();
}

void mlp_adam(Mlp& m, float lr, hipStream_t st) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    ++m.step;
    const float c1 = 1.f - std::pow(b1, (float)m.step), c2 = 1.f - std::pow(b2, (float)m.step);
    hipLaunchKernelGGL(k_adam, dim3((NP + 255) / 256), dim3(256), 0, st, m.P.p, m.M.p, m.V.p, (const float*)m.g.p,
                       lr, b1, b2, eps, c1, c2);
    ITS_LAUNCH_CHECK();
}

Mlp& mlp_of(its_mlp net) { return net->m; }
its_ctx mlp_ctx(its_mlp net) { return net->m.ctx; }
void mlp_set(its_mlp net, const float* h, hipStream_t st) {
    ITS_HIP(hipMemcpyAsync(net->m.P.p, h, NP * sizeof(float), hipMemcpyHostToDevice, st));
    ITS_HIP(hipMemsetAsync(net->m.M.p, 0, NP * sizeof(float), st));
    ITS_HIP(hipMemsetAsync(net->m.V.p, 0, NP * sizeof(float), st));
    ITS_HIP(hipStreamSynchronize(st));
    net->m.step = 0;
}
void mlp_get(its_mlp net, float* h, hipStream_t st) {
    ITS_HIP(hipMemcpyAsync(h, net->m.P.p, NP * sizeof(float), hipMemcpyDeviceToHost, st));
    ITS_HIP(hipStreamSynchronize(st));
}
// gradient (NP values) and mean loss of the last mlp_grad
void mlp_read_grad(its_mlp net, float* grad, double* loss, int64_t n, hipStream_t st) {
    std::vector<float> h(NP + 1);
    ITS_HIP(hipMemcpyAsync(h.data(), net->m.g.p, (NP + 1) * sizeof(float), hipMemcpyDeviceToHost, st));
    ITS_HIP(hipStreamSynchronize(st));
    if (grad) std::copy(h.begin(), h.begin() + NP, grad);
    if (loss) *loss = (double)h[NP] / (double)n;
}
namespace {
__global__ void k_scale_to_f64(int64_t n, const float* y, const float* s, double* w) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) w[i] = (double)y[i] * (double)s[i];
}
}  // namespace
void scale_to_f64(int64_t n, const float* y, const float* s, double* w, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_scale_to_f64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n, y, s, w);
    ITS_LAUNCH_CHECK();
}
its_mlp mlp_new() { return new its_mlp_s(); }
void mlp_delete(its_mlp net) { delete net; }

}  // namespace its
