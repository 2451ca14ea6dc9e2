// The learned warm start's network on gfx950: BF16 MFMA GEMMs with FP32
// accumulation, SiLU hidden layers, MSE loss, Adam on FP32 master weights.
//
// No counterpart in the reference. The network maps a point's live
// parameters to its operating point; Newton then starts from that guess.
//
// One GEMM kernel serves the forward pass and both backward products: a
// 256-thread workgroup computes a 64x64 tile of C = A B with A(i, k) and
// B(k, j) read through arbitrary strides (so transposed operands need no
// copies), staging 64x32 panels of each through LDS and running
// v_mfma_f32_16x16x32_bf16 (four waves, each a 32x32 quarter as 2x2 MFMA
// tiles).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "esl_mlp_internal.h"

namespace esl {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// FP32 -> BF16, round to nearest even (NaN stays NaN).
__device__ inline uint16_t f2bf(float f)
{
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ inline float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ inline float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ inline float silu_grad(float z)
{
    const float s = 1.0f / (1.0f + expf(-z));
    return s * (1.0f + z * (1.0f - s));
}

constexpr int kTile = 64, kKStep = 32, kLdsRow = kKStep + 8;   // row pitch 80 B

__global__ __launch_bounds__(256) void mlp_gemm(GemmArgs g)
{
    __shared__ __attribute__((aligned(16))) uint16_t As[kTile * kLdsRow];
    __shared__ __attribute__((aligned(16))) uint16_t Bs[kTile * kLdsRow];
    const uint32_t t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint32_t m0 = blockIdx.y * kTile, n0 = blockIdx.x * kTile;
    const uint32_t wr = wave >> 1, wc = wave & 1;
    f32x4 acc[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) acc[a][b] = f32x4{0.f, 0.f, 0.f, 0.f};
    for (uint32_t k0 = 0; k0 < g.K; k0 += kKStep) {
        // stage: 2048 elements of each panel, 8 per thread, k fastest
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const uint32_t e = t * 8 + r, row = e / kKStep, kk = e % kKStep;
            const uint32_t i = m0 + row, j = n0 + row, k = k0 + kk;
            uint16_t av = 0, bv = 0;
            if (i < g.M && k < g.K) av = g.A[(int64_t)i * g.sa_m + (int64_t)k * g.sa_k];
            if (j < g.N && k < g.K) bv = g.B[(int64_t)k * g.sb_k + (int64_t)j * g.sb_n];
            As[row * kLdsRow + kk] = av;
            Bs[row * kLdsRow + kk] = bv;
        }
        __syncthreads();
        const uint32_t kq = 8 * (lane >> 4), r16 = lane & 15;
#pragma unroll
        for (int a = 0; a < 2; a++) {
            const uint4 av = *reinterpret_cast<const uint4 *>(&As[(wr * 32 + a * 16 + r16) * kLdsRow + kq]);
            const bf16x8 af = __builtin_bit_cast(bf16x8, av);
#pragma unroll
            for (int b = 0; b < 2; b++) {
                const uint4 bv = *reinterpret_cast<const uint4 *>(&Bs[(wc * 32 + b * 16 + r16) * kLdsRow + kq]);
                const bf16x8 bf = __builtin_bit_cast(bf16x8, bv);
                acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af, bf, acc[a][b], 0, 0, 0);
            }
        }
        __syncthreads();
    }
    // C tile (16x16): lane holds column lane % 16, rows 4 (lane / 16) + 0..3
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const uint32_t i = m0 + wr * 32 + a * 16 + 4 * (lane >> 4) + r;
                const uint32_t j = n0 + wc * 32 + b * 16 + (lane & 15);
                if (i >= g.M || j >= g.N) continue;
                float v = acc[a][b][r];
                if (g.bias) v += g.bias[j];
                if (g.Cf) g.Cf[(int64_t)i * g.sc_m + (int64_t)j * g.sc_n] = v;
                if (g.Cb) g.Cb[(int64_t)i * g.sd_m + (int64_t)j * g.sd_n] = f2bf(g.act ? silu(v) : v);
            }
}

// X[r][c] = bf16((in[idx[r]][c] - mean[c]) / scale[c]) for the rows of a batch.
__global__ void mlp_inputs(const double *in, const uint32_t *idx, uint32_t rows, uint32_t cols, const double *mean,
                           const double *scale, uint16_t *X)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * cols) return;
    const uint32_t r = e / cols, c = e % cols;
    const uint32_t src = idx ? idx[r] : r;
    X[e] = f2bf((float)((in[(size_t)src * cols + c] - mean[c]) / scale[c]));
}

// Loss and output gradient: d = Y - T (T normalized targets of the batch's
// rows); dY = 2 d / (rows * cols) in FP32 and BF16; row_loss[r] = sum d^2.
__global__ void mlp_loss(const float *Y, const float *T, const uint32_t *idx, uint32_t rows, uint32_t cols,
                         float *dYf, uint16_t *dYb, float *row_loss)
{
    const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= rows) return;
    const uint32_t src = idx ? idx[r] : r;
    const float k = 2.0f / ((float)rows * (float)cols);
    float s = 0.f;
    for (uint32_t c = 0; c < cols; c++) {
        const float d = Y[(size_t)r * cols + c] - T[(size_t)src * cols + c];
        s += d * d;
        const float gd = k * d;
        dYf[(size_t)r * cols + c] = gd;
        dYb[(size_t)r * cols + c] = f2bf(gd);
    }
    row_loss[r] = s;
}

// dZ = dH * silu'(Z), FP32 and BF16.
__global__ void mlp_act_back(const float *dH, const float *Z, uint32_t count, float *dZf, uint16_t *dZb)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    const float v = dH[e] * silu_grad(Z[e]);
    dZf[e] = v;
    dZb[e] = f2bf(v);
}

// db[c] = sum over rows of dZ[r][c]: a block of kBiasCols columns x
// kBiasGroups row groups; each thread sums its group's rows (reads of a
// row coalesced across the block's columns), then the groups add up in LDS
// in a fixed tree (deterministic).
constexpr uint32_t kBiasCols = 64, kBiasGroups = 16;
__global__ __launch_bounds__(kBiasCols * kBiasGroups) void mlp_bias_grad(const float *dZ, uint32_t rows,
                                                                          uint32_t cols, float *db)
{
    __shared__ float part[kBiasGroups][kBiasCols];
    const uint32_t x = threadIdx.x % kBiasCols, g = threadIdx.x / kBiasCols;
    const uint32_t c = blockIdx.x * kBiasCols + x;
    float s = 0.f;
    if (c < cols)
        for (uint32_t r = g; r < rows; r += kBiasGroups) s += dZ[(size_t)r * cols + c];
    part[g][x] = s;
    __syncthreads();
    for (uint32_t h = kBiasGroups / 2; h > 0; h >>= 1) {
        if (g < h) part[g][x] += part[g + h][x];
        __syncthreads();
    }
    if (g == 0 && c < cols) db[c] = part[0][x];
}

// Sum of row losses into loss[0]: 256 threads, strided partial sums, then a
// fixed tree in LDS (deterministic).
__global__ __launch_bounds__(256) void mlp_loss_sum(const float *row_loss, uint32_t rows, uint32_t cols,
                                                    double *acc)
{
    __shared__ double part[256];
    double s = 0.0;
    for (uint32_t r = threadIdx.x; r < rows; r += 256) s += row_loss[r];
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        acc[0] += part[0] / ((double)rows * (double)cols);
        acc[1] += 1.0;
    }
}

// Adam (decoupled weight decay) on the FP32 master parameters; BF16 copies
// of the weights (not the biases) for the GEMMs.
__global__ void mlp_adam(float *p, const float *grad, float *m, float *v, uint32_t count, AdamStep s,
                         const uint8_t *is_weight, uint16_t *pb)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    const float gr = grad[e];
    const float mm = s.beta1 * m[e] + (1.0f - s.beta1) * gr;
    const float vv = s.beta2 * v[e] + (1.0f - s.beta2) * gr * gr;
    m[e] = mm;
    v[e] = vv;
    const float mh = mm / s.bc1, vh = vv / s.bc2;
    float w = p[e];
    if (is_weight[e]) w -= s.lr * s.weight_decay * w;
    w -= s.lr * mh / (sqrtf(vh) + s.eps);
    p[e] = w;
    pb[e] = f2bf(w);
}

// BF16 copies of all parameters (after set_params).
__global__ void mlp_to_bf16(const float *p, uint32_t count, uint16_t *pb)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < count) pb[e] = f2bf(p[e]);
}

// Network outputs -> start vectors: x[r][0] = 0 (ground), x[r][1 + c] =
// Y[r][c] * scale[c] + mean[c].
__global__ void mlp_outputs(const float *Y, uint32_t rows, uint32_t cols, const double *mean, const double *scale,
                            double *x)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * (cols + 1)) return;
    const uint32_t r = e / (cols + 1), c = e % (cols + 1);
    x[e] = c == 0 ? 0.0 : (double)Y[(size_t)r * cols + c - 1] * scale[c - 1] + mean[c - 1];
}

// Normalized training targets: T[r][c] = (x[r][1 + c] - mean[c]) / scale[c].
__global__ void mlp_targets(const double *x, uint32_t rows, uint32_t cols, const double *mean, const double *scale,
                            float *T)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= rows * cols) return;
    const uint32_t r = e / cols, c = e % cols;
    T[e] = (float)((x[(size_t)r * (cols + 1) + 1 + c] - mean[c]) / scale[c]);
}

}  // namespace esl

namespace {
inline uint32_t blocks(uint64_t n, uint32_t b) { return (uint32_t)((n + b - 1) / b); }
}

extern "C" {
hipError_t esl_mlp_launch_gemm(const esl::GemmArgs *g, hipStream_t s)
{
    dim3 grid(blocks(g->N, esl::kTile), blocks(g->M, esl::kTile));
    hipLaunchKernelGGL(esl::mlp_gemm, grid, dim3(256), 0, s, *g);
    return hipGetLastError();
}
hipError_t esl_mlp_launch_inputs(const double *in, const uint32_t *idx, uint32_t rows, uint32_t cols,
                                 const double *mean, const double *scale, uint16_t *X, hipStream_t s)
{
    hipLaunchKernelGGL(esl::mlp_inputs, dim3(blocks((uint64_t)rows * cols, 256)), dim3(256), 0, s, in, idx, rows,
                       cols, mean, scale, X);
    return hipGetLastError();
}
hipError_t esl_mlp_launch_loss(const float *Y, const float *T, const uint32_t *idx, uint32_t rows, uint32_t cols,
                               float *dYf, uint16_t *dYb, float *row_loss, double *acc, hipStream_t s)
{
    hipLaunchKernelGGL(esl::mlp_loss, dim3(blocks(rows, 128)), dim3(128), 0, s, Y, T, idx, rows, cols, dYf, dYb,
                       row_loss);
    hipLaunchKernelGGL(esl::mlp_loss_sum, dim3(1), dim3(256), 0, s, row_loss, rows, cols, acc);
    return hipGetLastError();
}
hipError_t esl_mlp_launch_act_back(const float *dH, const float *Z, uint32_t count, float *dZf, uint16_t *dZb,
                                   hipStream_t s)
{
    hipLaunchKernelGGL(esl::mlp_act_back, dim3(blocks(count, 256)), dim3(256), 0, s, dH, Z, count, dZf, dZb);
    return hipGetLastError();
}
hipError_t esl_mlp_launch_bias_grad(const float *dZ, uint32_t rows, uint32_t cols, float *db, hipStream_t s)
{
    hipLaunchKernelGGL(esl::mlp_bias_grad, dim3(blocks(cols, esl::kBiasCols)), dim3(esl::kBiasCols * esl::kBiasGroups),
                       0, s, dZ, rows, cols, db);
    return hipGetLastError();
}
hipError_t esl_mlp_launch_adam(float *p, const float *grad, float *m, float *v, uint32_t count,
                               const esl::AdamStep *st, const uint8_t *is_weight, uint16_t *pb, hipStream_t s)
{
    hipLaunchKernelGGL(esl::mlp_adam, dim3(blocks(count, 256)), dim3(256), 0, s, p, grad, m, v, count, *st,
                       is_weight, pb);
    return hipGetLastError();
}
hipError_t esl_mlp_launch_to_bf16(const float *p, uint32_t count, uint16_t *pb, hipStream_t s)
{
    hipLaunchKernelGGL(esl::mlp_to_bf16, dim3(blocks(count, 256)), dim3(256), 0, s, p, count, pb);
    return hipGetLastError();
}
hipError_t esl_mlp_launch_outputs(const float *Y, uint32_t rows, uint32_t cols, const double *mean,
                                  const double *scale, double *x, hipStream_t s)
{
    hipLaunchKernelGGL(esl::mlp_outputs, dim3(blocks((uint64_t)rows * (cols + 1), 256)), dim3(256), 0, s, Y, rows,
                       cols, mean, scale, x);
    return hipGetLastError();
}
hipError_t esl_mlp_launch_targets(const double *x, uint32_t rows, uint32_t cols, const double *mean,
                                  const double *scale, float *T, hipStream_t s)
{
    hipLaunchKernelGGL(esl::mlp_targets, dim3(blocks((uint64_t)rows * cols, 256)), dim3(256), 0, s, x, rows, cols,
                       mean, scale, T);
    return hipGetLastError();
}
}
