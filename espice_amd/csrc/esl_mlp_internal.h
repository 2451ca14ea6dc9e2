// Shared between the network's host code and its kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace esl {

// C = A B (+ bias), M x N, inner K; A(i, k) = A[i * sa_m + k * sa_k],
// B(k, j) = B[k * sb_k + j * sb_n], both BF16 bits. Cf (FP32, before the
// activation) and Cb (BF16, after it) are optional.
struct GemmArgs {
    const uint16_t *A;
    int64_t sa_m, sa_k;
    const uint16_t *B;
    int64_t sb_k, sb_n;
    uint32_t M, N, K, act;   // act: 0 identity, 1 SiLU
    const float *bias;       // [N] or null
    float *Cf;
    int64_t sc_m, sc_n;
    uint16_t *Cb;
    int64_t sd_m, sd_n;
};

struct AdamStep {
    float lr, beta1, beta2, eps, weight_decay, bc1, bc2, pad;
};

}  // namespace esl

extern "C" {
hipError_t esl_mlp_launch_gemm(const esl::GemmArgs *g, hipStream_t s);
hipError_t esl_mlp_launch_inputs(const double *in, const uint32_t *idx, uint32_t rows, uint32_t cols,
                                 const double *mean, const double *scale, uint16_t *X, hipStream_t s);
hipError_t esl_mlp_launch_loss(const float *Y, const float *T, const uint32_t *idx, uint32_t rows, uint32_t cols,
                               float *dYf, uint16_t *dYb, float *row_loss, double *acc, hipStream_t s);
hipError_t esl_mlp_launch_act_back(const float *dH, const float *Z, uint32_t count, float *dZf, uint16_t *dZb,
                                   hipStream_t s);
hipError_t esl_mlp_launch_bias_grad(const float *dZ, uint32_t rows, uint32_t cols, float *db, hipStream_t s);
hipError_t esl_mlp_launch_adam(float *p, const float *grad, float *m, float *v, uint32_t count,
                               const esl::AdamStep *st, const uint8_t *is_weight, uint16_t *pb, hipStream_t s);
hipError_t esl_mlp_launch_to_bf16(const float *p, uint32_t count, uint16_t *pb, hipStream_t s);
hipError_t esl_mlp_launch_outputs(const float *Y, uint32_t rows, uint32_t cols, const double *mean,
                                  const double *scale, double *x, hipStream_t s);
hipError_t esl_mlp_launch_targets(const double *x, uint32_t rows, uint32_t cols, const double *mean,
                                  const double *scale, float *T, hipStream_t s);
}
