// Learned-closure kernels: BF16 GEMM with FP32 accumulation and the element-wise steps of a small
// multilayer perceptron (see cwx_mlp.hip).  All pointers are device pointers; BF16 values are uint16_t.
#pragma once

#include <cstdint>

#include "cwx_common.h"

namespace cwx {

// C (M x N, row stride ldc) = alpha op(A) op(B) + beta C; A(m, k) = A[m sam + k sak], B(k, n) = B[k sbk + n sbn].
void gemm_bf16(int M, int N, int K, const uint16_t* A, long sam, long sak, const uint16_t* B, long sbk, long sbn,
               float alpha, float beta, float* C, long ldc, hipStream_t s);
// One dense layer in one launch: Z = X W^T + bias, H = act(Z) (X: M x K rows of stride ldx, W: N x K rows of
// stride ldw, both BF16 and contiguous along K; K, ldx, ldw multiples of 4).  Z, H (BF16), Hf (FP32) may be null.
void mlp_dense_fwd(int M, int N, int K, const uint16_t* X, long ldx, const uint16_t* W, long ldw, const float* bias,
                   int act, float* Z, uint16_t* H, float* Hf, hipStream_t s);
// Z = Zacc + bias (per column); H = bf16(act(Z)), Hf = act(Z) (act 0: identity, 1: GELU); outputs may be null.
void mlp_act_fwd(int M, int N, int act, const float* Zacc, const float* bias, float* Z, uint16_t* H, float* Hf,
                 hipStream_t s);
// dZ = dH act'(Z) (FP32 and/or BF16), dbias = column sums of dZ.
void mlp_act_bwd(int M, int N, int act, const float* Z, const float* dH, float* dZ, uint16_t* dZb, float* dbias,
                 hipStream_t s);
// loss += mean((Y - T)^2); dY = 2 (Y - T) / (M N).
void mlp_mse(int M, int N, const float* Y, const float* T, float* dY, uint16_t* dYb, float* loss, hipStream_t s);
// One Adam step (bias-corrected, step >= 1) on FP32 weights w with gradient g; wb receives the BF16 copy.
void mlp_adam(long n, float lr, float b1, float b2, float eps, int step, float* w, const float* g, float* m, float* v,
              uint16_t* wb, hipStream_t s);
void to_bf16(long n, const float* x, uint16_t* y, hipStream_t s);

}  // namespace cwx
