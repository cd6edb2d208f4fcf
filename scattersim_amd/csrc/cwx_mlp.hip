// Kernels of the learned closure: a small multilayer perceptron trained and run in BF16 with FP32
// accumulation on the matrix cores.
//
// GEMM: C (FP32) = alpha op(A) op(B) + beta C with A and B in BF16 under arbitrary element strides, so one
// kernel serves the forward product (X W^T), the data gradient (dY W) and the weight gradient (dY^T X).
// A 256-thread workgroup computes a 64 x 64 tile of C: 64 x 16 and 16 x 64 slices of A and B are staged
// through LDS per step of k, and each of the four wavefronts accumulates a 32 x 32 block as 2 x 2
// v_mfma_f32_16x16x16_bf16 products.  For that instruction lane l supplies A[l % 16][4(l/16) + 0..3] and
// B[4(l/16) + 0..3][l % 16], and holds D[4(l/16) + q][l % 16] in register q.
// A dense layer (the forward product, both operands contiguous along k) whose widths are multiples of 8
// runs k_dense_bf16: LDS-staged tiles on the double-depth instruction, bias and activation in the
// epilogue.  At widths that are multiples of 4 only, k_gemm_bf16_kc reads the fragments straight from
// memory instead.
// The element-wise kernels: bias + GELU (forward and backward, with the bias gradient), the mean-squared
// error and its gradient, Adam on FP32 master weights with BF16 copies, and FP32 -> BF16 rounding (to
// nearest even).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "cwx_common.h"
#include "cwx_mlp.h"

namespace cwx {
namespace {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ inline uint16_t to_bf16(float f) {
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float from_bf16(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

constexpr int GT = 64;   // C tile
constexpr int GK = 16;   // k step

__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const uint16_t* __restrict__ A, long sam,
                                                   long sak, const uint16_t* __restrict__ B, long sbk, long sbn,
                                                   float alpha, float beta, float* __restrict__ C, long ldc) {
  __shared__ uint16_t As[GT][GK + 2];
  __shared__ uint16_t Bs[GK][GT + 2];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
  const int wm = (w >> 1) * 32, wn = (w & 1) * 32;
  f32x4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int li = lane & 15, lk = (lane >> 4) * 4;
  for (int k0 = 0; k0 < K; k0 += GK) {
    for (int e = tid; e < GT * GK; e += 256) {
      const int r = e / GK, kk = e % GK, gm = m0 + r, gk = k0 + kk;
      As[r][kk] = (gm < M && gk < K) ? A[(long)gm * sam + (long)gk * sak] : (uint16_t)0;
    }
    for (int e = tid; e < GT * GK; e += 256) {
      const int kk = e / GT, c = e % GT, gk = k0 + kk, gn = n0 + c;
      Bs[kk][c] = (gk < K && gn < N) ? B[(long)gk * sbk + (long)gn * sbn] : (uint16_t)0;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const int ar = wm + 16 * i + li;
      const bf16x4 a = {(short)As[ar][lk], (short)As[ar][lk + 1], (short)As[ar][lk + 2], (short)As[ar][lk + 3]};
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int bc = wn + 16 * j + li;
        const bf16x4 b = {(short)Bs[lk][bc], (short)Bs[lk + 1][bc], (short)Bs[lk + 2][bc], (short)Bs[lk + 3][bc]};
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc[i][j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = m0 + wm + 16 * i + lk + q, col = n0 + wn + 16 * j + li;
        if (row < M && col < N) {
          float* c = C + (long)row * ldc + col;
          const float v = alpha * acc[i][j][q];
          *c = beta == 0.f ? v : v + beta * *c;
        }
      }
}

// Both operands contiguous along k (A(m, k) = A[m lda + k], B(k, n) = B[n ldb + k]; k, lda, ldb multiples of 4,
// 8-byte aligned bases): every lane reads its MFMA fragments straight from memory as 8-byte loads, no LDS.
// Each wavefront accumulates a 64 x 64 block as 4 x 4 products per step of k (8 loads per 16 MFMAs); four
// wavefronts make a 128 x 128 tile, whose rows and columns they share through the L1 cache.
constexpr int KC_W = 64;    // per-wavefront block
constexpr int KC_T = 128;   // workgroup tile
__device__ inline float gelu(float z, float& dg);

// Epilogue of the fast GEMM: with bias or an activation it writes Z = acc + bias (FP32, may be null),
// H = act(Z) as BF16 and/or FP32 (row stride N) instead of alpha acc + beta C.
struct DenseOut {
  const float* bias = nullptr;
  int act = -1;                 // -1: plain GEMM epilogue; 0: identity; 1: GELU
  float* Z = nullptr;
  uint16_t* H = nullptr;
  float* Hf = nullptr;
};

__global__ __launch_bounds__(256) void k_gemm_bf16_kc(int M, int N, int K, const uint16_t* __restrict__ A, long lda,
                                                      const uint16_t* __restrict__ B, long ldb, float alpha,
                                                      float beta, float* __restrict__ C, long ldc, DenseOut d) {
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int m0 = blockIdx.y * KC_T + (w >> 1) * KC_W, n0 = blockIdx.x * KC_T + (w & 1) * KC_W;
  const int li = lane & 15, lk = (lane >> 4) * 4;
  f32x4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const bf16x4 zero = {0, 0, 0, 0};
  const uint16_t* arow[4];
  const uint16_t* bcol[4];
  bool aok[4], bok[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int r = m0 + 16 * i + li, c = n0 + 16 * i + li;
    aok[i] = r < M;
    bok[i] = c < N;
    arow[i] = A + (long)(aok[i] ? r : 0) * lda + lk;
    bcol[i] = B + (long)(bok[i] ? c : 0) * ldb + lk;
  }
  for (int k = 0; k < K; k += GK) {
    const bool kin = k + lk < K;
    bf16x4 a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a[i] = (aok[i] && kin) ? *reinterpret_cast<const bf16x4*>(arow[i] + k) : zero;
      b[i] = (bok[i] && kin) ? *reinterpret_cast<const bf16x4*>(bcol[i] + k) : zero;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
  }
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int row = m0 + 16 * i + lk + q, col = n0 + 16 * j + li;
        if (row < M && col < N) {
          if (d.act < 0) {
            float* c = C + (long)row * ldc + col;
            const float v = alpha * acc[i][j][q];
            *c = beta == 0.f ? v : v + beta * *c;
          } else {
            const long e = (long)row * N + col;
            const float z = alpha * acc[i][j][q] + (d.bias ? d.bias[col] : 0.f);
            float dg;
            const float h = d.act ? gelu(z, dg) : z;
            if (d.Z) d.Z[e] = z;
            if (d.H) d.H[e] = to_bf16(h);
            if (d.Hf) d.Hf[e] = h;
          }
        }
      }
}

// Dense layer on the gfx950 double-depth BF16 instruction (v_mfma_f32_16x16x32_bf16: lane l supplies
// A[l % 16][8(l/16) + 0..7] and B[8(l/16) + 0..7][l % 16], and holds D[4(l/16) + q][l % 16] in register q).
// A 256-thread workgroup computes a BM x BN tile of X W^T (64 x 64 as launched); each wavefront a
// (BM/2) x (BN/2) block per 32-deep step of k.  The slices of X and W are loaded as 16-byte vectors into registers one step
// ahead and stored into the other half of a double-buffered LDS, so one barrier per step separates the
// stores from the reads.  LDS rows are padded to 40 values (80 bytes): the 16-byte fragment reads of 16
// consecutive rows then fall into distinct bank groups.  The epilogue stages each wavefront's block of
// Z = acc + bias in LDS, 32 rows at a time, and writes whole rows with vector stores (Z, act(Z) in FP32
// and in BF16).
// Needs K, ldx, ldw multiples of 8 and 16-byte aligned X, W.
constexpr int DL_K = 32, DL_LD = DL_K + 8;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <int BM, int BN>
__global__ __launch_bounds__(256) void k_dense_bf16(int M, int N, int K, const uint16_t* __restrict__ X, long ldx,
                                                    const uint16_t* __restrict__ W, long ldw, DenseOut d) {
  constexpr int TI = BM / 32, TJ = BN / 32;           // 16 x 16 products per wavefront block
  constexpr int WBM = BM / 2, WBN = BN / 2, ELD = WBN + 4;
  constexpr int STAGE = 2 * (BM + BN) * DL_LD * 2, EPI = 4 * 32 * ELD * 4;
  __shared__ __attribute__((aligned(16))) unsigned char smem[STAGE > EPI ? STAGE : EPI];
  uint16_t* const As = reinterpret_cast<uint16_t*>(smem);                 // [2][BM * DL_LD]
  uint16_t* const Bs = As + 2 * BM * DL_LD;                                 // [2][BN * DL_LD]
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  const int wm = (w >> 1) * WBM, wn = (w & 1) * WBN;
  // staging: thread t moves rows t/4 (+ 64 h) of both slices, 8 values at column 8 (t % 4)
  const int sr = tid >> 2, sk = (tid & 3) * 8;
  constexpr int HA = BM / 64, HB2 = BN / 64;
  const uint16_t* xa[HA];
  const uint16_t* wb[HB2];
  bool xo[HA], wo[HB2];
#pragma unroll
  for (int h = 0; h < HA; ++h) {
    const int r = m0 + sr + 64 * h;
    xo[h] = r < M;
    xa[h] = X + (long)(xo[h] ? r : 0) * ldx + sk;
  }
#pragma unroll
  for (int h = 0; h < HB2; ++h) {
    const int c = n0 + sr + 64 * h;
    wo[h] = c < N;
    wb[h] = W + (long)(wo[h] ? c : 0) * ldw + sk;
  }
  const uint4 z4 = make_uint4(0u, 0u, 0u, 0u);
  uint4 ra[HA], rb[HB2];
  auto fetch = [&](int k0) {
    const bool kin = k0 + sk < K;
#pragma unroll
    for (int h = 0; h < HA; ++h) ra[h] = (xo[h] && kin) ? *reinterpret_cast<const uint4*>(xa[h] + k0) : z4;
#pragma unroll
    for (int h = 0; h < HB2; ++h) rb[h] = (wo[h] && kin) ? *reinterpret_cast<const uint4*>(wb[h] + k0) : z4;
  };
  auto stash = [&](int buf) {
#pragma unroll
    for (int h = 0; h < HA; ++h)
      *reinterpret_cast<uint4*>(&As[buf * BM * DL_LD + (sr + 64 * h) * DL_LD + sk]) = ra[h];
#pragma unroll
    for (int h = 0; h < HB2; ++h)
      *reinterpret_cast<uint4*>(&Bs[buf * BN * DL_LD + (sr + 64 * h) * DL_LD + sk]) = rb[h];
  };
  f32x4 acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int li = lane & 15, lk = (lane >> 4) * 8;
  fetch(0);
  stash(0);
  __syncthreads();
  int buf = 0;
  for (int k0 = 0; k0 < K; k0 += DL_K) {
    const bool more = k0 + DL_K < K;
    if (more) fetch(k0 + DL_K);
    bf16x8 a[TI], b[TJ];
#pragma unroll
    for (int i = 0; i < TI; ++i)
      a[i] = *reinterpret_cast<const bf16x8*>(&As[buf * BM * DL_LD + (wm + 16 * i + li) * DL_LD + lk]);
#pragma unroll
    for (int j = 0; j < TJ; ++j)
      b[j] = *reinterpret_cast<const bf16x8*>(&Bs[buf * BN * DL_LD + (wn + 16 * j + li) * DL_LD + lk]);
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    if (more) stash(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }
  // epilogue: Z = acc + bias into this wavefront's LDS block (the staging buffers are free after the last
  // barrier), 32 rows at a time, then rows of 4-value vectors to memory
  constexpr int ER = 32;               // rows staged per pass
  float* const E = reinterpret_cast<float*>(smem) + w * ER * ELD;
  constexpr int LPR = WBN / 4;         // lanes per row (4 values each)
  constexpr int RPI = 64 / LPR;        // rows per store pass
  const int cq = (lane % LPR) * 4, col = n0 + wn + cq;
  const bool vec = (N & 3) == 0;
#pragma unroll
  for (int half = 0; half < WBM / ER; ++half) {
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
      const int cl = 16 * j + li, cg = n0 + wn + cl;
      const float bj = (d.bias && cg < N) ? d.bias[cg] : 0.f;
#pragma unroll
      for (int i = half * (ER / 16); i < (half + 1) * (ER / 16); ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) E[(16 * i - half * ER + (lane >> 4) * 4 + q) * ELD + cl] = acc[i][j][q] + bj;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int r = lane / LPR; r < ER; r += RPI) {
      const int row = m0 + wm + half * ER + r;
      if (row >= M || col >= N) continue;
      const f32x4 z = *reinterpret_cast<const f32x4*>(&E[r * ELD + cq]);
      f32x4 h = z;
      if (d.act) {
        float dg;
#pragma unroll
        for (int q = 0; q < 4; ++q) h[q] = gelu(z[q], dg);
      }
      const long e = (long)row * N + col;
      if (vec) {
        if (d.Z) *reinterpret_cast<f32x4*>(d.Z + e) = z;
        if (d.Hf) *reinterpret_cast<f32x4*>(d.Hf + e) = h;
        if (d.H) {
          const unsigned lo = (unsigned)to_bf16(h[0]) | ((unsigned)to_bf16(h[1]) << 16);
          const unsigned hi = (unsigned)to_bf16(h[2]) | ((unsigned)to_bf16(h[3]) << 16);
          *reinterpret_cast<uint2*>(d.H + e) = make_uint2(lo, hi);
        }
      } else {
        for (int q = 0; q < 4 && col + q < N; ++q) {
          if (d.Z) d.Z[e + q] = z[q];
          if (d.Hf) d.Hf[e + q] = h[q];
          if (d.H) d.H[e + q] = to_bf16(h[q]);
        }
      }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
}

__device__ inline float gelu(float z, float& dg) {
  const float c = 0.7978845608028654f, a = 0.044715f;
  const float u = c * (z + a * z * z * z);
  const float t = tanhf(u);
  dg = 0.5f * (1.f + t) + 0.5f * z * (1.f - t * t) * c * (1.f + 3.f * a * z * z);
  return 0.5f * z * (1.f + t);
}

__global__ void k_act_fwd(int M, int N, int act, const float* __restrict__ Zacc, const float* __restrict__ bias,
                          float* __restrict__ Z, uint16_t* __restrict__ H, float* __restrict__ Hf) {
  const long n = (long)M * N;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const float z = Zacc[e] + (bias ? bias[e % N] : 0.f);
    float dg;
    const float h = act ? gelu(z, dg) : z;
    if (Z) Z[e] = z;
    if (H) H[e] = to_bf16(h);
    if (Hf) Hf[e] = h;
  }
}

// dZ = dH * act'(Z); one thread per column also sums the bias gradient over the rows.
__global__ void k_act_bwd(int M, int N, int act, const float* __restrict__ Z, const float* __restrict__ dH,
                          float* __restrict__ dZ, uint16_t* __restrict__ dZb, float* __restrict__ dbias) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < N; c += gridDim.x * blockDim.x) {
    float s = 0.f;
    for (int r = 0; r < M; ++r) {
      const long e = (long)r * N + c;
      float dg = 1.f;
      if (act) (void)gelu(Z[e], dg);
      const float g = dH[e] * dg;
      if (dZ) dZ[e] = g;
      if (dZb) dZb[e] = to_bf16(g);
      s += g;
    }
    if (dbias) dbias[c] = s;
  }
}

// Mean squared error over M x N; dY = 2 (Y - T) / (M N); loss accumulated per block.
__global__ void k_mse(int M, int N, const float* __restrict__ Y, const float* __restrict__ T, float* __restrict__ dY,
                      uint16_t* __restrict__ dYb, float* __restrict__ loss) {
  __shared__ float part[256];
  const long n = (long)M * N;
  const float sc = 2.f / (float)n;
  float s = 0.f;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const float d = Y[e] - T[e];
    s += d * d;
    if (dY) dY[e] = sc * d;
    if (dYb) dYb[e] = to_bf16(sc * d);
  }
  part[threadIdx.x] = s;
  __syncthreads();
  for (int o = blockDim.x / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) part[threadIdx.x] += part[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0 && loss) atomicAdd(loss, part[0] / (float)n);
}

__global__ void k_adam(long n, float lr, float b1, float b2, float eps, float c1, float c2, float* __restrict__ w,
                       const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       uint16_t* __restrict__ wb) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const float gi = g[e];
    const float mi = b1 * m[e] + (1.f - b1) * gi;
    const float vi = b2 * v[e] + (1.f - b2) * gi * gi;
    m[e] = mi;
    v[e] = vi;
    const float wi = w[e] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[e] = wi;
    if (wb) wb[e] = to_bf16(wi);
  }
}

__global__ void k_to_bf16(long n, const float* __restrict__ x, uint16_t* __restrict__ y) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    y[e] = to_bf16(x[e]);
}

unsigned grid1(long n) { return (unsigned)std::min<long>(65535, std::max<long>(1, (n + 255) / 256)); }

}  // namespace

void gemm_bf16(int M, int N, int K, const uint16_t* A, long sam, long sak, const uint16_t* B, long sbk, long sbn,
               float alpha, float beta, float* C, long ldc, hipStream_t s) {
  CWX_CHECK(M >= 0 && N >= 0 && K >= 0, "negative GEMM dimension");
  if (M == 0 || N == 0) return;
  CWX_CHECK((M + GT - 1) / GT <= 65535, "GEMM: too many row tiles");
  const bool kc = sak == 1 && sbk == 1 && K % 4 == 0 && sam % 4 == 0 && sbn % 4 == 0 &&
                  ((uintptr_t)A & 7) == 0 && ((uintptr_t)B & 7) == 0;
  if (kc)
    hipLaunchKernelGGL(k_gemm_bf16_kc, dim3((N + KC_T - 1) / KC_T, (M + KC_T - 1) / KC_T), dim3(256), 0, s, M, N, K,
                       A, sam, B, sbn, alpha, beta, C, ldc, DenseOut{});
  else
    hipLaunchKernelGGL(k_gemm_bf16, dim3((N + GT - 1) / GT, (M + GT - 1) / GT), dim3(256), 0, s, M, N, K, A, sam,
                       sak, B, sbk, sbn, alpha, beta, C, ldc);
  CWX_HIP(hipGetLastError());
}

void mlp_dense_fwd(int M, int N, int K, const uint16_t* X, long ldx, const uint16_t* W, long ldw, const float* bias,
                   int act, float* Z, uint16_t* H, float* Hf, hipStream_t s) {
  CWX_CHECK(M >= 0 && N >= 1 && K >= 0 && (act == 0 || act == 1), "bad dense-layer arguments");
  CWX_CHECK(K % 4 == 0 && ldx % 4 == 0 && ldw % 4 == 0 && ((uintptr_t)X & 7) == 0 && ((uintptr_t)W & 7) == 0,
            "dense layer: the input width and both row strides must be multiples of 4 BF16 values, 8-byte aligned");
  if (M == 0) return;
  CWX_CHECK((M + KC_T - 1) / KC_T <= 65535, "dense layer: too many row tiles");
  DenseOut d;
  d.bias = bias;
  d.act = act;
  d.Z = Z;
  d.H = H;
  d.Hf = Hf;
  if (K % 8 == 0 && ldx % 8 == 0 && ldw % 8 == 0 && ((uintptr_t)X & 15) == 0 && ((uintptr_t)W & 15) == 0) {
    // 64 x 64 tiles: at the closure's shapes (batch 16384, widths 256 and 2700) they ran the forward pass in
    // 0.120 ms against 0.168 ms with 128 x 128 tiles, whose occupancy (140 VGPRs, one or two workgroups per
    // compute unit) leaves the global loads and the epilogue's stores exposed
    CWX_CHECK((M + 63) / 64 <= 65535, "dense layer: too many row tiles");
    hipLaunchKernelGGL((k_dense_bf16<64, 64>), dim3((N + 63) / 64, (M + 63) / 64), dim3(256), 0, s, M, N, K, X, ldx, W,
                       ldw, d);
    CWX_HIP(hipGetLastError());
    return;
  }
  hipLaunchKernelGGL(k_gemm_bf16_kc, dim3((N + KC_T - 1) / KC_T, (M + KC_T - 1) / KC_T), dim3(256), 0, s, M, N, K, X,
                     ldx, W, ldw, 1.f, 0.f, nullptr, (long)N, d);
  CWX_HIP(hipGetLastError());
}

void mlp_act_fwd(int M, int N, int act, const float* Zacc, const float* bias, float* Z, uint16_t* H, float* Hf,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_act_fwd, dim3(grid1((long)M * N)), dim3(256), 0, s, M, N, act, Zacc, bias, Z, H, Hf);
  CWX_HIP(hipGetLastError());
}

void mlp_act_bwd(int M, int N, int act, const float* Z, const float* dH, float* dZ, uint16_t* dZb, float* dbias,
                 hipStream_t s) {
  hipLaunchKernelGGL(k_act_bwd, dim3(grid1(N)), dim3(256), 0, s, M, N, act, Z, dH, dZ, dZb, dbias);
  CWX_HIP(hipGetLastError());
}

void mlp_mse(int M, int N, const float* Y, const float* T, float* dY, uint16_t* dYb, float* loss, hipStream_t s) {
  hipLaunchKernelGGL(k_mse, dim3(std::min<unsigned>(grid1((long)M * N), 1024)), dim3(256), 0, s, M, N, Y, T, dY, dYb,
                     loss);
  CWX_HIP(hipGetLastError());
}

void mlp_adam(long n, float lr, float b1, float b2, float eps, int step, float* w, const float* g, float* m, float* v,
              uint16_t* wb, hipStream_t s) {
  CWX_CHECK(step >= 1, "Adam steps count from 1");
  const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
  hipLaunchKernelGGL(k_adam, dim3(grid1(n)), dim3(256), 0, s, n, lr, b1, b2, eps, c1, c2, w, g, m, v, wb);
  CWX_HIP(hipGetLastError());
}

void to_bf16(long n, const float* x, uint16_t* y, hipStream_t s) {
  hipLaunchKernelGGL(k_to_bf16, dim3(grid1(n)), dim3(256), 0, s, n, x, y);
  CWX_HIP(hipGetLastError());
}

}  // namespace cwx
