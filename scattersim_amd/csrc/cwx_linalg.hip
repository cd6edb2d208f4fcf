// Dense complex FP64 linear algebra for the coupled-wave solve: batched GEMM, blocked LU with
// partial pivoting and the matching triangular solves.
//
// Replaces the LAPACK/BLAS seam of third_Lapack.h:47-85 (LAPACKE_zgesv, the zgesv-with-identity
// inverse and cblas_zgemm).  Storage is column-major complex<double> with explicit leading
// dimensions, as LAPACK; every entry point is batched over matrices placed at a fixed element stride.
#include "cwx_common.h"
#include "cwx_kernels.h"
#include "cwx_tune.h"

namespace cwx {

namespace {

// Pointers into global memory: loads and stores through them are global_* instructions (flat ones count
// against the LDS wait counter too, so an LDS wait would also wait for every load in flight).
typedef __attribute__((address_space(1))) const zc gzc;
typedef __attribute__((address_space(1))) zc gzw;
typedef double dv2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) dv2 gdv2;
// An element of a global-memory view: reads and writes are 16-byte global vector accesses.
struct GRef {
  gzw* p;
  __device__ operator zc() const {
    const dv2 v = *(const gdv2*)p;
    return mk(v.x, v.y);
  }
  __device__ GRef& operator=(zc v) {
    *(gdv2*)p = dv2{v.re, v.im};
    return *this;
  }
  __device__ GRef& operator=(const GRef& o) { return *this = (zc)o; }
};
// A complex array in global memory, indexed like a pointer.
struct GP {
  gzw* p;
  __device__ GRef operator[](long i) const { return GRef{p + i}; }
  __device__ GP operator+(long i) const { return GP{p + i}; }
};
__device__ inline GP gview(const zc* p) { return GP{(gzw*)p}; }

// ------------------------------------------------------------------------------------------------
// GEMM: 64x64 output tile per 256-lane workgroup, K staged through LDS in steps of 16, each lane
// owning a 4x4 block of outputs at stride 16 (conflict-free broadcast reads from LDS).
// ------------------------------------------------------------------------------------------------
constexpr int GM = 64, GN = 64, GK = 16;

__device__ inline zc load_op(const zc* __restrict__ A, int ld, Op op, int r, int c) {
  // element (r, c) of op(A)
  if (op == OpN) return A[(long)c * ld + r];
  zc v = A[(long)r * ld + c];
  return op == OpC ? conj(v) : v;
}

__global__ __launch_bounds__(256) void k_zgemm(Op opa, Op opb, int m, int n, int k, zc alpha, MatSet Aset,
                                               int lda, MatSet Bset, int ldb, zc beta, MatSet Cset, int ldc) {
  __shared__ zc As[GK][GM];
  __shared__ zc Bs[GK][GN];
  const int b = blockIdx.z;
  const zc* __restrict__ A = Aset.at(b);
  const zc* __restrict__ B = Bset.at(b);
  zc* __restrict__ C = Cset.at(b);
  const int m0 = blockIdx.x * GM, n0 = blockIdx.y * GN;
  const int t = threadIdx.x;
  const int tx = t & 15, ty = t >> 4;
  zc acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = mk(0.0);

  for (int k0 = 0; k0 < k; k0 += GK) {
    // A tile: GM x GK = 1024 elements, 4 per lane; lanes run along rows for op N (coalesced)
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      int e = t + 256 * u;
      int r, kk;
      if (opa == OpN) { r = e & (GM - 1); kk = e >> 6; }
      else { kk = e & (GK - 1); r = e >> 4; }
      int gr = m0 + r, gk = k0 + kk;
      As[kk][r] = (gr < m && gk < k) ? load_op(A, lda, opa, gr, gk) : mk(0.0);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      int e = t + 256 * u;
      int kk, c;
      if (opb == OpN) { kk = e & (GK - 1); c = e >> 4; }
      else { c = e & (GN - 1); kk = e >> 6; }
      int gk = k0 + kk, gc = n0 + c;
      Bs[kk][c] = (gk < k && gc < n) ? load_op(B, ldb, opb, gk, gc) : mk(0.0);
    }
    __syncthreads();
#pragma unroll 4
    for (int kk = 0; kk < GK; ++kk) {
      zc a[4], bb[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][tx + 16 * i];
#pragma unroll
      for (int j = 0; j < 4; ++j) bb[j] = Bs[kk][ty + 16 * j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) fma_acc(acc[i][j], a[i], bb[j]);
    }
    __syncthreads();
  }
  const bool beta0 = (beta.re == 0.0 && beta.im == 0.0);
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    int gc = n0 + ty + 16 * j;
    if (gc >= n) continue;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      int gr = m0 + tx + 16 * i;
      if (gr >= m) continue;
      zc* cp = C + (long)gc * ldc + gr;
      zc v = alpha * acc[i][j];
      if (!beta0) v = v + beta * (*cp);
      *cp = v;
    }
  }
}

// ------------------------------------------------------------------------------------------------
// Matrix-vector products (n <= 4 right-hand sides, op(A) = A): 16 rows per workgroup; the 64 column groups
// (16 wavefronts x 4 quarter-wavefronts, each quarter on the 16 rows) take the columns c = g (mod 64), four
// columns in flight per lane, and the partial sums are combined in LDS in a fixed order (results do not
// depend on scheduling or on the batch).  A 900-row matrix keeps 57 workgroups busy, each lane with about
// four rounds of loads (64 rows per workgroup took fourteen).
// ------------------------------------------------------------------------------------------------
constexpr int GV_MAXN = 4, GV_WAVES = 16, GV_UNROLL = 4, GV_ROWS = 16, GV_GROUPS = GV_WAVES * 64 / GV_ROWS;

__global__ __launch_bounds__(64 * GV_WAVES) void k_zgemv(int m, int n, int k, zc alpha, MatSet Aset, int lda,
                                                         MatSet Bset, int ldb, zc beta, MatSet Cset, int ldc) {
  __shared__ zc part[GV_GROUPS][GV_ROWS];
  const int b = blockIdx.y;
  const GP A = gview(Aset.at(b));
  const GP B = gview(Bset.at(b));
  zc* __restrict__ C = Cset.at(b);
  const int rr = threadIdx.x % GV_ROWS, g = threadIdx.x / GV_ROWS;
  const int r = blockIdx.x * GV_ROWS + rr;
  zc acc[GV_MAXN][GV_UNROLL];
#pragma unroll
  for (int j = 0; j < GV_MAXN; ++j)
#pragma unroll
    for (int u = 0; u < GV_UNROLL; ++u) acc[j][u] = mk(0.0);
  if (r < m) {
    constexpr int step = GV_GROUPS * GV_UNROLL;
    int c = g;
    for (; c + (GV_UNROLL - 1) * GV_GROUPS < k; c += step) {
      zc a[GV_UNROLL];
#pragma unroll
      for (int u = 0; u < GV_UNROLL; ++u) a[u] = A[(long)(c + u * GV_GROUPS) * lda + r];
#pragma unroll
      for (int j = 0; j < GV_MAXN; ++j)
        if (j < n) {
#pragma unroll
          for (int u = 0; u < GV_UNROLL; ++u) fma_acc(acc[j][u], a[u], B[(long)j * ldb + c + u * GV_GROUPS]);
        }
    }
    for (; c < k; c += GV_GROUPS) {
      const zc a0 = A[(long)c * lda + r];
#pragma unroll
      for (int j = 0; j < GV_MAXN; ++j)
        if (j < n) fma_acc(acc[j][0], a0, B[(long)j * ldb + c]);
    }
  }
  const bool beta0 = (beta.re == 0.0 && beta.im == 0.0);
#pragma unroll
  for (int j = 0; j < GV_MAXN; ++j) {
    if (j >= n) break;
    if (j > 0) __syncthreads();  // the previous column's partials have been read
    part[g][rr] = (acc[j][0] + acc[j][1]) + (acc[j][2] + acc[j][3]);
    __syncthreads();
    if (threadIdx.x < GV_ROWS && r < m) {
      // the 64 groups in a fixed order, four running sums
      zc t0 = part[0][rr], t1 = part[1][rr], t2 = part[2][rr], t3 = part[3][rr];
#pragma unroll
      for (int w = 4; w < GV_GROUPS; w += 4) {
        t0 = t0 + part[w][rr];
        t1 = t1 + part[w + 1][rr];
        t2 = t2 + part[w + 2][rr];
        t3 = t3 + part[w + 3][rr];
      }
      zc* cp = C + (long)j * ldc + r;
      zc v = alpha * ((t0 + t1) + (t2 + t3));
      if (!beta0) v = v + beta * (*cp);
      *cp = v;
    }
  }
}

// ------------------------------------------------------------------------------------------------
// GEMM on the FP64 matrix cores (v_mfma_f64_16x16x4f64).  On MI355X one such instruction keeps its SIMD's
// matrix unit busy for 64 cycles, so the kernel is built to issue them back to back: each wavefront owns
// TM x TN tiles of 16x16 outputs (the workgroup 2 x 2 wavefronts, a (32 TM) x (32 TN) output tile), the
// next 16-deep k panel is loaded from global memory into registers while the current one is multiplied out
// of LDS, and the operands sit in LDS as (re, im) pairs, read as 16-byte words.  A complex product is four
// real products into two accumulators, Re += Br Ar - Bi Ai and Im += Bi Ar + Br Ai.  The operands enter the
// instruction swapped (the B fragment first), so each lane's results are consecutive rows of the
// column-major C.  Lane layout of v_mfma_f64_16x16x4f64: A[i][k] from lane i + 16k, B[k][j] from lane
// j + 16k, D[(l / 16) + 4q][l % 16] in register q of lane l.
// ------------------------------------------------------------------------------------------------
typedef double d4 __attribute__((ext_vector_type(4)));
constexpr int MK = 16;

// Operands are read through global-address-space pointers: a flat load would also count against the LDS
// wait counter, so every LDS read's wait would wait for the prefetched panel too.
__device__ inline dv2 gload2(gzc* p) { return *(const __attribute__((address_space(1))) dv2*)(p); }

template <int TM, int TN>
__global__ __launch_bounds__(256, 2) void k_zgemm_mfma(Op opa, Op opb, int m, int n, int k, zc alpha, MatSet Aset,
                                                    int lda, MatSet Bset, int ldb, zc beta, MatSet Cset, int ldc) {
  constexpr int BM = 32 * TM, BN = 32 * TN;
  constexpr int NA = BM * MK / 256, NB = MK * BN / 256;
  __shared__ double2 As[MK][BM + 1], Bs[MK][BN + 1];
  const int b = blockIdx.z;
  gzc* A = (gzc*)Aset.at(b);
  gzc* B = (gzc*)Bset.at(b);
  zc* __restrict__ C = Cset.at(b);
  const int m0 = blockIdx.x * BM, n0 = blockIdx.y * BN;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int wm = (wv & 1) * 16 * TM, wn = (wv >> 1) * 16 * TN;
  // element e of a k panel: (row, k) of op(A), (k, column) of op(B); the fastest index follows memory
  auto a_at = [&](int e, int& r, int& kk) {
    if (opa == OpN) { r = e % BM; kk = e / BM; }
    else { kk = e % MK; r = e / MK; }
  };
  auto b_at = [&](int e, int& kk, int& c) {
    if (opb == OpN) { kk = e % MK; c = e / MK; }
    else { c = e % BN; kk = e / BN; }
  };
  // The next panel's loads are issued unconditionally (out-of-range elements read element (0, 0), which
  // exists) and masked when they are written to LDS: a load under a branch, or a select right after it,
  // would make the wavefront wait for it before the multiplications.
  dv2 pa[NA], pb[NB];
  unsigned va = 0, vb = 0;
  auto fetch = [&](int k0) {
    va = vb = 0;
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      int r, kk;
      a_at(t + 256 * u, r, kk);
      const int gr = m0 + r, gk = k0 + kk;
      const bool ok = gr < m && gk < k;
      va |= ok ? 1u << u : 0u;
      const long off = ok ? (opa == OpN ? (long)gk * lda + gr : (long)gr * lda + gk) : 0;
      pa[u] = gload2(A + off);
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      int kk, c;
      b_at(t + 256 * u, kk, c);
      const int gk = k0 + kk, gc = n0 + c;
      const bool ok = gk < k && gc < n;
      vb |= ok ? 1u << u : 0u;
      const long off = ok ? (opb == OpN ? (long)gc * ldb + gk : (long)gk * ldb + gc) : 0;
      pb[u] = gload2(B + off);
    }
  };
  const double sa = opa == OpC ? -1.0 : 1.0, sb = opb == OpC ? -1.0 : 1.0;
  d4 Re[TM][TN], Im[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) Re[i][j] = Im[i][j] = d4{0, 0, 0, 0};
  fetch(0);
  for (int k0 = 0; k0 < k; k0 += MK) {
#pragma unroll
    for (int u = 0; u < NA; ++u) {
      int r, kk;
      a_at(t + 256 * u, r, kk);
      const bool ok = (va >> u) & 1u;
      As[kk][r] = double2{ok ? pa[u].x : 0.0, ok ? sa * pa[u].y : 0.0};
    }
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      int kk, c;
      b_at(t + 256 * u, kk, c);
      const bool ok = (vb >> u) & 1u;
      Bs[kk][c] = double2{ok ? pb[u].x : 0.0, ok ? sb * pb[u].y : 0.0};
    }
    __syncthreads();
    if (k0 + MK < k) fetch(k0 + MK);  // in flight while this panel is multiplied
    auto step = [&](const double2 (&a)[TM], const double2 (&bb)[TN]) {
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          Re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bb[j].x, a[i].x, Re[i][j], 0, 0, 0);
          Im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bb[j].y, a[i].x, Im[i][j], 0, 0, 0);
        }
#pragma unroll
      for (int i = 0; i < TM; ++i)
#pragma unroll
        for (int j = 0; j < TN; ++j) {
          Re[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(-bb[j].y, a[i].y, Re[i][j], 0, 0, 0);
          Im[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bb[j].x, a[i].y, Im[i][j], 0, 0, 0);
        }
    };
    auto frags = [&](int k4, double2 (&a)[TM], double2 (&bb)[TN]) {
      const int kk = k4 + (lane >> 4);
#pragma unroll
      for (int i = 0; i < TM; ++i) a[i] = As[kk][wm + 16 * i + (lane & 15)];
#pragma unroll
      for (int j = 0; j < TN; ++j) bb[j] = Bs[kk][wn + 16 * j + (lane & 15)];
    };
#pragma unroll
    for (int k4 = 0; k4 < MK; k4 += 4) {
      double2 a[TM], bb[TN];
      frags(k4, a, bb);
      step(a, bb);
    }
    __syncthreads();
  }
  const bool beta0 = (beta.re == 0.0 && beta.im == 0.0);
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int gr = m0 + wm + 16 * i + (lane & 15);
        const int gc = n0 + wn + 16 * j + (lane >> 4) + 4 * q;
        if (gr < m && gc < n) {
          zc* cp = C + (long)gc * ldc + gr;
          zc v = alpha * mk(Re[i][j][q], Im[i][j][q]);
          if (!beta0) v = v + beta * (*cp);
          *cp = v;
        }
      }
}

// ------------------------------------------------------------------------------------------------
// Mixed-precision GEMM on the FP32 matrix cores (v_mfma_f32_16x16x4f32): every FP64 operand is split
// into hi = fl32(x) and lo = fl32(x - hi), and each real product X Y is taken as Xh Yh + Xh Yl + Xl Yh
// (Xl Yl, about 2^-48 of it, is dropped); with SPLIT false only Xh Yh is taken, a plain FP32 product of
// the rounded operands.  The matrix core accumulates 16 values of k in FP32; each
// such partial sum is then added to FP64 accumulators.  A complex product is four real products, so a
// complex tile costs 12 FP32 matrix instructions where the FP64 kernel issues 4 FP64 ones, against a
// 2:1 FP32:FP64 matrix rate on MI355X; the FP32 rounding inside each 16-term partial sum bounds the
// accuracy near 2^-24 of the partial sums.  The tiling of k_zgemm_mfma<2, 2>: a 64x64 output tile per
// 256-lane workgroup, each wavefront a 32x32 quarter as 2x2 tiles of 16x16.
// Lane layout of v_mfma_f32_16x16x4f32: A[i][k] from lane i + 16k, B[k][j] from lane j + 16k,
// D[4 (l / 16) + q][l % 16] in register q of lane l.
// ------------------------------------------------------------------------------------------------
typedef float f4v __attribute__((ext_vector_type(4)));

template <bool SPLIT>
__global__ __launch_bounds__(256, 2) void k_zgemm_split32(Op opa, Op opb, int m, int n, int k, zc alpha,
                                                       MatSet Aset, int lda, MatSet Bset, int ldb, zc beta,
                                                       MatSet Cset, int ldc) {
  // planes: 0 re-hi, 1 re-lo, 2 im-hi, 3 im-lo
  __shared__ float As[4][MK][GM + 1], Bs[4][MK][GN + 1];
  const int b = blockIdx.z;
  const zc* __restrict__ A = Aset.at(b);
  const zc* __restrict__ B = Bset.at(b);
  zc* __restrict__ C = Cset.at(b);
  const int m0 = blockIdx.x * GM, n0 = blockIdx.y * GN;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int wm = (wv & 1) * 32, wn = (wv >> 1) * 32;
  double PD[2][2][4], QD[2][2][4], RD[2][2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) PD[i][j][q] = QD[i][j][q] = RD[i][j][q] = 0.0;
  auto split = [](double x, float& hi, float& lo) {
    hi = (float)x;
    lo = (float)(x - (double)hi);
  };
  for (int k0 = 0; k0 < k; k0 += MK) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = t + 256 * u;
      int r, kk;
      if (opa == OpN) { r = e & (GM - 1); kk = e >> 6; }
      else { kk = e & (MK - 1); r = e >> 4; }
      const int gr = m0 + r, gk = k0 + kk;
      const zc v = (gr < m && gk < k) ? load_op(A, lda, opa, gr, gk) : mk(0.0);
      split(v.re, As[0][kk][r], As[1][kk][r]);
      split(v.im, As[2][kk][r], As[3][kk][r]);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int e = t + 256 * u;
      int kk, c;
      if (opb == OpN) { kk = e & (MK - 1); c = e >> 4; }
      else { c = e & (GN - 1); kk = e >> 6; }
      const int gk = k0 + kk, gc = n0 + c;
      const zc v = (gk < k && gc < n) ? load_op(B, ldb, opb, gk, gc) : mk(0.0);
      split(v.re, Bs[0][kk][c], Bs[1][kk][c]);
      split(v.im, Bs[2][kk][c], Bs[3][kk][c]);
    }
    __syncthreads();
    f4v P[2][2], Q[2][2], R[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) P[i][j] = Q[i][j] = R[i][j] = f4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k4 = 0; k4 < MK; k4 += 4) {
      const int kk = k4 + (lane >> 4);
      float a[4][2], bb[4][2];  // [plane][tile]
#pragma unroll
      for (int pl = 0; pl < 4; ++pl)
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          a[pl][i] = As[pl][kk][wm + 16 * i + (lane & 15)];
          bb[pl][i] = Bs[pl][kk][wn + 16 * i + (lane & 15)];
        }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          // Re Re, Im Im, Re Im + Im Re: the hi hi terms, then (split form) hi lo and lo hi
          P[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0][i], bb[0][j], P[i][j], 0, 0, 0);
          Q[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[2][i], bb[2][j], Q[i][j], 0, 0, 0);
          R[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0][i], bb[2][j], R[i][j], 0, 0, 0);
          R[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[2][i], bb[0][j], R[i][j], 0, 0, 0);
          if (SPLIT) {
            P[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0][i], bb[1][j], P[i][j], 0, 0, 0);
            P[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1][i], bb[0][j], P[i][j], 0, 0, 0);
            Q[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[2][i], bb[3][j], Q[i][j], 0, 0, 0);
            Q[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[3][i], bb[2][j], Q[i][j], 0, 0, 0);
            R[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0][i], bb[3][j], R[i][j], 0, 0, 0);
            R[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[1][i], bb[2][j], R[i][j], 0, 0, 0);
            R[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[2][i], bb[1][j], R[i][j], 0, 0, 0);
            R[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[3][i], bb[0][j], R[i][j], 0, 0, 0);
          }
        }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          PD[i][j][q] += (double)P[i][j][q];
          QD[i][j][q] += (double)Q[i][j][q];
          RD[i][j][q] += (double)R[i][j][q];
        }
    __syncthreads();
  }
  const bool beta0 = (beta.re == 0.0 && beta.im == 0.0);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const int gr = m0 + wm + 16 * i + 4 * (lane >> 4) + q;
        const int gc = n0 + wn + 16 * j + (lane & 15);
        if (gr < m && gc < n) {
          zc* cp = C + (long)gc * ldc + gr;
          zc v = alpha * mk(PD[i][j][q] - QD[i][j][q], RD[i][j][q]);
          if (!beta0) v = v + beta * (*cp);
          *cp = v;
        }
      }
}

// ------------------------------------------------------------------------------------------------
// LU panel: one workgroup per matrix factors columns [j0, j0+nb) over rows [j0, n) with partial
// pivoting (pivot = first maximum of |re|+|im|, as izamax).  Swaps are applied to the panel columns
// only; the other columns are swapped by k_lu_swap_trsm.
// ------------------------------------------------------------------------------------------------
constexpr int LU_NB = 32;
constexpr int LU_THREADS = 256;

__device__ inline void block_argmax(double v, int idx, double* sv, int* si, double& outv, int& outi) {
  // reduce (value, index) pairs: larger value wins, ties -> smaller index.  After one barrier every
  // lane reduces the per-wavefront partials itself.  The partials are rewritten by the next call, so
  // callers put a barrier between consecutive calls.
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  // within each row of 16 lanes: rotations by 1, 2, 4, 8 (DPP row_ror) leave the row's best pair in
  // every lane of the row; then across rows through LDS permutes
  auto ror = [](double x, int ctrl_sel) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
    unsigned lo, hi;
    switch (ctrl_sel) {
      case 1: lo = __builtin_amdgcn_mov_dpp((int)(unsigned)u, 0x121, 0xF, 0xF, false);
              hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), 0x121, 0xF, 0xF, false); break;
      case 2: lo = __builtin_amdgcn_mov_dpp((int)(unsigned)u, 0x122, 0xF, 0xF, false);
              hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), 0x122, 0xF, 0xF, false); break;
      case 4: lo = __builtin_amdgcn_mov_dpp((int)(unsigned)u, 0x124, 0xF, 0xF, false);
              hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), 0x124, 0xF, 0xF, false); break;
      default: lo = __builtin_amdgcn_mov_dpp((int)(unsigned)u, 0x128, 0xF, 0xF, false);
               hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), 0x128, 0xF, 0xF, false); break;
    }
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
  };
  auto rori = [](int x, int ctrl_sel) {
    switch (ctrl_sel) {
      case 1: return __builtin_amdgcn_mov_dpp(x, 0x121, 0xF, 0xF, false);
      case 2: return __builtin_amdgcn_mov_dpp(x, 0x122, 0xF, 0xF, false);
      case 4: return __builtin_amdgcn_mov_dpp(x, 0x124, 0xF, 0xF, false);
      default: return __builtin_amdgcn_mov_dpp(x, 0x128, 0xF, 0xF, false);
    }
  };
#pragma unroll
  for (int sh = 1; sh <= 8; sh <<= 1) {
    const double ov = ror(v, sh);
    const int oi = rori(idx, sh);
    if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
  }
  for (int off = 16; off < 64; off <<= 1) {
    double ov = __shfl_xor(v, off);
    int oi = __shfl_xor(idx, off);
    if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
  }
  if (lane == 0) { sv[wid] = v; si[wid] = idx; }
  __syncthreads();
  double bv = sv[0];
  int bi = si[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
    if (sv[w] > bv || (sv[w] == bv && si[w] < bi)) { bv = sv[w]; bi = si[w]; }
  outv = bv;
  outi = bi;
}

__global__ __launch_bounds__(LU_THREADS) void k_lu_panel(int n, int j0, int nb, MatSet Aset, int lda,
                                                         IntSet pset, int* info) {
  __shared__ double sv[LU_THREADS / 64];
  __shared__ int si[LU_THREADS / 64];
  __shared__ zc urow[LU_NB];
  const int b = blockIdx.x;
  zc* A = Aset.at(b);
  int* ipiv = pset.at(b);
  for (int jj = 0; jj < nb; ++jj) {
    const int j = j0 + jj;
    // pivot search over rows j..n-1 of column j
    double best = -1.0;
    int bidx = n;
    for (int r = j + threadIdx.x; r < n; r += blockDim.x) {
      double v = abs1(A[(long)j * lda + r]);
      if (v > best) { best = v; bidx = r; }
    }
    int p;
    double pval;
    block_argmax(best, bidx, sv, si, pval, p);
    if (threadIdx.x == 0) {
      ipiv[j] = p;
      if (pval == 0.0 && info[b] == 0) info[b] = j + 1;
    }
    // swap rows j and p inside the panel columns
    if (p != j) {
      for (int c = j0 + threadIdx.x; c < j0 + nb; c += blockDim.x) {
        zc t = A[(long)c * lda + j];
        A[(long)c * lda + j] = A[(long)c * lda + p];
        A[(long)c * lda + p] = t;
      }
    }
    __syncthreads();
    const zc piv = A[(long)j * lda + j];
    if (pval != 0.0) {
      const zc rinv = zdiv(mk(1.0), piv);
      for (int r = j + 1 + threadIdx.x; r < n; r += blockDim.x) A[(long)j * lda + r] = A[(long)j * lda + r] * rinv;
    }
    __syncthreads();
    // rank-1 update of the remaining panel columns: the pivot row's entries are staged in LDS, and each
    // row's entries are loaded together, updated and stored together (no load waits on a store)
    const int ncols = j0 + nb - (j + 1);
    if (ncols > 0) {
      for (int c = threadIdx.x; c < ncols; c += blockDim.x) urow[c] = A[(long)(j + 1 + c) * lda + j];
      __syncthreads();
      for (int r = j + 1 + threadIdx.x; r < n; r += blockDim.x) {
        const zc l = A[(long)j * lda + r];
        zc av[LU_NB];
#pragma unroll
        for (int c = 0; c < LU_NB; ++c)
          if (c < ncols) av[c] = A[(long)(j + 1 + c) * lda + r];
#pragma unroll
        for (int c = 0; c < LU_NB; ++c)
          if (c < ncols) A[(long)(j + 1 + c) * lda + r] = av[c] - l * urow[c];
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------------------
// LU panel held in registers (same pivots and arithmetic order as k_lu_panel): 512 lanes as 256
// pairs; pair p owns panel rows p + 256 q (q < Q), and the panel is factored in two halves of 16
// columns, each lane of a pair holding 8 of the half's columns.  Between the halves the first half's
// row interchanges are applied to the second half (one gather of the composite permutation),
// U12 = L11^-1 A12 is solved in registers and A22 -= L21 U12 is applied in registers; at the end the
// second half's interchanges are applied to the first half.  Panels of at most 256 Q rows; Q is the
// smallest of 1, 2, 4 that covers the panel, since every lane issues the work of all Q rows.
// ------------------------------------------------------------------------------------------------
constexpr int LP_THREADS = 512, LP_QMAX = 4, LP_W = 8;

// The value of lane ^ 1 (a DPP quad permutation [1, 0, 3, 2], no LDS round trip).
__device__ inline double pair_swap(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)u, 0xB1, 0xF, 0xF, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), 0xB1, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// The interchanges (t, piv[t]), t0 <= t < t1 (panel-relative rows, applied in order), on the columns
// [cl, cl + ncol) of the panel.  Only the rows they touch change: rowmap (LDS, R entries) tracks
// which original row each position holds, one lane updating it in O(1) per interchange; then every
// (touched row, column) pair reads its source before any of them writes.  Needs
// 2 (t1 - t0) ncol <= blockDim.x.
__device__ void lp_gather_swaps(zc* A, int lda, int j0, const int* piv, int t0, int t1, int cl, int ncol,
                                int* rowmap, int R) {
  for (int i = threadIdx.x; i < R; i += blockDim.x) rowmap[i] = i;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int t = t0; t < t1; ++t) {
      const int a = t, b = piv[t];
      if (a == b) continue;
      const int tmp = rowmap[a];
      rowmap[a] = rowmap[b];
      rowmap[b] = tmp;
    }
  }
  __syncthreads();
  const int nt = t1 - t0, e = threadIdx.x;
  const bool act = ncol > 0 && e < 2 * nt * ncol;
  int row = 0, c = 0;
  zc val = mk(0.0);
  if (act) {
    const int cand = e / ncol;
    c = e % ncol;
    row = cand < nt ? t0 + cand : piv[t0 + cand - nt];  // duplicates move the same value twice
    val = A[(long)(j0 + cl + c) * lda + j0 + rowmap[row]];
  }
  __syncthreads();
  if (act) A[(long)(j0 + cl + c) * lda + j0 + row] = val;
  __syncthreads();
}

template <int Q>
__global__ __launch_bounds__(LP_THREADS) void k_lu_panel_reg(int n, int j0, int nb, MatSet Aset, int lda,
                                                             IntSet pset, int* info) {
  __shared__ double sv[LP_THREADS / 64];
  __shared__ int si[LP_THREADS / 64];
  __shared__ zc prow[16], orow[16];
  __shared__ zc L11[16][17], U12[16][17];
  __shared__ int spiv[32], rowmap[256 * Q];
  const int b = blockIdx.x;
  zc* A = Aset.at(b);
  int* ipiv = pset.at(b);
  const int R = n - j0;
  const int tid = threadIdx.x, p = tid >> 1, h = tid & 1;
  for (int half = 0; half < 2; ++half) {
    const int c0 = half * 16;
    const int hn = min(16, nb - c0);
    if (hn <= 0) break;
    if (half == 1) {
      // the first half's interchanges on this half's columns, then U12 = L11^-1 A12
      lp_gather_swaps(A, lda, j0, spiv, 0, 16, 16, hn, rowmap, R);
      for (int e = tid; e < 16 * 16; e += LP_THREADS) {
        const int r = e & 15, c = e >> 4;
        L11[r][c] = A[(long)(j0 + c) * lda + j0 + r];
        U12[r][c] = (c < hn) ? A[(long)(j0 + 16 + c) * lda + j0 + r] : mk(0.0);
      }
      __syncthreads();
      if (tid < hn) {
        // forward substitution down one column, held in registers
        zc u[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) u[r] = U12[r][tid];
#pragma unroll
        for (int r = 1; r < 16; ++r)
#pragma unroll
          for (int q = 0; q < r; ++q) u[r] = u[r] - L11[r][q] * u[q];
#pragma unroll
        for (int r = 1; r < 16; ++r) U12[r][tid] = u[r];
      }
      __syncthreads();
    }
    zc x[Q][LP_W];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int r = p + 256 * q;
#pragma unroll
      for (int c = 0; c < LP_W; ++c) {
        const int col = c0 + h * LP_W + c;
        x[q][c] = (r < R && col < c0 + hn) ? A[(long)(j0 + col) * lda + j0 + r] : mk(0.0);
      }
    }
    if (half == 1) {
      // A22 -= L21 U12 (rows below 16), rows 0..15 become U12
      // L21 is read from memory two columns at a time, so that the loads of a pair are in flight
      // together
#pragma unroll 1
      for (int k0 = 0; k0 < 16; k0 += 2) {
        zc lk[2][Q];
#pragma unroll
        for (int kk = 0; kk < 2; ++kk)
#pragma unroll
          for (int q = 0; q < Q; ++q) {
            const int r = p + 256 * q;
            lk[kk][q] = (r >= 16 && r < R) ? A[(long)(j0 + k0 + kk) * lda + j0 + r] : mk(0.0);
          }
#pragma unroll
        for (int kk = 0; kk < 2; ++kk) {
          zc u[LP_W];
#pragma unroll
          for (int c = 0; c < LP_W; ++c) u[c] = U12[k0 + kk][h * LP_W + c];
#pragma unroll
          for (int q = 0; q < Q; ++q) {
            const int r = p + 256 * q;
            if (r >= 16 && r < R) {
#pragma unroll
              for (int c = 0; c < LP_W; ++c) x[q][c] = x[q][c] - lk[kk][q] * u[c];
            }
          }
        }
      }
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int r = p + 256 * q;
        if (r < 16) {
#pragma unroll
          for (int c = 0; c < LP_W; ++c) x[q][c] = U12[r][h * LP_W + c];
        }
      }
    }
    // the column loop stays rolled; register indices are kept static by predicated selects
#pragma unroll 1
    for (int jj = 0; jj < hn; ++jj) {
      const int j = c0 + jj;  // panel-relative row and column
      const int hj = jj >> 3, cj = jj & 7;
      double best = -1.0;
      int bidx = R;
      if (h == hj) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
          const int r = p + 256 * q;
          zc v = mk(0.0);
#pragma unroll
          for (int c = 0; c < LP_W; ++c)
            if (c == cj) v = x[q][c];
          if (r >= j && r < R && abs1(v) > best) {
            best = abs1(v);
            bidx = r;
          }
        }
      }
      int pr;
      double pval;
      block_argmax(best, bidx, sv, si, pval, pr);
      if (tid == 0) {
        ipiv[j0 + j] = j0 + pr;
        spiv[j] = pr;
        if (pval == 0.0 && info[b] == 0) info[b] = j0 + j + 1;
      }
      // interchange rows j and pr within this half
      const int pp = pr & 255, qp = pr >> 8;
      if (p == pp) {
#pragma unroll
        for (int q = 0; q < Q; ++q)
          if (q == qp) {
#pragma unroll
            for (int c = 0; c < LP_W; ++c) prow[h * LP_W + c] = x[q][c];
          }
      }
      if (p == j) {
#pragma unroll
        for (int c = 0; c < LP_W; ++c) orow[h * LP_W + c] = x[0][c];
      }
      __syncthreads();
      if (pr != j) {
        if (p == j) {
#pragma unroll
          for (int c = 0; c < LP_W; ++c) x[0][c] = prow[h * LP_W + c];
        }
        if (p == pp) {
#pragma unroll
          for (int q = 0; q < Q; ++q)
            if (q == qp) {
#pragma unroll
              for (int c = 0; c < LP_W; ++c) x[q][c] = orow[h * LP_W + c];
            }
        }
      }
      const zc inv = (pval != 0.0) ? zdiv(mk(1.0), prow[jj]) : mk(0.0);
#pragma unroll
      for (int q = 0; q < Q; ++q) {
        const int r = p + 256 * q;
        const bool act = (r > j && r < R);
        zc v = mk(0.0);
#pragma unroll
        for (int c = 0; c < LP_W; ++c)
          if (c == cj) v = x[q][c];
        zc l = (h == hj) ? v * inv : mk(0.0);
        zc lo;
        lo.re = pair_swap(l.re);
        lo.im = pair_swap(l.im);
        if (h != hj) l = lo;
        if (act) {
#pragma unroll
          for (int c = 0; c < LP_W; ++c) {
            const int col = h * LP_W + c;
            if (col == jj) x[q][c] = l;
            else if (col > jj) x[q][c] = x[q][c] - l * prow[col];
          }
        }
      }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int r = p + 256 * q;
#pragma unroll
      for (int c = 0; c < LP_W; ++c) {
        const int col = c0 + h * LP_W + c;
        if (r < R && col < c0 + hn) A[(long)(j0 + col) * lda + j0 + r] = x[q][c];
      }
    }
    __syncthreads();
  }
  // the second half's interchanges on the first half's columns
  if (nb > 16) lp_gather_swaps(A, lda, j0, spiv, 16, nb, 0, 16, rowmap, R);
}

// ------------------------------------------------------------------------------------------------
// LU panel with one column pair per wavefront (same pivots as k_lu_panel: largest |re| + |im| of the
// interchanged column, ties to the first, as izamax).  Wavefront w holds panel columns 2w and 2w + 1 of
// every row: lane l, register i holds physical row l + 64 i, so once a column pair is factored its
// wavefront has no more work and the later columns' updates are not predicated per lane.  Rows are not
// moved: each carries its logical position (the row order after the interchanges so far); a pivot is
// chosen among the rows whose position is not yet taken, an interchange swaps two positions, and every
// row is stored at its final position at the end.  Per column: the owning wavefront finds the pivot
// (cross-lane reductions, no barrier), scales the column and publishes the multipliers in LDS
// (two buffers, used alternately); one barrier; every wavefront with later columns reads them and
// updates its columns.  NR registers of rows cover panels of up to 64 NR rows; 2 WAVES columns at most.
// The update does not skip the rows above the pivot (their multipliers are zero): with a pivot row that
// is not finite they become NaN too, where LAPACK leaves them (the factors are not finite either way).
// ------------------------------------------------------------------------------------------------
// Lane src's value in every lane (src uniform): scalar reads, no LDS round trip.
__device__ inline zc shfl_z(zc v, int src) {
  const double re = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v.re), src),
                                     __builtin_amdgcn_readlane(__double2loint(v.re), src));
  const double im = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v.im), src),
                                     __builtin_amdgcn_readlane(__double2loint(v.im), src));
  return mk(re, im);
}

// (value, position) maximum over the wavefront, ties to the smaller position; every lane gets the result.
// Rotations by 1, 2, 4, 8 inside each row of 16 lanes (DPP row_ror), then across rows by lane exchanges.
__device__ inline void wave_argmax(double& v, int& idx) {
  // The largest value by DPP alone (rotations inside rows of 16, then row broadcasts 15 and 31: lane 63
  // ends with the maximum), then the smallest position among the lanes that hold it: usually one lane, read
  // directly; with ties, a minimum over those lanes (rotations and lane exchanges).
  auto dpp64 = [](double x, double old, int ctrl, int rows) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
    const unsigned long long o = __builtin_bit_cast(unsigned long long, old);
    unsigned lo, hi;
    switch (ctrl) {
      case 0x121: lo = __builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)u, 0x121, 0xF, 0xF, false);
                  hi = __builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(u >> 32), 0x121, 0xF, 0xF, false);
                  break;
      case 0x122: lo = __builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)u, 0x122, 0xF, 0xF, false);
                  hi = __builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(u >> 32), 0x122, 0xF, 0xF, false);
                  break;
      case 0x124: lo = __builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)u, 0x124, 0xF, 0xF, false);
                  hi = __builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(u >> 32), 0x124, 0xF, 0xF, false);
                  break;
      case 0x128: lo = __builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)u, 0x128, 0xF, 0xF, false);
                  hi = __builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(u >> 32), 0x128, 0xF, 0xF, false);
                  break;
      case 0x142: lo = __builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)u, 0x142, 0xA, 0xF, false);
                  hi = __builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(u >> 32), 0x142, 0xA, 0xF, false);
                  break;
      default:    lo = __builtin_amdgcn_update_dpp((int)(unsigned)o, (int)(unsigned)u, 0x143, 0xC, 0xF, false);
                  hi = __builtin_amdgcn_update_dpp((int)(unsigned)(o >> 32), (int)(unsigned)(u >> 32), 0x143, 0xC, 0xF, false);
                  break;
    }
    (void)rows;
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
  };
  double m = v;
  m = fmax(m, dpp64(m, m, 0x121, 0));
  m = fmax(m, dpp64(m, m, 0x122, 0));
  m = fmax(m, dpp64(m, m, 0x124, 0));
  m = fmax(m, dpp64(m, m, 0x128, 0));
  m = fmax(m, dpp64(m, m, 0x142, 0));
  m = fmax(m, dpp64(m, m, 0x143, 0));
  const double vmax = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(m), 63),
                                       __builtin_amdgcn_readlane(__double2loint(m), 63));
  const unsigned long long tie = __ballot(v == vmax);
  int imin;
  if (__popcll(tie) == 1) {
    imin = __builtin_amdgcn_readlane(idx, __ffsll((long long)tie) - 1);
  } else {
    int k = (v == vmax) ? idx : 0x7fffffff;
#pragma unroll
    for (int sh = 1; sh < 64; sh <<= 1) k = min(k, __shfl_xor(k, sh));
    imin = k;
  }
  v = vmax;
  idx = imin;
}

// 1 / p from the hardware reciprocal estimate and two Newton steps, p scaled by a power of two so that
// |p|^2 neither overflows nor underflows (within a few ulps of the division; p = 0 gives infinities).
__device__ inline zc zrcp_scaled(zc p) {
  const int e = ilogb(fmax(fabs(p.re), fabs(p.im)));
  const double a = ldexp(p.re, -e), b = ldexp(p.im, -e);
  const double d = fma(a, a, b * b);
  double r = __builtin_amdgcn_rcp(d);
  double t = fma(-d, r, 1.0);
  r = fma(r, t, r);
  t = fma(-d, r, 1.0);
  r = fma(r, t, r);
  return mk(ldexp(a * r, -e), ldexp(-b * r, -e));
}

// x - a b with four fused multiply-adds
__device__ inline zc zfms(zc x, zc a, zc b) {
  return mk(fma(-a.re, b.re, fma(a.im, b.im, x.re)), fma(-a.re, b.im, fma(-a.im, b.re, x.im)));
}

// c ? a : b by components (a select of the structure itself would keep x out of registers)
__device__ inline zc zsel(bool c, zc a, zc b) { return mk(c ? a.re : b.re, c ? a.im : b.im); }

// The owning wavefront's part of column jj (its column CO): pivot, multipliers into LDS and into x.
// Row loops are branch-free (selects), so that their loads and stores are issued together.
template <int CO, int NR>
__device__ inline void cw_owner_step(zc (&x)[NR][2], const int (&lp)[NR], int jj, int lane, zc* Lc, int* s_pr,
                                     int* ipiv, int j0, int* info, int b) {
  double best = -1.0;
  int bl = 0x7fffffff;
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    const double a = (lp[i] >= jj) ? abs1(x[i][CO]) : -1.0;
    const int li = (lp[i] >= jj) ? lp[i] : 0x7fffffff;
    const bool take = a > best || (a == best && li < bl);
    best = take ? a : best;
    bl = take ? li : bl;
  }
  wave_argmax(best, bl);
  const int pr = bl;
  zc P = mk(0.0);
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    const unsigned long long m = __ballot(lp[i] == pr);
    if (m) P = shfl_z(x[i][CO], __ffsll((long long)m) - 1);
  }
  const zc rinv = (best != 0.0) ? zrcp_scaled(P) : mk(0.0);
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    const bool below = lp[i] >= jj && lp[i] != pr;
    const zc l = x[i][CO] * rinv;
    Lc[lane + 64 * i] = zsel(below, l, mk(0.0));
    x[i][CO] = zsel(below, l, x[i][CO]);
  }
  if (lane == 0) {
    *s_pr = pr;
    ipiv[j0 + jj] = j0 + pr;
    if (best == 0.0 && info[b] == 0) info[b] = j0 + jj + 1;
  }
}

template <int NR, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_lu_panel_cw(int n, int j0, int nb, MatSet Aset, int lda, IntSet pset,
                                                            int* info) {
  __shared__ zc Lc[2][64 * NR];
  __shared__ int s_pr[2];
  const int b = blockIdx.x;
  const GP A = gview(Aset.at(b));
  int* ipiv = pset.at(b);
  const int R = n - j0;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int ca = 2 * w, cb = 2 * w + 1;  // this wavefront's panel columns
  zc x[NR][2];
  int lp[NR];  // logical position of each register's row; -1 past the panel's end
#pragma unroll
  for (int i = 0; i < NR; ++i) {
    const int r = lane + 64 * i;
    const bool ok = r < R;
    lp[i] = ok ? r : -1;
    x[i][0] = (ok && ca < nb) ? A[(long)(j0 + ca) * lda + j0 + r] : mk(0.0);
    x[i][1] = (ok && cb < nb) ? A[(long)(j0 + cb) * lda + j0 + r] : mk(0.0);
  }
#pragma unroll 1
  for (int jj = 0; jj < nb; ++jj) {
    const int buf = jj & 1;
    if (w == (jj >> 1)) {
      if (jj & 1)
        cw_owner_step<1, NR>(x, lp, jj, lane, Lc[buf], &s_pr[buf], ipiv, j0, info, b);
      else
        cw_owner_step<0, NR>(x, lp, jj, lane, Lc[buf], &s_pr[buf], ipiv, j0, info, b);
    }
    __syncthreads();
    const int pr = s_pr[buf];
    const bool liveA = ca > jj && ca < nb, liveB = cb > jj && cb < nb;
    if (liveA || liveB) {
      // the pivot row's entries in this wavefront's columns, then the rank-1 update below it
      zc u0 = mk(0.0), u1 = mk(0.0);
#pragma unroll
      for (int i = 0; i < NR; ++i) {
        const unsigned long long m = __ballot(lp[i] == pr);
        if (m) {
          const int src = __ffsll((long long)m) - 1;
          u0 = shfl_z(x[i][0], src);
          u1 = shfl_z(x[i][1], src);
        }
      }
      // The multipliers are zero in the rows that are not below the pivot (the owner wrote zeros there), so
      // every row takes x - l u: unchanged in value in those rows while the pivot row is finite.  Multipliers
      // are read four registers at a time.
#pragma unroll
      for (int i0 = 0; i0 < NR; i0 += 4) {
        zc l[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) l[k] = Lc[buf][lane + 64 * (i0 + k)];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
          if (liveA) x[i0 + k][0] = x[i0 + k][0] - l[k] * u0;
          if (liveB) x[i0 + k][1] = x[i0 + k][1] - l[k] * u1;
        }
      }
    }
    // positions jj and pr are interchanged
#pragma unroll
    for (int i = 0; i < NR; ++i) lp[i] = (lp[i] == jj) ? pr : ((lp[i] == pr) ? jj : lp[i]);
  }
#pragma unroll
  for (int i = 0; i < NR; ++i)
    if (lp[i] >= 0) {
      if (ca < nb) A[(long)(j0 + ca) * lda + j0 + lp[i]] = x[i][0];
      if (cb < nb) A[(long)(j0 + cb) * lda + j0 + lp[i]] = x[i][1];
    }
}

// ------------------------------------------------------------------------------------------------
// LU panel with whole rows in registers (same pivots as k_lu_panel).  Lane t holds panel rows t + NT q
// (q < NR), all NB columns, and the column loop is unrolled, so every column index is static: a column's
// update touches only the columns after it, with no per-column predication.  Rows are not moved: each
// carries its logical position (LAPACK's row order after the interchanges so far), a pivot is chosen
// among the rows whose position is not yet taken (largest |re| + |im|, ties to the smaller position, as
// izamax on the interchanged column), an interchange swaps two positions, and every row is stored at its
// final position at the end.  Per column: a wavefront reduction and a reduction over the wavefronts'
// results in LDS (one barrier), then the pivot row's lane publishes the row (second barrier), and every
// row below it takes its multiplier and updates its later columns.  LDS areas alternate between columns,
// so two barriers per column suffice.  The update does not skip the rows above the pivot (their
// multipliers are zero): with a pivot row that is not finite they become NaN too, where LAPACK leaves
// them (the factors are not finite either way).
// ------------------------------------------------------------------------------------------------
template <int NT, int NR, int NB>
__global__ __launch_bounds__(NT) void k_lu_panel_rw(int n, int j0, int nb, MatSet Aset, int lda, IntSet pset,
                                                    int* info) {
  constexpr int NW = NT / 64;
  __shared__ double s_v[2][NW];
  __shared__ int s_i[2][NW];
  __shared__ zc s_row[2][NB];
  const int b = blockIdx.x;
  const GP A = gview(Aset.at(b));
  int* ipiv = pset.at(b);
  const int R = n - j0, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  zc x[NR][NB];
  int lp[NR];  // logical position of each row held; -1 past the panel's end
#pragma unroll
  for (int q = 0; q < NR; ++q) {
    const int r = tid + NT * q;
    lp[q] = r < R ? r : -1;
#pragma unroll
    for (int c = 0; c < NB; ++c) x[q][c] = (r < R && c < nb) ? A[(long)(j0 + c) * lda + j0 + r] : mk(0.0);
  }
#pragma unroll
  for (int jj = 0; jj < NB; ++jj) {
    if (jj < nb) {
    const int buf = jj & 1;
    double a = -1.0;
    int idx = 0x7fffffff;
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      const double aq = (lp[q] >= jj) ? abs1(x[q][jj]) : -1.0;
      const int iq = (lp[q] >= jj) ? lp[q] : 0x7fffffff;
      const bool take = aq > a || (aq == a && iq < idx);
      a = take ? aq : a;
      idx = take ? iq : idx;
    }
    wave_argmax(a, idx);
    if (lane == 0) {
      s_v[buf][w] = a;
      s_i[buf][w] = idx;
    }
    __syncthreads();
    double best = s_v[buf][0];
    int pr = s_i[buf][0];
#pragma unroll
    for (int k = 1; k < NW; ++k) {
      const double ov = s_v[buf][k];
      const int oi = s_i[buf][k];
      const bool take = ov > best || (ov == best && oi < pr);
      best = take ? ov : best;
      pr = take ? oi : pr;
    }
#pragma unroll
    for (int q = 0; q < NR; ++q)
      if (lp[q] == pr) {
#pragma unroll
        for (int c = jj; c < NB; ++c) s_row[buf][c] = x[q][c];
      }
    if (tid == 0) {
      ipiv[j0 + jj] = j0 + pr;
      if (best == 0.0 && info[b] == 0) info[b] = j0 + jj + 1;
    }
    __syncthreads();
    const zc rinv = (best != 0.0) ? zrcp_scaled(s_row[buf][jj]) : mk(0.0);
    zc le[NR];
#pragma unroll
    for (int q = 0; q < NR; ++q) {
      const bool below = lp[q] >= jj && lp[q] != pr;
      const zc l = x[q][jj] * rinv;
      le[q] = zsel(below, l, mk(0.0));
      x[q][jj] = zsel(below, l, x[q][jj]);
      lp[q] = (lp[q] == jj) ? pr : ((lp[q] == pr) ? jj : lp[q]);
    }
#pragma unroll
    for (int c = jj + 1; c < NB; ++c) {
      const zc u = s_row[buf][c];
#pragma unroll
      for (int q = 0; q < NR; ++q) x[q][c] = x[q][c] - le[q] * u;
    }
    }
  }
#pragma unroll
  for (int q = 0; q < NR; ++q)
    if (lp[q] >= 0) {
#pragma unroll
      for (int c = 0; c < NB; ++c)
        if (c < nb) A[(long)(j0 + c) * lda + j0 + lp[q]] = x[q][c];
    }
}

// Apply the panel's swaps to the columns outside [j0, j0+nb) and, for columns right of the panel,
// solve U12 = L11^{-1} A12 (L11 unit lower triangular).  One lane per column.  The interchanges
// (applied in order) are composed once per workgroup into a gather over the rows they touch: rows
// j0..j0+nb-1 and the pivot rows below the panel, at most 2 nb.  Each lane then loads all of them at
// once and stores them permuted, instead of nb dependent load/store round trips.
__global__ __launch_bounds__(64) void k_lu_swap_trsm(int n, int j0, int nb, MatSet Aset, int lda,
                                                     IntSet pset) {
  __shared__ zc L11[LU_NB][LU_NB];
  __shared__ int trow[2 * LU_NB], tsrc[2 * LU_NB];
  __shared__ int s_nt;
  const int b = blockIdx.y;
  const GP A = gview(Aset.at(b));
  const int* ipiv = pset.at(b);
  // the whole LU_NB x LU_NB block, zero outside nb x nb
  for (int e = threadIdx.x; e < LU_NB * LU_NB; e += blockDim.x) {
    const int r = e % LU_NB, c = e / LU_NB;
    L11[r][c] = (r < nb && c < nb) ? (zc)A[(long)(j0 + c) * lda + j0 + r] : mk(0.0);
  }
  {
    // Lane t holds touched row t (my_row) and the original row whose value ends in it (my_src): the
    // panel's rows first, then each pivot row below the panel when first met (pivots are >= their
    // row).  A lookup is a ballot and an interchange two lane reads, so the composition is cheap.
    const int lane = threadIdx.x;
    int my_row = (lane < nb) ? j0 + lane : -1, my_src = my_row;
    const int pv = (lane < nb) ? ipiv[j0 + lane] : 0;
    int nt = nb;
    for (int jj = 0; jj < nb; ++jj) {
      const int p = __builtin_amdgcn_readlane(pv, jj);
      if (p == j0 + jj) continue;
      int tp = p - j0;
      if (p >= j0 + nb) {
        const unsigned long long hit = __ballot(my_row == p);
        if (hit) {
          tp = __ffsll((long long)hit) - 1;
        } else {
          tp = nt++;
          if (lane == tp) my_row = my_src = p;
        }
      }
      const int sj = __builtin_amdgcn_readlane(my_src, jj), sp = __builtin_amdgcn_readlane(my_src, tp);
      if (lane == jj) my_src = sp;
      else if (lane == tp) my_src = sj;
    }
    trow[lane] = my_row;
    tsrc[lane] = my_src;
    if (lane == 0) s_nt = nt;
  }
  __syncthreads();
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  // columns [0, j0) and [j0+nb, n)
  if (c >= j0) c += nb;
  if (c >= n) return;
  const GP col = A + (long)c * lda;
  const int ne = s_nt - nb;  // pivot rows below the panel
  zc x[LU_NB], w[LU_NB];      // the panel's rows, then those pivot rows
  // every load from a valid row (indices clamped, the extra values dropped by selects), all in flight
  // together rather than one conditional load after another
#pragma unroll
  for (int t = 0; t < LU_NB; ++t) {
    x[t] = col[tsrc[min(t, nb - 1)]];
    w[t] = col[tsrc[ne > 0 ? nb + min(t, ne - 1) : 0]];
  }
#pragma unroll
  for (int t = 0; t < LU_NB; ++t) {
    x[t] = zsel(t < nb, x[t], mk(0.0));
    w[t] = zsel(t < ne, w[t], mk(0.0));
  }
#pragma unroll
  for (int t = 0; t < LU_NB; ++t)
    if (t < ne) col[trow[nb + t]] = w[t];
  if (c < j0) {
#pragma unroll
    for (int t = 0; t < LU_NB; ++t)
      if (t < nb) col[j0 + t] = x[t];
    return;
  }
  // one uniform branch per column; inside, every row below is updated (L11 is zero past nb)
#pragma unroll
  for (int r = 0; r < LU_NB; ++r)
    if (r < nb) {
#pragma unroll
      for (int q = r + 1; q < LU_NB; ++q) x[q] = x[q] - L11[q][r] * x[r];
    }
#pragma unroll
  for (int r = 0; r < LU_NB; ++r)
    if (r < nb) col[j0 + r] = x[r];
}

// The same step with four lanes per column (256 threads, 64 columns per workgroup): lane p of a column's
// quad holds the touched rows p, p + 4, ... of the panel (and of the pivot rows below it), so each lane
// issues a quarter of the loads and stores, and in the forward substitution the row that becomes final
// is passed to the other three lanes by a DPP quad broadcast (the row index is a compile-time constant,
// so are all register indices).  L11 sits in LDS column-major and zero outside nb x nb; the rows at or
// above the diagonal take a zero multiplier.
template <int Q>
__device__ inline double quad_bcast(double x) {
  constexpr int ctrl = Q | (Q << 2) | (Q << 4) | (Q << 6);  // quad_perm [Q, Q, Q, Q]
  const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)u, ctrl, 0xF, 0xF, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), ctrl, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
template <int Q>
__device__ inline zc quad_bcast_z(zc v) { return mk(quad_bcast<Q>(v.re), quad_bcast<Q>(v.im)); }

template <int R>
__device__ inline void st4_step(zc (&x)[LU_NB / 4], const zc (*Lc)[LU_NB], int part) {
  const zc xr = quad_bcast_z<R & 3>(x[R >> 2]);
#pragma unroll
  for (int k = R >> 2; k < LU_NB / 4; ++k) {
    const int t = part + 4 * k;
    const zc l = zsel(t > R, Lc[R][t], mk(0.0));
    x[k] = x[k] - l * xr;
  }
  if constexpr (R + 1 < LU_NB) st4_step<R + 1>(x, Lc, part);
}

__global__ __launch_bounds__(256) void k_lu_swap_trsm4(int n, int j0, int nb, MatSet Aset, int lda, IntSet pset) {
  __shared__ zc Lc[LU_NB][LU_NB];  // Lc[c][r] = L11(r, c)
  __shared__ int trow[2 * LU_NB], tsrc[2 * LU_NB];
  __shared__ int s_nt;
  const int b = blockIdx.y;
  const GP A = gview(Aset.at(b));
  const int* ipiv = pset.at(b);
  const int tid = threadIdx.x;
  {
    zc v[LU_NB * LU_NB / 256];  // loads in flight together, addresses clamped into the block
#pragma unroll
    for (int k = 0; k < LU_NB * LU_NB / 256; ++k) {
      const int e = tid + 256 * k, r = e % LU_NB, c = e / LU_NB;
      v[k] = A[(long)(j0 + min(c, nb - 1)) * lda + j0 + min(r, nb - 1)];
    }
#pragma unroll
    for (int k = 0; k < LU_NB * LU_NB / 256; ++k) {
      const int e = tid + 256 * k, r = e % LU_NB, c = e / LU_NB;
      Lc[c][r] = zsel(r < nb && c < nb, v[k], mk(0.0));
    }
  }
  if (tid < 64) {
    // the composition of the panel's interchanges, as in k_lu_swap_trsm
    const int lane = tid;
    int my_row = (lane < nb) ? j0 + lane : -1, my_src = my_row;
    const int pv = (lane < nb) ? ipiv[j0 + lane] : 0;
    int nt = nb;
    for (int jj = 0; jj < nb; ++jj) {
      const int p = __builtin_amdgcn_readlane(pv, jj);
      if (p == j0 + jj) continue;
      int tp = p - j0;
      if (p >= j0 + nb) {
        const unsigned long long hit = __ballot(my_row == p);
        if (hit) {
          tp = __ffsll((long long)hit) - 1;
        } else {
          tp = nt++;
          if (lane == tp) my_row = my_src = p;
        }
      }
      const int sj = __builtin_amdgcn_readlane(my_src, jj), sp = __builtin_amdgcn_readlane(my_src, tp);
      if (lane == jj) my_src = sp;
      else if (lane == tp) my_src = sj;
    }
    trow[lane] = my_row;
    tsrc[lane] = my_src;
    if (lane == 0) s_nt = nt;
  }
  __syncthreads();
  const int part = tid & 3;
  int c = blockIdx.x * 64 + (tid >> 2);
  if (c >= j0) c += nb;  // columns [0, j0) and [j0 + nb, n)
  const bool active = c < n;
  const GP col = A + (long)(active ? c : 0) * lda;
  const int ne = s_nt - nb;  // pivot rows below the panel
  zc x[LU_NB / 4], w[LU_NB / 4];
#pragma unroll
  for (int k = 0; k < LU_NB / 4; ++k) {
    const int t = part + 4 * k;
    x[k] = col[tsrc[min(t, nb - 1)]];
    w[k] = col[tsrc[ne > 0 ? nb + min(t, ne - 1) : 0]];
  }
#pragma unroll
  for (int k = 0; k < LU_NB / 4; ++k) {
    const int t = part + 4 * k;
    x[k] = zsel(t < nb, x[k], mk(0.0));
    w[k] = zsel(t < ne, w[k], mk(0.0));
  }
  // every lane of a column has its sources before any lane stores (the touched rows are shared by the
  // column's four lanes)
  __syncthreads();
  if (!active) return;
#pragma unroll
  for (int k = 0; k < LU_NB / 4; ++k) {
    const int t = part + 4 * k;
    if (t < ne) col[trow[nb + t]] = w[k];
  }
  if (c >= j0) st4_step<0>(x, Lc, part);
#pragma unroll
  for (int k = 0; k < LU_NB / 4; ++k) {
    const int t = part + 4 * k;
    if (t < nb) col[j0 + t] = x[k];
  }
}

// ------------------------------------------------------------------------------------------------
// getrs pieces: row interchanges on B, and small diagonal-block triangular solves.
// ------------------------------------------------------------------------------------------------
// Row interchanges ipiv (zlaswp order: rows 0..n-1 in turn) on the nrhs columns of B, as one gather.
// The interchanges compose to B'(i, :) = B(src(i), :); src is formed in LDS by one lane (n dependent
// steps, once per matrix), then columns pass through LDS in chunks of ch: read along the column and
// written back permuted, both coalesced, instead of n dependent swaps of scattered 16-byte elements
// per column.  Workgroups (matrix, y) take the chunks y, y + gridDim.y, ... of their matrix, each forming
// the map itself; dynamic LDS: ch x n complex, then n ints.
constexpr int SWP_THREADS = 512;
// B <- P B for the interchanges (t, ipiv[t]), t = 0..n-1 in order (zlaswp).  The row of P B at position i
// comes from the row found by undoing the interchanges from the last back to the first.  With zgetrf's
// pivots (ipiv[t] >= t) only t <= i can move it, so each position is traced by one thread over t = i .. 0,
// all positions at once, instead of one thread applying n dependent interchanges; other pivot lists are
// traced over all t.  Then each chunk of columns is staged in LDS and written back permuted.  Pivots
// outside [0, n) are taken as no interchange (they would address rows outside B).
__global__ __launch_bounds__(SWP_THREADS) void k_laswp(int n, int nrhs, int ch, MatSet Bset, int ldb, IntSet pset) {
  extern __shared__ __align__(16) unsigned char swp_lds[];
  zc* buf = reinterpret_cast<zc*>(swp_lds);
  int* src = reinterpret_cast<int*>(swp_lds + sizeof(zc) * (size_t)ch * n);
  int* pv = src + n;
  const int b = blockIdx.x;
  const GP B = gview(Bset.at(b));
  const int* ipiv = pset.at(b);
  int below = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const int p = ipiv[i];
    pv[i] = (p >= 0 && p < n) ? p : i;
    below |= (p >= 0 && p < i) ? 1 : 0;
  }
  const bool general = __syncthreads_or(below);
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    int k = i, t = general ? n - 1 : i;
    for (; t >= 7; t -= 8) {
      int p[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) p[u] = pv[t - u];
#pragma unroll
      for (int u = 0; u < 8; ++u) k = (k == t - u) ? p[u] : ((k == p[u]) ? t - u : k);
    }
    for (; t >= 0; --t) {
      const int q = pv[t];
      k = (k == t) ? q : ((k == q) ? t : k);
    }
    src[i] = k;
  }
  __syncthreads();
  for (int c0 = blockIdx.y * ch; c0 < nrhs; c0 += gridDim.y * ch) {
    const int cn = min(ch, nrhs - c0);
    for (int e = threadIdx.x; e < cn * n; e += blockDim.x) {
      const int c = e / n, r = e % n;
      buf[e] = B[(long)(c0 + c) * ldb + r];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < cn * n; e += blockDim.x) {
      const int c = e / n, r = e % n;
      B[(long)(c0 + c) * ldb + r] = buf[c * n + src[r]];
    }
    __syncthreads();
  }
}

// Solve T x = b for the nb x nb diagonal block starting at j0 (unit lower if lower, else upper),
// for every right-hand-side column; one lane per column.
__global__ __launch_bounds__(64) void k_trsm_block(int j0, int nb, int lower, MatSet LUset, int lda, int nrhs,
                                                   MatSet Bset, int ldb) {
  __shared__ zc T[LU_NB][LU_NB];
  const int b = blockIdx.y;
  const GP LU = gview(LUset.at(b));
  const GP B = gview(Bset.at(b));
  for (int e = threadIdx.x; e < nb * nb; e += blockDim.x) {
    int r = e % nb, c = e / nb;
    T[r][c] = LU[(long)(j0 + c) * lda + j0 + r];
  }
  __syncthreads();
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nrhs) return;
  const GP col = B + (long)c * ldb + j0;
  zc x[LU_NB];
#pragma unroll
  for (int r = 0; r < LU_NB; ++r) x[r] = col[min(r, nb - 1)];  // all loads in flight together
#pragma unroll
  for (int r = 0; r < LU_NB; ++r) x[r] = zsel(r < nb, x[r], mk(0.0));
  if (lower) {
#pragma unroll
    for (int r = 0; r < LU_NB; ++r) {
#pragma unroll
      for (int q = r + 1; q < LU_NB; ++q)
        if (q < nb) x[q] = x[q] - T[q][r] * x[r];
    }
  } else {
#pragma unroll
    for (int r = LU_NB - 1; r >= 0; --r) {
      if (r < nb) {
        x[r] = zdiv(x[r], T[r][r]);
#pragma unroll
        for (int q = 0; q < r; ++q) x[q] = x[q] - T[q][r] * x[r];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < LU_NB; ++r)
    if (r < nb) col[r] = x[r];
}

// Forward and back substitution with the LU factors for one right-hand side (its row interchanges already
// applied), one workgroup per matrix, x held in LDS.  Blocks of 32 rows: one wavefront solves the
// diagonal block with its rows in registers (lane r holds row r), then all lanes update the rows still
// to come with the block's 32 values.  Replaces two launches per block of the many-column path.
constexpr int GS_THREADS = 512, GS_MAXN = 4000;

// sum_q LU(r, j0 + q) xs[j0 + q], q < nb: every load issued before the first is used (addresses clamped to
// the block, the terms past nb dropped by selects), two partial sums.
__device__ inline zc gs_dot(const GP LU, int lda, int j0, int nb, int r, const zc* xs) {
  zc a[LU_NB];
#pragma unroll
  for (int q = 0; q < LU_NB; ++q) a[q] = LU[(long)(j0 + min(q, nb - 1)) * lda + r];
  zc acc0 = mk(0.0), acc1 = mk(0.0);
#pragma unroll
  for (int q = 0; q < LU_NB; q += 2) {
    fma_acc(acc0, zsel(q < nb, a[q], mk(0.0)), xs[j0 + min(q, nb - 1)]);
    fma_acc(acc1, zsel(q + 1 < nb, a[q + 1], mk(0.0)), xs[j0 + min(q + 1, nb - 1)]);
  }
  return acc0 + acc1;
}

__global__ __launch_bounds__(GS_THREADS) void k_getrs1(int n, MatSet LUset, int lda, MatSet Bset) {
  extern __shared__ __align__(16) unsigned char gs_lds[];
  zc* xs = reinterpret_cast<zc*>(gs_lds);
  const GP LU = gview(LUset.at(blockIdx.x));
  const GP B = gview(Bset.at(blockIdx.x));
  const int tid = threadIdx.x, lane = tid & 63;
  for (int r = tid; r < n; r += GS_THREADS) xs[r] = B[r];
  __syncthreads();
  // L (unit lower), blocks from the top
  for (int j0 = 0; j0 < n; j0 += LU_NB) {
    const int nb = min(LU_NB, n - j0);
    if (tid < 64) {
      // loads from rows inside the block for every lane (selected afterwards), so that they are all in
      // flight together instead of one conditional load after another
      const int rl = j0 + min(lane, nb - 1);
      zc lrow[LU_NB];
#pragma unroll
      for (int q = 0; q < LU_NB; ++q) lrow[q] = LU[(long)(j0 + min(q, nb - 1)) * lda + rl];
#pragma unroll
      for (int q = 0; q < LU_NB; ++q) lrow[q] = zsel(q < lane && lane < nb, lrow[q], mk(0.0));
      zc xr = (lane < nb) ? xs[j0 + lane] : mk(0.0);
#pragma unroll
      for (int q = 0; q < LU_NB - 1; ++q)
        if (q < nb - 1) xr = xr - lrow[q] * shfl_z(xr, q);
      if (lane < nb) xs[j0 + lane] = xr;
    }
    __syncthreads();
    for (int r = j0 + nb + tid; r < n; r += GS_THREADS) xs[r] = xs[r] - gs_dot(LU, lda, j0, nb, r, xs);
    __syncthreads();
  }
  // U (upper, with its diagonal), blocks from the bottom
  for (int j1 = n; j1 > 0; j1 -= LU_NB) {
    const int j0 = max(0, j1 - LU_NB), nb = j1 - j0;
    if (tid < 64) {
      const int rl = j0 + min(lane, nb - 1);
      zc urow[LU_NB];
#pragma unroll
      for (int q = 0; q < LU_NB; ++q) urow[q] = LU[(long)(j0 + min(q, nb - 1)) * lda + rl];
#pragma unroll
      for (int q = 0; q < LU_NB; ++q) urow[q] = zsel(lane <= q && q < nb, urow[q], mk(0.0));
      zc xr = (lane < nb) ? xs[j0 + lane] : mk(0.0);
#pragma unroll
      for (int q = LU_NB - 1; q >= 0; --q) {
        if (q < nb) {
          if (lane == q) xr = zdiv(xr, urow[q]);
          const zc xq = shfl_z(xr, q);
          if (lane < q) xr = xr - urow[q] * xq;
        }
      }
      if (lane < nb) xs[j0 + lane] = xr;
    }
    __syncthreads();
    for (int r = tid; r < j0; r += GS_THREADS) {
      xs[r] = xs[r] - gs_dot(LU, lda, j0, nb, r, xs);
    }
    __syncthreads();
  }
  for (int r = tid; r < n; r += GS_THREADS) B[r] = xs[r];
}

__global__ void k_fill_int(int n, int* p, int v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

// ------------------------------------------------------------------------------------------------
// Inverses of the 64 x 64 diagonal blocks of an LU factorisation, for the solves that apply them as
// matrix products.  Workgroup (j, f, b): block j of matrix b, f = 0 the unit lower factor, f = 1 the
// upper one.  The block sits in LDS (column-major, 65 KiB, two workgroups per compute unit); a block cut
// by the matrix's end is completed with the identity, whose inverse it keeps.  Lane (c, p) of the 256
// (c = tid / 4, p = tid % 4) holds rows p, p + 4, ..., p + 60 of column c of the inverse in registers:
// column elimination with the final row passed inside the quad by DPP, no barrier.  The output is
// column-major, leading dimension 64: block j of L^-1 at j 64^2, of U^-1 at (nblk + j) 64^2 of the
// matrix's area.
// ------------------------------------------------------------------------------------------------
constexpr int TI_NB = 64, TI_LD = TI_NB + 1;
constexpr size_t TI_LDS = sizeof(zc) * (TI_NB * TI_LD + TI_NB);

template <int Q>
__device__ inline void ti_lower(zc (&x)[16], const zc* Tc, int part) {
  // row Q is final: the rows below take x_r -= L[r][Q] x_Q (a zero multiplier at and above the diagonal)
  const zc xq = quad_bcast_z<Q & 3>(x[Q >> 2]);
  const zc* tq = Tc + Q * TI_LD;
#pragma unroll
  for (int k = Q >> 2; k < 16; ++k) {
    const int r = part + 4 * k;
    x[k] = x[k] - zsel(r > Q, tq[r], mk(0.0)) * xq;
  }
  __builtin_amdgcn_sched_barrier(0);  // one row's LDS reads at a time (registers)
  if constexpr (Q + 2 < TI_NB) ti_lower<Q + 1>(x, Tc, part);
}

template <int Q>
__device__ inline void ti_upper(zc (&x)[16], const zc* Tc, const zc* rdiag, int part) {
  // x_Q times 1 / U[Q][Q] in the lane that owns row Q, then the rows above take x_r -= U[r][Q] x_Q
  x[Q >> 2] = zsel(part == (Q & 3), x[Q >> 2] * rdiag[Q], x[Q >> 2]);
  const zc xq = quad_bcast_z<Q & 3>(x[Q >> 2]);
  const zc* tq = Tc + Q * TI_LD;
#pragma unroll
  for (int k = 0; k <= (Q >> 2); ++k) {
    const int r = part + 4 * k;
    x[k] = x[k] - zsel(r < Q, tq[r], mk(0.0)) * xq;
  }
  __builtin_amdgcn_sched_barrier(0);
  if constexpr (Q > 0) ti_upper<Q - 1>(x, Tc, rdiag, part);
}

__global__ __launch_bounds__(256) void k_trtri_diag(int n, MatSet LUset, int lda, MatSet Dset) {
  extern __shared__ __align__(16) unsigned char ti_lds[];
  zc* Tc = reinterpret_cast<zc*>(ti_lds);  // Tc[c * TI_LD + r] = block(r, c)
  zc* rdiag = Tc + TI_NB * TI_LD;          // 1 / U[q][q]
  const int j = blockIdx.x, upper = blockIdx.y, b = blockIdx.z;
  const int nblk = (n + TI_NB - 1) / TI_NB;
  const int j0 = j * TI_NB, nbj = min(TI_NB, n - j0);
  const GP LU = gview(LUset.at(b));
  const GP D = gview(Dset.at(b)) + (long)(upper * nblk + j) * TI_NB * TI_NB;
  {
    constexpr int PER = TI_NB * TI_NB / 256;
    zc v[PER];  // all loads in flight together (addresses clamped into the block, selected below)
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int e = threadIdx.x + 256 * k, r = e & (TI_NB - 1), c = e / TI_NB;
      v[k] = LU[(long)(j0 + min(c, nbj - 1)) * lda + j0 + min(r, nbj - 1)];
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int e = threadIdx.x + 256 * k, r = e & (TI_NB - 1), c = e / TI_NB;
      Tc[c * TI_LD + r] = zsel(r < nbj && c < nbj, v[k], mk(r == c ? 1.0 : 0.0));
    }
  }
  __syncthreads();
  if (upper && threadIdx.x < TI_NB) rdiag[threadIdx.x] = zdiv(mk(1.0), Tc[threadIdx.x * TI_LD + threadIdx.x]);
  __syncthreads();
  // lane (c, p) holds rows p, p + 4, ..., p + 60 of column c of the inverse; a quad is one column, so the
  // elimination needs no barrier (rows pass between the four lanes by DPP)
  const int c = threadIdx.x >> 2, part = threadIdx.x & 3;
  zc x[16];
#pragma unroll
  for (int k = 0; k < 16; ++k) x[k] = mk(part + 4 * k == c ? 1.0 : 0.0);
  if (upper)
    ti_upper<TI_NB - 1>(x, Tc, rdiag, part);
  else
    ti_lower<0>(x, Tc, part);
#pragma unroll
  for (int k = 0; k < 16; ++k) D[(long)c * TI_NB + part + 4 * k] = x[k];
}

}  // namespace

void zgemm_set(Op opa, Op opb, int m, int n, int k, zc alpha, MatSet A, int lda, MatSet B, int ldb, zc beta,
               MatSet C, int ldc, int batch, hipStream_t s) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  CWX_CHECK(batch <= 65535, "zgemm: batch too large");
  CWX_CHECK(k > 0, "zgemm: k must be positive");
  if (n <= GV_MAXN && opa == OpN && opb == OpN) {
    hipLaunchKernelGGL(k_zgemv, dim3(cdiv(m, GV_ROWS), batch), dim3(64 * GV_WAVES), 0, s, m, n, k, alpha, A, lda, B,
                       ldb, beta, C, ldc);
    CWX_HIP(hipGetLastError());
    return;
  }
  // the matrix-core kernel with 64x64 tiles; the others stay selectable (cwx_set_param "zgemm_kernel")
  const int kind = (int)tune_lookup("zgemm_kernel", m, n, 1);
  if (kind == 1) {
    hipLaunchKernelGGL((k_zgemm_mfma<2, 2>), dim3(cdiv(m, 64), cdiv(n, 64), batch), dim3(256), 0, s, opa, opb, m, n,
                       k, alpha, A, lda, B, ldb, beta, C, ldc);
  } else if (kind == 2) {
    hipLaunchKernelGGL((k_zgemm_mfma<4, 2>), dim3(cdiv(m, 128), cdiv(n, 64), batch), dim3(256), 0, s, opa, opb, m,
                       n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  } else if (kind == 3) {
    hipLaunchKernelGGL((k_zgemm_mfma<2, 4>), dim3(cdiv(m, 64), cdiv(n, 128), batch), dim3(256), 0, s, opa, opb, m,
                       n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  } else {
    dim3 grid(cdiv(m, GM), cdiv(n, GN), batch);
    hipLaunchKernelGGL(k_zgemm, grid, dim3(256), 0, s, opa, opb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc);
  }
  CWX_HIP(hipGetLastError());
}

void zgemm_split32_set(Op opa, Op opb, int m, int n, int k, zc alpha, MatSet A, int lda, MatSet B, int ldb, zc beta,
                       MatSet C, int ldc, int batch, bool split, hipStream_t s) {
  if (m <= 0 || n <= 0 || batch <= 0) return;
  CWX_CHECK(batch <= 65535, "zgemm: batch too large");
  CWX_CHECK(k > 0, "zgemm: k must be positive");
  dim3 grid(cdiv(m, GM), cdiv(n, GN), batch);
  if (split)
    hipLaunchKernelGGL(k_zgemm_split32<true>, grid, dim3(256), 0, s, opa, opb, m, n, k, alpha, A, lda, B, ldb, beta, C,
                       ldc);
  else
    hipLaunchKernelGGL(k_zgemm_split32<false>, grid, dim3(256), 0, s, opa, opb, m, n, k, alpha, A, lda, B, ldb, beta,
                       C, ldc);
  CWX_HIP(hipGetLastError());
}

void zgetrf_set(int n, MatSet A, int lda, IntSet ipiv, int* info, int batch, hipStream_t s) {
  if (n <= 0 || batch <= 0) return;
  CWX_CHECK(batch <= 65535, "zgetrf: batch too large");
  // a fill kernel rather than hipMemsetAsync keeps the sequence capturable in a HIP graph
  hipLaunchKernelGGL(k_fill_int, dim3(cdiv(batch, 256)), dim3(256), 0, s, batch, info, 0);
  // panels with whole rows per lane (2, the default), one column pair per wavefront (1) or rows per lane
  // pair (0); the first two keep 32 columns up to 512 rows and 16 columns up to 1024 rows
  const int kind = (int)tune_lookup("lu_panel", n, 0, 2);
  for (int j0 = 0, nb = 0; j0 < n; j0 += nb) {
    const int R = n - j0;
    nb = std::min((kind != 0 && R > 512 && R <= 1024) ? 16 : LU_NB, R);
    if (kind == 2 && R <= 256)
      hipLaunchKernelGGL((k_lu_panel_rw<256, 1, 32>), dim3(batch), dim3(256), 0, s, n, j0, nb, A, lda, ipiv, info);
    else if (kind == 2 && R <= 512)
      hipLaunchKernelGGL((k_lu_panel_rw<512, 1, 32>), dim3(batch), dim3(512), 0, s, n, j0, nb, A, lda, ipiv, info);
    else if (kind == 2 && R <= 1024)
      hipLaunchKernelGGL((k_lu_panel_rw<512, 2, 16>), dim3(batch), dim3(512), 0, s, n, j0, nb, A, lda, ipiv, info);
    else if (kind == 1 && R <= 256)
      hipLaunchKernelGGL((k_lu_panel_cw<4, 16>), dim3(batch), dim3(1024), 0, s, n, j0, nb, A, lda, ipiv, info);
    else if (kind == 1 && R <= 512)
      hipLaunchKernelGGL((k_lu_panel_cw<8, 16>), dim3(batch), dim3(1024), 0, s, n, j0, nb, A, lda, ipiv, info);
    else if (kind == 1 && R <= 1024)
      hipLaunchKernelGGL((k_lu_panel_cw<16, 8>), dim3(batch), dim3(512), 0, s, n, j0, nb, A, lda, ipiv, info);
    else if (n - j0 <= 256)
      hipLaunchKernelGGL(k_lu_panel_reg<1>, dim3(batch), dim3(LP_THREADS), 0, s, n, j0, nb, A, lda, ipiv, info);
    else if (n - j0 <= 512)
      hipLaunchKernelGGL(k_lu_panel_reg<2>, dim3(batch), dim3(LP_THREADS), 0, s, n, j0, nb, A, lda, ipiv, info);
    else if (n - j0 <= 256 * LP_QMAX)
      hipLaunchKernelGGL(k_lu_panel_reg<4>, dim3(batch), dim3(LP_THREADS), 0, s, n, j0, nb, A, lda, ipiv, info);
    else
      hipLaunchKernelGGL(k_lu_panel, dim3(batch), dim3(LU_THREADS), 0, s, n, j0, nb, A, lda, ipiv, info);
    const int others = n - nb;
    if (others > 0 && kind != 0)
      hipLaunchKernelGGL(k_lu_swap_trsm4, dim3(cdiv(others, 64), batch), dim3(256), 0, s, n, j0, nb, A, lda, ipiv);
    else if (others > 0)
      hipLaunchKernelGGL(k_lu_swap_trsm, dim3(cdiv(others, 64), batch), dim3(64), 0, s, n, j0, nb, A, lda, ipiv);
    const int rest = n - j0 - nb;
    if (rest > 0)
      zgemm_set(OpN, OpN, rest, rest, nb, mk(-1.0), A.shifted((long)j0 * lda + j0 + nb), lda,
                A.shifted((long)(j0 + nb) * lda + j0), lda, mk(1.0), A.shifted((long)(j0 + nb) * lda + j0 + nb), lda,
                batch, s);
  }
  CWX_HIP(hipGetLastError());
}

// The row interchanges of zgetrf on B (LAPACK zlaswp, forward).
void laswp_set(int n, int nrhs, MatSet B, int ldb, IntSet ipiv, int batch, hipStream_t s) {
  // columns staged per chunk: as many as fit 64 KiB of LDS beside the n-entry source map (the device's
  // whole LDS allowance when one column does not fit in 64 KiB)
  const size_t need = (2 * sizeof(int) + sizeof(zc)) * (size_t)n;
  const size_t budget = need <= 64 * 1024 ? 64 * 1024 : lds_limit();
  CWX_CHECK(need <= budget, "zgetrs: n too large for the LDS row map");
  const size_t avail = budget - 2 * sizeof(int) * (size_t)n;
  const int ch = (int)std::max<size_t>(1, std::min<size_t>((size_t)nrhs, avail / (sizeof(zc) * (size_t)n)));
  const size_t lds = sizeof(zc) * (size_t)ch * n + 2 * sizeof(int) * (size_t)n;
  // enough workgroups to spread the column chunks over the device (about 2048 in all)
  const int chunks = (nrhs + ch - 1) / ch;
  const int gy = std::max(1, std::min({chunks, 65535, 2048 / batch}));
  hipLaunchKernelGGL(k_laswp, dim3(batch, gy), dim3(SWP_THREADS), lds, s, n, nrhs, ch, B, ldb, ipiv);
}

int trtri_diag_blocks(int n) { return (n + TI_NB - 1) / TI_NB; }

long trtri_diag_size(int n) { return 2L * trtri_diag_blocks(n) * TI_NB * TI_NB; }

void ztrtri_diag_set(int n, MatSet LU, int lda, MatSet Dinv, int batch, hipStream_t s) {
  if (n <= 0 || batch <= 0) return;
  CWX_CHECK(batch <= 65535, "ztrtri_diag: batch too large");
  hipLaunchKernelGGL(k_trtri_diag, dim3(trtri_diag_blocks(n), 2, batch), dim3(256), TI_LDS, s, n, LU, lda, Dinv);
  CWX_HIP(hipGetLastError());
}

void zgetrs_inv_set(int n, int nrhs, MatSet LU, int lda, IntSet ipiv, MatSet Dinv, MatSet B, int ldb, MatSet Y,
                    int ldy, int batch, hipStream_t s) {
  if (n <= 0 || nrhs <= 0 || batch <= 0) return;
  CWX_CHECK(batch <= 65535, "zgetrs: batch too large");
  laswp_set(n, nrhs, B, ldb, ipiv, batch, s);
  const int nblk = trtri_diag_blocks(n);
  const long bs = (long)TI_NB * TI_NB;
  // forward, Y = L^-1 B: Y_j = (L_jj)^-1 B_j, then the rows below take B -= L(:, j) Y_j
  for (int j = 0; j < nblk; ++j) {
    const int j0 = j * TI_NB, nb = std::min(TI_NB, n - j0);
    zgemm_set(OpN, OpN, nb, nrhs, nb, mk(1.0), Dinv.shifted(j * bs), TI_NB, B.shifted(j0), ldb, mk(0.0),
              Y.shifted(j0), ldy, batch, s);
    const int rest = n - j0 - nb;
    if (rest > 0)
      zgemm_set(OpN, OpN, rest, nrhs, nb, mk(-1.0), LU.shifted((long)j0 * lda + j0 + nb), lda, Y.shifted(j0), ldy,
                mk(1.0), B.shifted(j0 + nb), ldb, batch, s);
  }
  // backward, X = U^-1 Y into B: X_j = (U_jj)^-1 Y_j, then the rows above take Y -= U(:, j) X_j
  for (int j = nblk - 1; j >= 0; --j) {
    const int j0 = j * TI_NB, nb = std::min(TI_NB, n - j0);
    zgemm_set(OpN, OpN, nb, nrhs, nb, mk(1.0), Dinv.shifted((nblk + j) * bs), TI_NB, Y.shifted(j0), ldy, mk(0.0),
              B.shifted(j0), ldb, batch, s);
    if (j0 > 0)
      zgemm_set(OpN, OpN, j0, nrhs, nb, mk(-1.0), LU.shifted((long)j0 * lda), lda, B.shifted(j0), ldb, mk(1.0), Y,
                ldy, batch, s);
  }
}

void zgetrs_set(int n, int nrhs, MatSet LU, int lda, IntSet ipiv, MatSet B, int ldb, int batch, hipStream_t s) {
  if (n <= 0 || nrhs <= 0 || batch <= 0) return;
  CWX_CHECK(batch <= 65535, "zgetrs: batch too large");
  laswp_set(n, nrhs, B, ldb, ipiv, batch, s);
  if (nrhs == 1 && n <= GS_MAXN) {
    hipLaunchKernelGGL(k_getrs1, dim3(batch), dim3(GS_THREADS), sizeof(zc) * (size_t)n, s, n, LU, lda, B);
    CWX_HIP(hipGetLastError());
    return;
  }
  // Blocks of TS_NB rows: inside a block, diagonal LU_NB blocks and updates of the block's own rows;
  // the rows below (above) are updated once per block, which bounds the passes over B.
  constexpr int TS_NB = 128;
  // forward: L (unit lower)
  for (int b0 = 0; b0 < n; b0 += TS_NB) {
    const int bn = std::min(TS_NB, n - b0);
    for (int j0 = b0; j0 < b0 + bn; j0 += LU_NB) {
      const int nb = std::min(LU_NB, b0 + bn - j0);
      hipLaunchKernelGGL(k_trsm_block, dim3(cdiv(nrhs, 64), batch), dim3(64), 0, s, j0, nb, 1, LU, lda, nrhs, B,
                         ldb);
      const int rest = b0 + bn - j0 - nb;
      if (rest > 0)
        zgemm_set(OpN, OpN, rest, nrhs, nb, mk(-1.0), LU.shifted((long)j0 * lda + j0 + nb), lda, B.shifted(j0), ldb,
                  mk(1.0), B.shifted(j0 + nb), ldb, batch, s);
    }
    const int below = n - b0 - bn;
    if (below > 0)
      zgemm_set(OpN, OpN, below, nrhs, bn, mk(-1.0), LU.shifted((long)b0 * lda + b0 + bn), lda, B.shifted(b0), ldb,
                mk(1.0), B.shifted(b0 + bn), ldb, batch, s);
  }
  // backward: U
  const int nbig = (n + TS_NB - 1) / TS_NB;
  for (int bi = nbig - 1; bi >= 0; --bi) {
    const int b0 = bi * TS_NB, bn = std::min(TS_NB, n - b0);
    const int nsub = (bn + LU_NB - 1) / LU_NB;
    for (int si = nsub - 1; si >= 0; --si) {
      const int j0 = b0 + si * LU_NB;
      const int nb = std::min(LU_NB, b0 + bn - j0);
      hipLaunchKernelGGL(k_trsm_block, dim3(cdiv(nrhs, 64), batch), dim3(64), 0, s, j0, nb, 0, LU, lda, nrhs, B,
                         ldb);
      if (j0 > b0)
        zgemm_set(OpN, OpN, j0 - b0, nrhs, nb, mk(-1.0), LU.shifted((long)j0 * lda + b0), lda, B.shifted(j0), ldb,
                  mk(1.0), B.shifted(b0), ldb, batch, s);
    }
    if (b0 > 0)
      zgemm_set(OpN, OpN, b0, nrhs, bn, mk(-1.0), LU.shifted((long)b0 * lda), lda, B.shifted(b0), ldb, mk(1.0), B, ldb,
                batch, s);
  }
  CWX_HIP(hipGetLastError());
}

}  // namespace cwx
