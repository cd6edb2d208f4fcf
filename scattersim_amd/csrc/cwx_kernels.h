// Internal launch interface of the cwx HIP kernels (not part of the C ABI).
#pragma once

#include "cwx_common.h"

namespace cwx {

// ---------------------------------------------------------------------------------------------
// field evaluation
// ---------------------------------------------------------------------------------------------
struct GridSpec {
  float x0, y0, z0;   // coordinates of grid point (0, 0, 0)
  float dx, dy, dz;   // steps; coordinate i is x0 + dx * (float)i as in gpuEvaluator.cu:78-79
  int nx, ny, nz;     // output is [nz][ny][nx][3]
};

struct FieldTable {
  const void* waves = nullptr;  // nwaves records {E0[3], k[3]}, complex<float> (48 B) or complex<double> (96 B)
  int precision = 0;            // 4 or 8
  long nwaves = 0;
  const float* z_layers = nullptr;  // device, layers
  const int* wb = nullptr;          // device, layers + 1
  const int* we = nullptr;
  int layers = 0;
};

void launch_field_direct(const FieldTable& t, const GridSpec& g, int tol_mode, int out_f64, void* out,
                         hipStream_t s);

// Separable evaluation of a solved structure (one homogeneous region per z): see cwx_fieldsep.hip.
struct SepRegion {
  // E(x,y,z) = sum_{a<nkx, b<nky} X[a](x) Y[b](y) C[a][b](z), C = A (nkx*nky*3 x nmodes) . e^{i kz z}
  int nmodes;
};

// ---------------------------------------------------------------------------------------------
// dense complex linear algebra (column-major, batched with element strides)
// ---------------------------------------------------------------------------------------------
enum Op { OpN = 0, OpT = 1, OpC = 2 };

// A set of `batch` matrices (or vectors): member b starts at base + stride * b, or at ptr[b] when a
// device array of pointers is given; `off` (elements) is added in both cases, so sub-blocks of every
// member form a set of their own.
template <typename T>
struct Set {
  T* base = nullptr;
  long stride = 0;
  T* const* ptr = nullptr;
  long off = 0;
  __host__ __device__ T* at(int b) const { return (ptr ? ptr[b] : base + stride * b) + off; }
  Set shifted(long o) const {
    Set s = *this;
    s.off += o;
    return s;
  }
};
using MatSet = Set<zc>;
using IntSet = Set<int>;
template <typename T>
inline Set<T> strided(const T* p, long stride) {
  Set<T> s;
  s.base = const_cast<T*>(p);
  s.stride = stride;
  return s;
}
template <typename T>
inline Set<T> listed(T* const* d_ptrs) {
  Set<T> s;
  s.ptr = d_ptrs;
  return s;
}

// C[b] = alpha * op(A[b]) * op(B[b]) + beta * C[b]
void zgemm_set(Op opa, Op opb, int m, int n, int k, zc alpha, MatSet A, int lda, MatSet B, int ldb, zc beta,
               MatSet C, int ldc, int batch, hipStream_t s);
inline void zgemm_batched(Op opa, Op opb, int m, int n, int k, zc alpha, const zc* A, int lda, long sA,
                          const zc* B, int ldb, long sB, zc beta, zc* C, int ldc, long sC, int batch,
                          hipStream_t s) {
  zgemm_set(opa, opb, m, n, k, alpha, strided(A, sA), lda, strided(B, sB), ldb, beta, strided(C, sC), ldc, batch,
            s);
}

// The same product on the FP32 matrix cores: split operands (hi * hi + hi * lo + lo * hi) or, with split
// false, the operands rounded to FP32; FP32 partial sums of 16 terms added in FP64.  About 1e-7 relative.
void zgemm_split32_set(Op opa, Op opb, int m, int n, int k, zc alpha, MatSet A, int lda, MatSet B, int ldb, zc beta,
                       MatSet C, int ldc, int batch, bool split, hipStream_t s);

// LU with partial pivoting, in place (LAPACK zgetrf contract, 0-based pivots).  info[b] = 0 or
// the 1-based index of the first exactly-zero pivot.
void zgetrf_set(int n, MatSet A, int lda, IntSet ipiv, int* info, int batch, hipStream_t s);
inline void zgetrf_batched(int n, zc* A, int lda, long sA, int* ipiv, long sP, int* info, int batch,
                           hipStream_t s) {
  zgetrf_set(n, strided(A, sA), lda, strided(ipiv, sP), info, batch, s);
}

// Inverses of the 64 x 64 diagonal blocks of the factors of zgetrf: trtri_diag_size(n) elements per matrix
// (block j of L^-1, then block j of U^-1 after all of L's, each 64 x 64 column-major).
int trtri_diag_blocks(int n);
long trtri_diag_size(int n);
void ztrtri_diag_set(int n, MatSet LU, int lda, MatSet Dinv, int batch, hipStream_t s);
// Solve A X = B as zgetrs does, with the diagonal blocks applied as products by their inverses (Dinv from
// ztrtri_diag_set) on the matrix cores; Y (n x nrhs per matrix, leading dimension ldy) is scratch.  The
// rounding differs from zgetrs in the last bits.
void zgetrs_inv_set(int n, int nrhs, MatSet LU, int lda, IntSet ipiv, MatSet Dinv, MatSet B, int ldb, MatSet Y,
                    int ldy, int batch, hipStream_t s);

// Solve A X = B with the factors of zgetrf (LAPACK zgetrs, no transpose).
void zgetrs_set(int n, int nrhs, MatSet LU, int lda, IntSet ipiv, MatSet B, int ldb, int batch, hipStream_t s);
inline void zgetrs_batched(int n, int nrhs, const zc* LU, int lda, long sA, const int* ipiv, long sP, zc* B,
                           int ldb, long sB, int batch, hipStream_t s) {
  zgetrs_set(n, nrhs, strided(LU, sA), lda, strided(ipiv, sP), strided(B, sB), ldb, batch, s);
}

// ---------------------------------------------------------------------------------------------
// non-symmetric eigensolver (LAPACK zgeev('N','V') contract)
// ---------------------------------------------------------------------------------------------
struct EigWork {
  DevBuf<zc> Z;      // n*n*batch Schur vectors
  DevBuf<zc> X;      // n*n*batch triangular eigenvectors of T
  DevBuf<zc> tau;    // n*batch
  DevBuf<int> info;  // batch
  DevBuf<zc> Y, T, W; // blocked Hessenberg reduction: panel Y, T of every panel, GEMM scratch
  DevBuf<double> ring; // QR: the current window's rotations of every matrix, read back as scalars
  DevBuf<int> team_ctl;       // QR teams (leader + helpers per matrix): counters, rotation slots, slot
  DevBuf<double> team_slots;  // descriptors
  DevBuf<int> team_desc;
  void ensure(int n, int batch);
  void ensure_team(int batch);
};


// A is destroyed (overwritten by the Schur form).  W receives eigenvalues, V the unit 2-norm right
// eigenvectors with their largest component real.  info (device, batch) is 0 or the LAPACK-style
// failure index.
// teams = false keeps one workgroup per matrix in the QR iteration (no cooperative launch, so that the
// launch can be captured in a graph).
void zgeev_batched(int n, zc* A, int lda, long sA, zc* W, long sW, zc* V, int ldv, long sV, int batch,
                   EigWork& work, int* info, hipStream_t s, bool teams = true);

}  // namespace cwx
