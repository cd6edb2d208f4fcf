// Batched non-symmetric complex eigensolver with the LAPACK zgeev('N','V') contract.
//
// Replaces MKL_eigensolve (third_Lapack.h:38-45, LAPACKE_zgeev with jobvl='N', jobvr='V'), which the
// reference calls once per layer (ScatterVolume.cpp:255-262).  Steps, one workgroup per matrix:
//   1. Householder reduction to upper Hessenberg form          (zgehd2 algorithm)
//   2. explicit formation of the unitary factor Q                (zung2r algorithm)
//   3. Schur form T and Schur vectors Z = Q * (QR transformations): on active blocks of 48 rows and more,
//      aggressive early deflation on a trailing 16 x 16 window (zlaqr3 algorithm) alternating with
//      sweeps of 16 chained single-shift bulges whose shifts are the window's undeflated eigenvalues
//      (zlaqr5-style multishift); below that, single-shift QR with Wilkinson / exceptional shifts and
//      the Ahues-Kahan deflation test (zlahqr algorithm).  Bulges are chased through windows of H held
//      in LDS; each window's rotations are then applied to the rest of H and to Z from registers, one
//      matrix line per lane.
//   4. eigenvectors of T by back-substitution (ztrevc algorithm), one wavefront per eigenvector
//   5. V = Z X on the GEMM kernel, then zgeev's normalisation (unit 2-norm, largest component real).
#include "cwx_common.h"
#include "cwx_kernels.h"
#include "cwx_tune.h"

namespace cwx {

namespace {

// Threads of the QR workgroup: 512 when the matrices do not fill the device (one workgroup per CU);
// 256 when there are at least two matrices per CU, so that two workgroups share a CU and the
// latency-bound chase of one overlaps the arithmetic-bound deferred products of the other.
constexpr int EIG_THREADS = 512, EIG_THREADS_SMALL = 256;
constexpr int QR_W = 32;  // rotations per window
constexpr double kUlp = 2.220446049250313e-16;      // dlamch('P')
constexpr double kSafmin = 2.2250738585072014e-308;  // dlamch('S')

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// block-wide sum of a complex value; every lane receives the result
__device__ inline zc block_sum(zc v, zc* red) {
  v.re = wave_sum(v.re);
  v.im = wave_sum(v.im);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) red[wid] = v;
  __syncthreads();
  zc s = mk(0.0);
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
  __syncthreads();
  return s;
}

// ---- hand-offs between the workgroups of one matrix (Hessenberg panels and QR iteration) ----
// Counters are polled relaxed (a coherent load, no cache invalidation); the acquire fence follows once
// the awaited value has been seen.  One thread of a workgroup performs each release and acquire, next
// to a workgroup barrier: the cache operations they stand for act on the caches the whole workgroup
// shares.  Every wait ends, on its counter, on an abort flag or after `limit` wall-clock ticks.
constexpr int CTL_INTS = 16, CTL_DONE = 8, TEAM_MAX = CTL_INTS - CTL_DONE + 1;

__device__ inline int ld_relaxed(const int* p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }
__device__ inline void st_release(int* p, int v) { __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT); }

// Every wave's stores have reached the L2 (vmcnt(0)); after the barrier one release store by thread 0
// (which writes the L2 back) publishes them all.
__device__ inline void wg_release_store(int* p, int v) {
  __builtin_amdgcn_s_waitcnt(0x0F70);
  __syncthreads();
  if (threadIdx.x == 0) st_release(p, v);
}

// Thread 0: waits until every counter p[0..cnt-1] is at least `target`; false on an abort (*abort set,
// by another workgroup or here on the time limit).
__device__ inline bool counters_reach(const int* p, int cnt, int target, int* abort, unsigned long long limit) {
  const unsigned long long t0 = wall_clock64();
  for (int h = 0; h < cnt; ++h) {
    while (ld_relaxed(p + h) < target) {
      if (ld_relaxed(abort) || wall_clock64() - t0 > limit) {
        st_release(abort, 1);
        return false;
      }
      __builtin_amdgcn_s_sleep(2);
    }
  }
  return true;
}

// The whole workgroup waits for counters_reach and then sees what was published; false on an abort.
__device__ inline bool wg_wait(const int* p, int cnt, int target, int* abort, unsigned long long limit, int* s_flag) {
  if (threadIdx.x == 0) {
    *s_flag = counters_reach(p, cnt, target, abort, limit) ? 1 : 0;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
  return __builtin_amdgcn_readfirstlane(*s_flag) != 0;
}

// A matrix's workgroups when several share it: G per matrix (G = 1: one workgroup), counters per matrix
// (ctl[0] published steps, ctl[2] abort, ctl[CTL_DONE + h - 1] steps finished by helper h).  The
// workgroups of one matrix share blockIdx.x mod 8 (one XCD under round-robin dispatch).
struct Team {
  int G = 1;
  int batch = 0;
  int* ctl = nullptr;
  unsigned long long limit = 0;
};

__device__ inline void team_role(const Team& t, int& b, int& role) {
  b = blockIdx.x;
  role = 0;
  if (t.G > 1) {
    const int k = blockIdx.x / 8;
    b = (int)(blockIdx.x % 8) + 8 * (k / t.G);
    role = k % t.G;
  }
}

// 1/x for normal x: v_rcp_f64 refined by two Newton steps (within an ulp of the rounded quotient).
__device__ inline double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}
// sqrt(x) for normal positive x: v_rsq_f64 and Goldschmidt's iteration for the root and its half
// reciprocal, then one correction of the root (within an ulp of the rounded root).
__device__ inline double sqrt_nr(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  double g = x * y, h = 0.5 * y;
  const double r = fma(-g, h, 0.5);
  g = fma(g, r, g);
  h = fma(h, r, h);
  return fma(fma(-g, g, x), h, g);
}

__device__ inline double dlapy3(double x, double y, double z) {
  double w = fmax(fabs(x), fmax(fabs(y), fabs(z)));
  if (w == 0.0) return fabs(x) + fabs(y) + fabs(z);
  return w * sqrt((x / w) * (x / w) + (y / w) * (y / w) + (z / w) * (z / w));
}

// zlarfg for n = 2: beta = -sign(Re alpha) |(alpha, x)|, tau = (beta - alpha) / beta,
// v2 = x / (alpha - beta).  When |(alpha, x)|^2 is comfortably inside the double range it is formed
// directly, so the chain is one square root and two independent reciprocals (|alpha - beta| >= |beta|,
// no cancellation), each from the hardware estimate and Newton steps; otherwise the scaled forms of
// zlarfg (dlapy3 and Smith's division) are used.
// Returns false when x = 0 and alpha is real (tau = 0, the identity).
__device__ inline bool reflector2(zc& alpha, zc x, zc& tau, zc& v2) {
  if (x.re == 0.0 && x.im == 0.0 && alpha.im == 0.0) return false;
  const double s2 = fma(alpha.re, alpha.re, fma(alpha.im, alpha.im, norm2(x)));
  double beta;
  if (s2 > 1e-280 && s2 < 1e280) {
    beta = -copysign(sqrt_nr(s2), alpha.re);
    const double rb = rcp_nr(beta);
    const zc d = mk(alpha.re - beta, alpha.im);
    const double rd = rcp_nr(norm2(d));
    tau = mk((beta - alpha.re) * rb, -alpha.im * rb);
    v2 = x * mk(d.re * rd, -d.im * rd);
  } else {
    beta = -copysign(dlapy3(alpha.re, alpha.im, absz(x)), alpha.re);
    tau = mk((beta - alpha.re) / beta, -alpha.im / beta);
    v2 = x * zdiv(mk(1.0), alpha - mk(beta));
  }
  alpha = mk(beta);
  return true;
}

// ------------------------------------------------------------------------------------------------
// Level-2 building blocks for one workgroup operating on a column-major block in global memory.
// Loads are unrolled with independent accumulators so that several are in flight per lane: the
// block lives in L2 and latency, not bandwidth, limits a single workgroup.
// ------------------------------------------------------------------------------------------------
constexpr int HH_THREADS = 1024;

// w[r] = scale * sum_c A(r, c) v[c], r < rows, c < cols.  Lanes run down the rows; when the rows do
// not fill the workgroup, G groups of lanes take interleaved columns and their partial sums are
// combined through `part` (LDS, at least blockDim.x entries).  Ends with a barrier.
__device__ inline void gemv_rows(const zc* __restrict__ A, int lda, int rows, int cols, const zc* v, zc scale,
                                 zc* w, zc* part) {
  const int R = (rows + 63) & ~63;
  const int G = rows >= (int)blockDim.x ? 1 : min(8, (int)blockDim.x / R);
  const int g = threadIdx.x / R;
  for (int r = threadIdx.x % R; r < rows; r += (G == 1 ? blockDim.x : R)) {
    if (g >= G) break;
    zc a0 = mk(0.0), a1 = mk(0.0), a2 = mk(0.0), a3 = mk(0.0);
    const zc* p = A + r;
    int c = g;
    for (; c + 3 * G < cols; c += 4 * G) {
      const zc x0 = p[(long)c * lda], x1 = p[(long)(c + G) * lda];
      const zc x2 = p[(long)(c + 2 * G) * lda], x3 = p[(long)(c + 3 * G) * lda];
      fma_acc(a0, x0, v[c]);
      fma_acc(a1, x1, v[c + G]);
      fma_acc(a2, x2, v[c + 2 * G]);
      fma_acc(a3, x3, v[c + 3 * G]);
    }
    for (; c < cols; c += G) fma_acc(a0, p[(long)c * lda], v[c]);
    const zc acc = (a0 + a1) + (a2 + a3);
    if (G == 1) w[r] = acc * scale;
    else part[g * rows + r] = acc;
  }
  __syncthreads();
  if (G > 1) {
    for (int r = threadIdx.x; r < rows; r += blockDim.x) {
      zc acc = part[r];
      for (int q = 1; q < G; ++q) acc += part[q * rows + r];
      w[r] = acc * scale;
    }
    __syncthreads();
  }
}

// out[r] = sum_c A(r, c) v[c] for r0 <= r < r1, c < cols, in a summation order fixed per row whatever
// r0 and r1 are: GV_GROUPS lanes take interleaved columns of a row (four accumulators each), and their
// partial sums are added in group order.  So the rows of one product may be split over workgroups and
// give the same bits as one workgroup.  part: LDS, blockDim.x entries.  Ends with a barrier.
constexpr int GV_GROUPS = 8;
__device__ inline void gemv_rows_det(const zc* __restrict__ A, int lda, int r0, int r1, int cols, const zc* v,
                                     zc* out, zc* part) {
  const int rpass = (int)blockDim.x / GV_GROUPS;
  const int g = threadIdx.x / rpass, rr = threadIdx.x % rpass;
  for (int base = r0; base < r1; base += rpass) {
    const int r = base + rr;
    zc a0 = mk(0.0), a1 = mk(0.0), a2 = mk(0.0), a3 = mk(0.0);
    if (r < r1) {
      const zc* p = A + r;
      int c = g;
      for (; c + 3 * GV_GROUPS < cols; c += 4 * GV_GROUPS) {
        const zc x0 = p[(long)c * lda], x1 = p[(long)(c + GV_GROUPS) * lda];
        const zc x2 = p[(long)(c + 2 * GV_GROUPS) * lda], x3 = p[(long)(c + 3 * GV_GROUPS) * lda];
        fma_acc(a0, x0, v[c]);
        fma_acc(a1, x1, v[c + GV_GROUPS]);
        fma_acc(a2, x2, v[c + 2 * GV_GROUPS]);
        fma_acc(a3, x3, v[c + 3 * GV_GROUPS]);
      }
      for (; c < cols; c += GV_GROUPS) fma_acc(a0, p[(long)c * lda], v[c]);
    }
    part[g * rpass + rr] = (a0 + a1) + (a2 + a3);
    __syncthreads();
    if ((int)threadIdx.x < rpass && base + (int)threadIdx.x < r1) {
      zc acc = part[threadIdx.x];
      for (int q = 1; q < GV_GROUPS; ++q) acc += part[q * rpass + threadIdx.x];
      out[base + threadIdx.x] = acc;
    }
    __syncthreads();
  }
}

// u[c] = scale * sum_r conj(v[r]) A(r, c) (wavefront per column, lanes down the rows)
__device__ inline void gemv_cols_conj(const zc* __restrict__ A, int lda, int rows, int cols, const zc* v,
                                      zc scale, zc* u) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int c = wid; c < cols; c += nw) {
    const zc* col = A + (long)c * lda;
    zc a0 = mk(0.0), a1 = mk(0.0);
    int r = lane;
    for (; r + 64 < rows; r += 128) {
      const zc x0 = col[r], x1 = col[r + 64];
      fma_acc(a0, conj(v[r]), x0);
      fma_acc(a1, conj(v[r + 64]), x1);
    }
    if (r < rows) fma_acc(a0, conj(v[r]), col[r]);
    zc acc = a0 + a1;
    acc.re = wave_sum(acc.re);
    acc.im = wave_sum(acc.im);
    if (lane == 0) u[c] = acc * scale;
  }
}

// A(r, c) -= x[r] * y[c] for r < rows, c < cols; y_conj conjugates y (flat, four elements per lane)
__device__ inline void rank1_sub(zc* __restrict__ A, int lda, int rows, int cols, const zc* x, const zc* y,
                                 bool y_conj) {
  const long total = (long)rows * cols;
  for (long base = (long)threadIdx.x * 4; base < total; base += (long)blockDim.x * 4) {
    zc val[4];
    long off[4];
    int rr[4], cc[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const long e = base + u;
      rr[u] = (int)(e % rows);
      cc[u] = (int)(e / rows);
      off[u] = (long)cc[u] * lda + rr[u];
      if (e < total) val[u] = A[off[u]];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      if (base + u < total) {
        const zc yc = y_conj ? conj(y[cc[u]]) : y[cc[u]];
        A[off[u]] = val[u] - x[rr[u]] * yc;
      }
    }
  }
}

// ------------------------------------------------------------------------------------------------
// 1. Hessenberg reduction (zgehd2).  A(j+1:n, j) -> beta e1, v stored below, tau[j].
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(HH_THREADS) void k_gehrd(int n, int j0, zc* A, int lda, long sA, zc* tau) {
  extern __shared__ zc smem[];
  zc* v = smem;           // n
  zc* w = smem + n;       // n
  zc* part = smem + 2 * n;  // blockDim.x
  __shared__ zc red[HH_THREADS / 64];
  const int b = blockIdx.x;
  A += sA * b;
  tau += (long)n * b;
  const int tid = threadIdx.x;
  for (int j = j0; j < n - 2; ++j) {
    const int m = n - j - 1;  // length of x = A(j+1:n, j)
    zc* x = A + (long)j * lda + j + 1;
    double ss = 0.0;
    for (int r = 1 + tid; r < m; r += blockDim.x) ss += norm2(x[r]);
    zc tot = block_sum(mk(ss), red);
    const double xnorm = sqrt(tot.re);
    const zc alpha = x[0];
    zc t = mk(0.0);
    double beta = 0.0;
    if (!(xnorm == 0.0 && alpha.im == 0.0)) {
      beta = -copysign(dlapy3(alpha.re, alpha.im, xnorm), alpha.re);
      t = mk((beta - alpha.re) / beta, -alpha.im / beta);
    }
    const bool trivial = (t.re == 0.0 && t.im == 0.0);
    if (!trivial) {
      const zc scal = zdiv(mk(1.0), alpha - mk(beta));
      for (int r = 1 + tid; r < m; r += blockDim.x) x[r] = x[r] * scal;
    }
    __syncthreads();
    for (int r = tid; r < m; r += blockDim.x) v[r] = (r == 0) ? mk(1.0) : x[r];
    if (tid == 0) {
      tau[j] = t;
      if (!trivial) x[0] = mk(beta);
    }
    __syncthreads();
    if (!trivial) {
      // right: A(0:n, j+1:n) -= (t A v) v^H
      gemv_rows(A + (long)(j + 1) * lda, lda, n, m, v, t, w, part);
      rank1_sub(A + (long)(j + 1) * lda, lda, n, m, w, v, true);
      __syncthreads();
      // left: A(j+1:n, j+1:n) -= v (conj(t) v^H A)
      gemv_cols_conj(A + (long)(j + 1) * lda + j + 1, lda, m, m, v, conj(t), w);
      __syncthreads();
      rank1_sub(A + (long)(j + 1) * lda + j + 1, lda, m, m, v, w, false);
      __syncthreads();
    }
  }
}

// ------------------------------------------------------------------------------------------------
// 2. Q = H(0) H(1) ... H(n-3) applied to the identity, from the last reflector backwards (zung2r on
//    the trailing block): this kernel sets Z = I and applies H(n-3) .. H(jlo); the reflectors of the
//    blocked panels follow as block reflectors (zlarfb) by batched GEMMs.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(HH_THREADS) void k_unghr(int n, int jlo, const zc* A, int lda, long sA, const zc* tau,
                                                      zc* Z, int ldz, long sZ) {
  extern __shared__ zc smem[];
  zc* v = smem;      // n
  zc* w = smem + n;  // n
  const int b = blockIdx.x;
  A += sA * b;
  tau += (long)n * b;
  Z += sZ * b;
  const int tid = threadIdx.x;
  for (long e = tid; e < (long)n * n; e += blockDim.x) {
    int r = (int)(e % n), c = (int)(e / n);
    Z[(long)c * ldz + r] = mk(r == c ? 1.0 : 0.0);
  }
  __syncthreads();
  for (int j = n - 3; j >= jlo; --j) {
    const zc t = tau[j];
    if (t.re == 0.0 && t.im == 0.0) continue;
    const int m = n - j - 1;
    for (int r = tid; r < m; r += blockDim.x) v[r] = (r == 0) ? mk(1.0) : A[(long)j * lda + j + 1 + r];
    __syncthreads();
    // Z(j+1:n, j+1:n) = (I - t v v^H) Z(j+1:n, j+1:n)
    zc* Zb = Z + (long)(j + 1) * ldz + j + 1;
    gemv_cols_conj(Zb, ldz, m, m, v, t, w);
    __syncthreads();
    rank1_sub(Zb, ldz, m, m, v, w, false);
    __syncthreads();
  }
}

// The Householder vectors below the subdiagonal are cleared so that A holds H.
__global__ void k_fill_i32(long n, int* a) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) a[i] = 0;
}

__global__ void k_clear_below_sub(int n, zc* A, int lda, long sA) {
  A += sA * blockIdx.y;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)n * n; e += (long)gridDim.x * blockDim.x) {
    const int r = (int)(e % n), c = (int)(e / n);
    if (r > c + 1) A[(long)c * lda + r] = mk(0.0);
  }
}

// ------------------------------------------------------------------------------------------------
// 1b. Blocked Hessenberg reduction (zgehrd with zlahr2 panels).  Panel i0 .. i0+nb-1 is reduced by
//     one workgroup per matrix, reading the trailing matrix once per column (the product with the
//     new Householder vector, corrected by the panel's earlier reflectors through Y and T); the rest
//     of the matrix is updated afterwards by batched GEMMs.  V is kept dense (unit lower trapezoidal,
//     explicit zeros and ones), V(r, j) for matrix row i0+1+r; T is upper triangular with
//     H(i0) ... H(i0+nb-1) = I - V T V^H; Y(i0+1:n, :) = A(i0+1:n, i0+1:n) V T.
// ------------------------------------------------------------------------------------------------
constexpr int HB = 32;     // panel width
constexpr int HB_NX = 64;  // the last columns are reduced unblocked

// Rows [r0, r1) of the trailing-matrix product of workgroup `role` in a team of G.
__device__ inline void lahr2_rows(int m, int G, int role, int& r0, int& r1) {
  r0 = (int)((long)m * role / G);
  r1 = (int)((long)m * (role + 1) / G);
}

// Helper `role` of a panel team: for every column of the panel, once the leader has published the
// column's Householder vector (V(:, j), counter i0 + j + 1), its rows of y = A(i0+1:n, c+1:n) v go to
// Y(:, j), where the leader picks them up.
__device__ void lahr2_help(int n, int i0, int nb, const zc* A, int lda, const zc* V, int ldv, zc* Y, int ldy,
                           const Team& t, int b, int role, zc* smem) {
  __shared__ int s_ok;
  zc* vv = smem + n;
  zc* part = smem + 2 * n;
  int* ctl = t.ctl + (long)b * CTL_INTS;
  const int m = n - i0 - 1;
  int r0, r1;
  lahr2_rows(m, t.G, role, r0, r1);
  for (int j = 0; j < nb; ++j) {
    const int c = i0 + j;
    if (!wg_wait(ctl, 1, c + 1, ctl + 2, t.limit, &s_ok)) return;
    for (int r = j + threadIdx.x; r < m; r += blockDim.x) vv[r - j] = V[(long)j * ldv + r];
    __syncthreads();
    gemv_rows_det(A + (long)(c + 1) * lda + i0 + 1, lda, r0, r1, m - j, vv, Y + (long)j * ldy, part);
    wg_release_store(ctl + CTL_DONE + role - 1, c + 1);
  }
}

__global__ __launch_bounds__(HH_THREADS) void k_lahr2(int n, int i0, int nb, zc* A, int lda, long sA, zc* V,
                                                      int ldv, long sV, zc* Y, int ldy, long sY, zc* T, long sT,
                                                      zc* tau, Team team) {
  extern __shared__ zc smem[];
  zc* bcol = smem;           // m: the current column, rows i0+1..n-1; then the new column of Y
  zc* vv = smem + n;         // m: the new Householder vector over columns c+1..n-1
  zc* part = smem + 2 * n;   // blockDim.x
  __shared__ zc red[HH_THREADS / 64];
  __shared__ zc w[HB], t2[HB];
  __shared__ int s_ok;
  int bi, role;
  team_role(team, bi, role);
  if (bi >= (team.G > 1 ? team.batch : (int)gridDim.x)) return;
  if (role > 0) {
    lahr2_help(n, i0, nb, A + sA * bi, lda, V + sV * bi, ldv, Y + sY * bi, ldy, team, bi, role, smem);
    return;
  }
  int* const ctl = team.G > 1 ? team.ctl + (long)bi * CTL_INTS : nullptr;
  A += sA * bi;
  V += sV * bi;
  Y += sY * bi;
  T += sT * bi;
  tau += (long)n * bi;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wid = tid >> 6, nw = nt >> 6;
  const int m = n - i0 - 1;
  auto Ap = [&](int r, int c) -> zc& { return A[(long)c * lda + r]; };
  auto Vp = [&](int r, int j) -> zc& { return V[(long)j * ldv + r]; };
  auto Yp = [&](int r, int j) -> zc& { return Y[(long)j * ldy + r]; };
  auto Tp = [&](int r, int j) -> zc& { return T[(long)j * HB + r]; };
  for (int j = 0; j < nb; ++j) {
    const int c = i0 + j;
    for (int r = tid; r < m; r += nt) bcol[r] = Ap(i0 + 1 + r, c);
    __syncthreads();
    if (j > 0) {
      // b -= Y(:, 0:j) conj(V(j-1, 0:j)): the right update of this column by the earlier reflectors
      // loads four columns at a time (addresses clamped below j), so that they are in flight together;
      // the terms are added in the same order as one at a time
      for (int r = tid; r < m; r += nt) {
        zc acc = bcol[r];
        for (int q0 = 0; q0 < j; q0 += 4) {
          zc yv[4], vc[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) {
            const int q = min(q0 + k, j - 1);
            yv[k] = Yp(r, q);
            vc[k] = Vp(j - 1, q);
          }
#pragma unroll
          for (int k = 0; k < 4; ++k)
            if (q0 + k < j) acc = acc - yv[k] * conj(vc[k]);
        }
        bcol[r] = acc;
      }
      __syncthreads();
      // b -= V T^H V^H b: the left update by the earlier reflectors
      for (int q = wid; q < j; q += nw) {
        zc acc = mk(0.0);
        int r = q + lane;
        for (; r + 192 < m; r += 256) {  // four loads in flight, the same order of terms
          const zc v0 = Vp(r, q), v1 = Vp(r + 64, q), v2 = Vp(r + 128, q), v3 = Vp(r + 192, q);
          fma_acc(acc, conj(v0), bcol[r]);
          fma_acc(acc, conj(v1), bcol[r + 64]);
          fma_acc(acc, conj(v2), bcol[r + 128]);
          fma_acc(acc, conj(v3), bcol[r + 192]);
        }
        for (; r < m; r += 64) fma_acc(acc, conj(Vp(r, q)), bcol[r]);
        acc.re = wave_sum(acc.re);
        acc.im = wave_sum(acc.im);
        if (lane == 0) w[q] = acc;
      }
      __syncthreads();
      if (tid < j) {
        zc acc = mk(0.0);
        for (int q = 0; q <= tid; ++q) acc = acc + conj(Tp(q, tid)) * w[q];
        t2[tid] = acc;
      }
      __syncthreads();
      for (int r = tid; r < m; r += nt) {
        zc acc = bcol[r];
        for (int q0 = 0; q0 < j; q0 += 8) {
          zc vr8[8];
#pragma unroll
          for (int k = 0; k < 8; ++k) vr8[k] = Vp(r, min(q0 + k, j - 1));
#pragma unroll
          for (int k = 0; k < 8; ++k)
            if (q0 + k < j) acc = acc - vr8[k] * t2[q0 + k];
        }
        bcol[r] = acc;
      }
      __syncthreads();
    }
    // reflector annihilating b(j+1 .. m-1) (zlarfg): alpha = b(j) is the subdiagonal entry A(c+1, c)
    double ss = 0.0;
    for (int r = j + 1 + tid; r < m; r += nt) ss += norm2(bcol[r]);
    const zc tot = block_sum(mk(ss), red);
    const double xnorm = sqrt(tot.re);
    const zc alpha = bcol[j];
    zc tj = mk(0.0);
    double beta = alpha.re;
    const bool trivial = (xnorm == 0.0 && alpha.im == 0.0);
    zc scal = mk(1.0);
    if (!trivial) {
      beta = -copysign(dlapy3(alpha.re, alpha.im, xnorm), alpha.re);
      tj = mk((beta - alpha.re) / beta, -alpha.im / beta);
      scal = zdiv(mk(1.0), alpha - mk(beta));
    }
    for (int r = tid; r < m; r += nt) {
      zc vr;
      if (r < j) {
        vr = mk(0.0);
        Ap(i0 + 1 + r, c) = bcol[r];
      } else if (r == j) {
        vr = mk(1.0);
        Ap(i0 + 1 + r, c) = trivial ? alpha : mk(beta);
      } else {
        vr = trivial ? bcol[r] : bcol[r] * scal;
        Ap(i0 + 1 + r, c) = vr;
      }
      Vp(r, j) = vr;
      if (r >= j) vv[r - j] = vr;
    }
    if (tid == 0) tau[c] = tj;
    __syncthreads();
    // y = A(i0+1:n, c+1:n) v  (the trailing matrix, not yet touched by this panel's right updates); in a
    // team the helpers take their rows once V(:, j) is published, and leave them in Y(:, j)
    int r0 = 0, r1 = m;
    if (ctl) {
      wg_release_store(ctl, c + 1);
      lahr2_rows(m, team.G, 0, r0, r1);
    }
    gemv_rows_det(A + (long)(c + 1) * lda + i0 + 1, lda, r0, r1, m - j, vv, bcol, part);
    if (ctl) {
      if (!wg_wait(ctl + CTL_DONE, team.G - 1, c + 1, ctl + 2, team.limit, &s_ok)) {
        // a helper did not answer: mark the panel failed (tau = NaN poisons the matrix, and the
        // eigensolver's non-finite check reports it)
        if (tid == 0) tau[c] = mk(__builtin_nan(""), 0.0);
        return;
      }
      for (int r = r1 + tid; r < m; r += nt) bcol[r] = Yp(r, j);
      __syncthreads();
    }
    // t = V(:, 0:j)^H v
    for (int q = wid; q < j; q += nw) {
      zc acc = mk(0.0);
      int r = j + lane;
      for (; r + 192 < m; r += 256) {  // four loads in flight, the same order of terms
        const zc v0 = Vp(r, q), v1 = Vp(r + 64, q), v2 = Vp(r + 128, q), v3 = Vp(r + 192, q);
        fma_acc(acc, conj(v0), vv[r - j]);
        fma_acc(acc, conj(v1), vv[r + 64 - j]);
        fma_acc(acc, conj(v2), vv[r + 128 - j]);
        fma_acc(acc, conj(v3), vv[r + 192 - j]);
      }
      for (; r < m; r += 64) fma_acc(acc, conj(Vp(r, q)), vv[r - j]);
      acc.re = wave_sum(acc.re);
      acc.im = wave_sum(acc.im);
      if (lane == 0) w[q] = acc;
    }
    __syncthreads();
    // Y(:, j) = tau (y - Y(:, 0:j) t)
    for (int r = tid; r < m; r += nt) {
      zc acc = bcol[r];
      for (int q0 = 0; q0 < j; q0 += 8) {
        zc yr8[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) yr8[k] = Yp(r, min(q0 + k, j - 1));
#pragma unroll
        for (int k = 0; k < 8; ++k)
          if (q0 + k < j) acc = acc - yr8[k] * w[q0 + k];
      }
      Yp(r, j) = tj * acc;
    }
    // T(0:j, j) = -tau T(0:j, 0:j) t,  T(j, j) = tau
    if (tid <= j) {
      zc acc = mk(0.0);
      if (tid < j) {
        for (int q = tid; q < j; ++q) acc = acc + Tp(tid, q) * w[q];
        acc = -(tj * acc);
      } else {
        acc = tj;
      }
      Tp(tid, j) = acc;
    }
    if (tid >= j + 1 && tid < HB) Tp(tid, j) = mk(0.0);
    __syncthreads();
  }
}

// ------------------------------------------------------------------------------------------------
// 3. Schur form by single-shift QR (zlahqr), windowed bulge chase.
// ------------------------------------------------------------------------------------------------
struct Rot {
  zc t1;
  zc v2;
  double t2;
};

// 2-element Householder (zlarfg with n = 2): returns (t1, v2, t2 = Re(t1 v2)) and beta in alpha.
__device__ inline Rot make_reflector(zc& alpha, zc x) {
  Rot g;
  if (!reflector2(alpha, x, g.t1, g.v2)) {
    g.t1 = mk(0.0);
    g.v2 = x;
    g.t2 = 0.0;
    return g;
  }
  g.t2 = (g.t1 * g.v2).re;
  return g;
}

// The rotations are written out as fused multiply-adds (12 of them and 2 additions each).
__device__ inline void rot_left(const Rot& g, zc& a, zc& b) {
  // rows k, k+1:  sum = conj(t1) a + t2 b; a -= sum; b -= sum v2
  zc s;
  s.re = fma(g.t1.re, a.re, fma(g.t1.im, a.im, g.t2 * b.re));
  s.im = fma(g.t1.re, a.im, fma(-g.t1.im, a.re, g.t2 * b.im));
  a.re -= s.re;
  a.im -= s.im;
  b.re = fma(-s.re, g.v2.re, fma(s.im, g.v2.im, b.re));
  b.im = fma(-s.re, g.v2.im, fma(-s.im, g.v2.re, b.im));
}
__device__ inline void rot_right(const Rot& g, zc& a, zc& b) {
  // cols k, k+1:  sum = t1 a + t2 b; a -= sum; b -= sum conj(v2)
  zc s;
  s.re = fma(g.t1.re, a.re, fma(-g.t1.im, a.im, g.t2 * b.re));
  s.im = fma(g.t1.re, a.im, fma(g.t1.im, a.re, g.t2 * b.im));
  a.re -= s.re;
  a.im -= s.im;
  b.re = fma(-s.re, g.v2.re, fma(-s.im, g.v2.im, b.re));
  b.im = fma(-s.im, g.v2.re, fma(s.re, g.v2.im, b.im));
}

__device__ inline bool subdiag_negligible(const zc* H, int ldh, int k, int ilo, int ihi, double smlnum) {
  auto h = [&](int r, int c) { return H[(long)c * ldh + r]; };
  const zc hkk1 = h(k, k - 1);
  if (abs1(hkk1) <= smlnum) return true;
  double tst = abs1(h(k - 1, k - 1)) + abs1(h(k, k));
  if (tst == 0.0) {
    if (k - 2 >= ilo) tst += fabs(h(k - 1, k - 2).re);
    if (k + 1 <= ihi) tst += fabs(h(k + 1, k).re);
  }
  if (fabs(hkk1.re) <= kUlp * tst) {
    const double ab = fmax(abs1(hkk1), abs1(h(k - 1, k)));
    const double ba = fmin(abs1(hkk1), abs1(h(k - 1, k)));
    const double aa = fmax(abs1(h(k, k)), abs1(h(k - 1, k - 1) - h(k, k)));
    const double bb = fmin(abs1(h(k, k)), abs1(h(k - 1, k - 1) - h(k, k)));
    const double s = aa + ab;
    if (ba * (ab / s) <= fmax(smlnum, kUlp * (bb * (aa / s)))) return true;
  }
  return false;
}

__device__ inline bool two_small(const zc* H, int ldh, int m, zc t) {
  auto h = [&](int r, int c) { return H[(long)c * ldh + r]; };
  const zc h11 = h(m, m), h22 = h(m + 1, m + 1);
  zc h11s = h11 - t;
  double h21 = h(m + 1, m).re;
  const double s = abs1(h11s) + fabs(h21);
  h11s = h11s * (1.0 / s);
  h21 = h21 / s;
  const double h10 = h(m, m - 1).re;
  return fabs(h10) * fabs(h21) <= kUlp * (abs1(h11s) * (abs1(h11) + abs1(h22)));
}

// ------------------------------------------------------------------------------------------------
// Multishift sweeps.  A chain of up to MS_NB single-shift bulges, introduced two rows apart at the top
// of the active block and chased down in lockstep: bulges two rows apart touch disjoint row pairs
// and column pairs, so within a step every left transformation commutes with every other left one,
// right ones likewise, and a left with a right; each bulge reads only entries no trailing bulge has
// written yet.  Each step is therefore three phases (reflectors; left products plus the bulge
// column; right products), and the result is the sequence of MS_NB single-shift sweeps with shifts
// fixed before the first one.  The chain moves through windows of MS_S steps held in LDS; the
// window's rotations are then applied to the rest of H and to Z from registers.
// ------------------------------------------------------------------------------------------------
constexpr int MS_NB = 16;                               // bulges (shifts) per sweep
constexpr int MS_S = 16;                                // chase steps per window
constexpr int MS_LEN = MS_S + 2 * (MS_NB - 1) + 2;      // rows spanned by one window (48)
constexpr int MS_NMIN = 48;                             // smallest active block swept this way
// Leading dimensions of the LDS window blocks: a column stride of 4 (mod 32) dwords puts the 16-byte
// accesses of 8 consecutive lanes walking along a row into distinct bank groups.
constexpr int MS_LD = MS_LEN + 1;                       // 49 complex = 196 dwords
constexpr int QR_LD = QR_W + 3;                         // 35 complex = 140 dwords

struct Rot2 {
  zc t1, v2, t2;  // reflector I - t1 v v^H with v = (1, v2); t2 = t1 v2
};

// 2-element Householder (zlarfg, n = 2) for a general complex second entry.
__device__ inline Rot2 make_reflector2(zc& alpha, zc x) {
  Rot2 g;
  if (!reflector2(alpha, x, g.t1, g.v2)) {
    g.t1 = g.v2 = g.t2 = mk(0.0);
    return g;
  }
  g.t2 = g.t1 * g.v2;
  return g;
}
__device__ inline void rot2_left(const Rot2& g, zc& a, zc& b) {
  // s = conj(t1) a + conj(t2) b; a -= s; b -= s v2
  zc s;
  s.re = fma(g.t1.re, a.re, fma(g.t1.im, a.im, fma(g.t2.re, b.re, g.t2.im * b.im)));
  s.im = fma(g.t1.re, a.im, fma(-g.t1.im, a.re, fma(g.t2.re, b.im, -(g.t2.im * b.re))));
  a.re -= s.re;
  a.im -= s.im;
  b.re = fma(-s.re, g.v2.re, fma(s.im, g.v2.im, b.re));
  b.im = fma(-s.re, g.v2.im, fma(-s.im, g.v2.re, b.im));
}
__device__ inline void rot2_right(const Rot2& g, zc& a, zc& b) {
  // s = t1 a + t2 b; a -= s; b -= s conj(v2)
  zc s;
  s.re = fma(g.t1.re, a.re, fma(-g.t1.im, a.im, fma(g.t2.re, b.re, -(g.t2.im * b.im))));
  s.im = fma(g.t1.re, a.im, fma(g.t1.im, a.re, fma(g.t2.re, b.im, g.t2.im * b.re)));
  a.re -= s.re;
  a.im -= s.im;
  b.re = fma(-s.re, g.v2.re, fma(-s.im, g.v2.im, b.re));
  b.im = fma(-s.im, g.v2.re, fma(s.re, g.v2.im, b.im));
}

// Wilkinson's shift for the trailing 2 x 2 block [[a, b], [c, d]] of an active block (zlahqr): the
// eigenvalue d - bc / (x + y) nearer to d, x = (a - d) / 2, y = +-sqrt(x^2 + bc) with Re(conj(x) y) >= 0.
// It is the same eigenvalue zlahqr forms from sqrt(b) sqrt(c), with one complex square root and one
// reciprocal.  The shift sets only the rate of convergence: a value that is not finite (range
// exceeded) falls back to d.
__device__ inline zc wilkinson_shift(zc a, zc b, zc c, zc d) {
  const zc x = 0.5 * (a - d), p = b * c;
  const zc q = x * x + p;
  const double m = sqrt_nr(fmax(norm2(q), 1e-300));
  const double tq = sqrt_nr(fmax(0.5 * (m + fabs(q.re)), 1e-300));
  const double h = 0.5 * q.im * rcp_nr(tq);
  zc y = (q.re >= 0.0) ? mk(tq, h) : mk(fabs(h), copysign(tq, q.im));
  if (x.re * y.re + x.im * y.im < 0.0) y = -y;
  const zc den = x + y;
  const double nd = norm2(den);
  zc t = d;
  if (nd > 0.0) {
    const double rn = rcp_nr(nd);
    const zc r = p * conj(den);
    t = d - mk(r.re * rn, r.im * rn);
  }
  if (!(fabs(t.re) + fabs(t.im) < 1e300)) t = d;
  return t;
}

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Deflation test of zlahqr with |h| = |re| + |im| for the subdiagonal entries (they may be complex).
__device__ inline bool negligible_c(const zc* S, int ld, int k, int lo, int hi, double smlnum) {
  auto h = [&](int r, int c) { return S[(long)c * ld + r]; };
  const zc hkk1 = h(k, k - 1);
  if (abs1(hkk1) <= smlnum) return true;
  double tst = abs1(h(k - 1, k - 1)) + abs1(h(k, k));
  if (tst == 0.0) {
    if (k - 2 >= lo) tst += abs1(h(k - 1, k - 2));
    if (k + 1 <= hi) tst += abs1(h(k + 1, k));
  }
  if (abs1(hkk1) <= kUlp * tst) {
    const double ab = fmax(abs1(hkk1), abs1(h(k - 1, k)));
    const double ba = fmin(abs1(hkk1), abs1(h(k - 1, k)));
    const double aa = fmax(abs1(h(k, k)), abs1(h(k - 1, k - 1) - h(k, k)));
    const double bb = fmin(abs1(h(k, k)), abs1(h(k - 1, k - 1) - h(k, k)));
    const double s = aa + ab;
    if (ba * (ab / s) <= fmax(smlnum, kUlp * (bb * (aa / s)))) return true;
  }
  return false;
}

// Eigenvalues of the m x m upper Hessenberg S (LDS, column-major, m <= 64) by single-shift QR
// without Schur form or vectors (zlahqr with wantt = wantz = false), run by one wavefront.
__device__ __attribute__((always_inline)) inline bool small_eigvals_wave(zc* S, int m, zc* w, double smlnum) {
  const int lane = threadIdx.x & 63;
  auto Sp = [&](int r, int c) -> zc& { return S[c * m + r]; };
  int i = m - 1, kdefl = 0, its = 0;
  const int itmax = 30 * max(10, m) * m;
  while (i >= 0) {
    if (++its > itmax) return false;
    const int k = lane + 1;
    const bool neg = (k <= i) && negligible_c(S, m, k, 0, m - 1, smlnum);
    const unsigned long long mask = __ballot(neg);
    const int l = mask ? (64 - __clzll(mask)) : 0;  // largest k with a negligible subdiagonal
    wave_sync();
    if (l > 0 && lane == 0) Sp(l, l - 1) = mk(0.0);
    wave_sync();
    if (l >= i) {
      if (lane == 0) w[i] = Sp(i, i);
      i = l - 1;
      kdefl = 0;
      wave_sync();
      continue;
    }
    ++kdefl;
    zc t;
    if (kdefl % 20 == 0) {
      t = mk(0.75 * abs1(Sp(i, i - 1))) + Sp(i, i);
    } else if (kdefl % 10 == 0) {
      t = mk(0.75 * abs1(Sp(l + 1, l))) + Sp(l, l);
    } else {
      t = wilkinson_shift(Sp(i - 1, i - 1), Sp(i - 1, i), Sp(i, i - 1), Sp(i, i));
    }
    for (int kk = l; kk < i; ++kk) {
      zc v1, v2;
      if (kk == l) {
        const zc h11s = Sp(l, l) - t, h21 = Sp(l + 1, l);
        const double sc = abs1(h11s) + abs1(h21);
        v1 = h11s * (1.0 / sc);
        v2 = h21 * (1.0 / sc);
      } else {
        v1 = Sp(kk, kk - 1);
        v2 = Sp(kk + 1, kk - 1);
      }
      const Rot2 g = make_reflector2(v1, v2);
      wave_sync();
      for (int c = kk + lane; c <= i; c += 64) rot2_left(g, Sp(kk, c), Sp(kk + 1, c));
      if (kk > l && lane == 0) {
        Sp(kk, kk - 1) = v1;
        Sp(kk + 1, kk - 1) = mk(0.0);
      }
      wave_sync();
      const int rmax = min(kk + 2, i);
      for (int r = l + lane; r <= rmax; r += 64) rot2_right(g, Sp(r, kk), Sp(r, kk + 1));
      wave_sync();
    }
  }
  return true;
}

// Make the subdiagonal H(L+1:I, L:I-1) real by the diagonal unitary similarity D = diag(d_L..d_I),
// d_i = d_{i-1} h_i / |h_i| for complex h_i (d_i = d_{i-1} for real h_i), applied to H (rows L..I
// from the subdiagonal rightwards, rows above L in columns L..I) and to Z.  Real subdiagonal entries
// keep their value, as in zlahqr's preparatory loop (a sequential scan whose scalings compound the
// same way).  d: LDS scratch of I - L + 1 entries.  Called by the whole workgroup.
__device__ void real_subdiag(zc* H, int ldh, zc* Z, int ldz, int n, int L, int I, zc* d) {
  __shared__ int s_first;
  auto Hp = [&](int r, int c) -> zc& { return H[(long)c * ldh + r]; };
  if (threadIdx.x == 0) s_first = I + 1;
  __syncthreads();
  for (int i = L + 1 + threadIdx.x; i <= I; i += blockDim.x)
    if (Hp(i, i - 1).im != 0.0) atomicMin(&s_first, i);
  __syncthreads();
  const int f = s_first;
  __syncthreads();
  if (f > I) return;
  const int m = I - f + 1;
  if (threadIdx.x == 0) {
    zc acc = mk(1.0);
    for (int i = f; i <= I; ++i) {
      const zc h = Hp(i, i - 1);
      if (h.im != 0.0) acc = acc * (h * (1.0 / absz(h)));
      d[i - f] = acc;
    }
  }
  __syncthreads();
  for (long e = threadIdx.x; e < (long)m * (n - f + 1); e += blockDim.x) {
    const int r = f + (int)(e % m), c = f - 1 + (int)(e / m);
    if (c < r - 1) continue;
    zc& h = Hp(r, c);
    if (c == r - 1) {
      if (h.im != 0.0) h = mk(absz(h));
      continue;
    }
    const zc dc = (c <= I) ? d[c - f] : mk(1.0);
    h = conj(d[r - f]) * h * dc;
  }
  for (long e = threadIdx.x; e < (long)f * m; e += blockDim.x) {
    const int r = (int)(e % f), c = f + (int)(e / f);
    Hp(r, c) = Hp(r, c) * d[c - f];
  }
  for (long e = threadIdx.x; e < (long)n * m; e += blockDim.x) {
    const int r = (int)(e % n), c = f + (int)(e / n);
    zc& z = Z[(long)c * ldz + r];
    z = z * d[c - f];
  }
  __syncthreads();
}

// ------------------------------------------------------------------------------------------------
// Aggressive early deflation (the zlaqr3 algorithm of Braman, Byers and Mathias).  Before a
// multishift sweep, the trailing AED_W x AED_W window of the active block is brought to Schur form
// T = V^H W V.  Its coupling to the rest of H is the single spike s = H(kw, kw-1), which becomes
// the vector s conj(V(0, :)) in the new basis: every bottom entry of T whose spike component is
// negligible deflates at once; the others are moved to the top of T by swaps of adjacent diagonal
// entries.  When something deflated, the spike is reflected back onto its first entry, the
// undeflated part is reduced to Hessenberg form again, and V is applied to the rest of H and to Z.
// The undeflated eigenvalues of the window are the next sweep's shifts.
// ------------------------------------------------------------------------------------------------
constexpr int AED_W = 16;          // window rows (at most 32: lanes 0..31 take T's lines, 32..63 V's rows)
constexpr int AED_LD = AED_W + 1;  // leading dimension of T and V in LDS

// Schur form of the m x m upper Hessenberg S (LDS, column-major, leading dimension ld, m <= 32), with
// the Schur vectors accumulated into V (initially the identity): zlahqr with wantt = wantz = true on
// one wavefront.  Left products touch S's columns (one per lane), right products S's rows (lanes
// 0..31) and V's rows (lanes 32..63) in the same phase.  Returns false at the iteration limit.
__device__ __attribute__((always_inline)) inline bool schur_wave(zc* S, zc* V, int m, int ld, double smlnum) {
  const int lane = threadIdx.x & 63;
  auto Sp = [&](int r, int c) -> zc& { return S[c * ld + r]; };
  int i = m - 1, kdefl = 0, its = 0;
  const int itmax = 30 * max(10, m) * m;
  while (i >= 0) {
    if (++its > itmax) return false;
    const int k = lane + 1;
    const bool neg = (k <= i) && negligible_c(S, ld, k, 0, m - 1, smlnum);
    const unsigned long long mask = __ballot(neg);
    const int l = mask ? (64 - __clzll(mask)) : 0;
    wave_sync();
    if (l > 0 && lane == 0) Sp(l, l - 1) = mk(0.0);
    wave_sync();
    if (l >= i) {
      i = l - 1;
      kdefl = 0;
      continue;
    }
    ++kdefl;
    zc t;
    if (kdefl % 20 == 0) {
      t = mk(0.75 * abs1(Sp(i, i - 1))) + Sp(i, i);
    } else if (kdefl % 10 == 0) {
      t = mk(0.75 * abs1(Sp(l + 1, l))) + Sp(l, l);
    } else {
      t = wilkinson_shift(Sp(i - 1, i - 1), Sp(i - 1, i), Sp(i, i - 1), Sp(i, i));
    }
    for (int kk = l; kk < i; ++kk) {
      zc v1, v2;
      if (kk == l) {
        const zc h11s = Sp(l, l) - t, h21 = Sp(l + 1, l);
        const double sc = abs1(h11s) + abs1(h21);
        v1 = h11s * (1.0 / sc);
        v2 = h21 * (1.0 / sc);
      } else {
        v1 = Sp(kk, kk - 1);
        v2 = Sp(kk + 1, kk - 1);
      }
      const Rot2 g = make_reflector2(v1, v2);
      wave_sync();
      if (kk + lane < m) rot2_left(g, Sp(kk, kk + lane), Sp(kk + 1, kk + lane));
      if (kk > l && lane == 0) {
        Sp(kk, kk - 1) = v1;
        Sp(kk + 1, kk - 1) = mk(0.0);
      }
      wave_sync();
      {
        // one code path for T's rows (lanes 0..31) and V's rows (lanes 32..63)
        const bool isv = lane >= 32;
        const int r = isv ? lane - 32 : lane;
        zc* const base = isv ? V : S;
        if (r <= (isv ? m - 1 : min(kk + 2, i))) rot2_right(g, base[kk * ld + r], base[(kk + 1) * ld + r]);
      }
      wave_sync();
    }
  }
  return true;
}

// Plane rotation [c s; -conj(s) c] with c real that maps (f, g) to (r, 0) (the zlartg contract):
// c = |f| / sqrt(|f|^2 + |g|^2), s = (f / |f|) conj(g) / sqrt(|f|^2 + |g|^2).  f and g are first scaled by
// the same (approximate) reciprocal of max(|f|, |g|)_1, which leaves c and s unchanged; then one
// reciprocal square root gives both, c = |f|^2 q and s = f conj(g) q with q = (|f|^2 (|f|^2 + |g|^2))^-1/2.
__device__ inline void lartg(zc f, zc g, double& c, zc& s) {
  const double af = abs1(f), ag = abs1(g);
  if (ag == 0.0) {
    c = 1.0;
    s = mk(0.0);
    return;
  }
  const double r = __builtin_amdgcn_rcp(fmax(af, ag));
  const zc fs = f * r, gs = g * r;
  const double f2 = norm2(fs), g2 = norm2(gs);
  if (af == 0.0 || f2 < 1e-200) {
    // |f| negligible against |g| (or zero): the general form of zlartg
    const double ga = absz(g);
    if (af == 0.0) {
      c = 0.0;
      s = conj(g) * (1.0 / ga);
      return;
    }
    const double fa = absz(f), nrm = hypot(fa, ga);
    c = fa / nrm;
    s = (f * (1.0 / fa)) * conj(g) * (1.0 / nrm);
    return;
  }
  const double q = rsqrt(f2 * (f2 + g2));
  c = f2 * q;
  s = fs * conj(gs) * q;
}

// Householder reflector I - tau u u^H (u_0 = 1) with (I - tau u u^H)^H (alpha, x) = (beta, 0) for the
// vector held by lanes lo..hi-1 of one wavefront (x0 on lane lo): the zlarfg algorithm.  Every lane
// receives tau and its own entry of u (zero outside lo..hi-1); returns beta.
__device__ inline double wave_larfg(zc xi, int lo, int hi, zc alpha, zc& tau, zc& ui) {
  const int lane = threadIdx.x & 63;
  const double xn2 = wave_sum((lane > lo && lane < hi) ? norm2(xi) : 0.0);
  if (xn2 == 0.0 && alpha.im == 0.0) {
    tau = mk(0.0);
    ui = (lane == lo) ? mk(1.0) : mk(0.0);
    return alpha.re;
  }
  const double beta = -copysign(sqrt(norm2(alpha) + xn2), alpha.re);
  tau = (mk(beta) - alpha) * (1.0 / beta);
  const zc sc = zdiv(mk(1.0), alpha - mk(beta));
  ui = (lane == lo) ? mk(1.0) : (lane > lo && lane < hi) ? xi * sc : mk(0.0);
  return beta;
}

// One wavefront, after schur_wave: the deflation test and reordering of zlaqr3 (ztrexc swaps), then,
// when something deflated and more than one entry did not, the spike reflection and the Hessenberg
// reduction (zgehrd) of the undeflated block, all accumulated into V.  hv: LDS vector of AED_W.
// Returns the number ns of undeflated eigenvalues, now T(0..ns-1) before the reduction (their values
// are written to w[0..ns-1]).
__device__ __attribute__((always_inline)) inline int aed_deflate_wave(zc* T, zc* V, zc s, zc* hv, zc* w, double smlnum) {
  const int lane = threadIdx.x & 63, m = AED_W, ld = AED_LD;
  auto Tp = [&](int r, int c) -> zc& { return T[c * ld + r]; };
  auto Vp = [&](int r, int c) -> zc& { return V[c * ld + r]; };
  int ns = m, ilst = 0;
  const double as = abs1(s);
  for (int knt = 0; knt < m; ++knt) {
    const int k = ns - 1;
    double foo = abs1(Tp(k, k));
    if (foo == 0.0) foo = as;
    if (as * abs1(Vp(0, k)) <= fmax(smlnum, kUlp * foo)) {
      --ns;
      continue;
    }
    for (int p = k - 1; p >= ilst; --p) {
      // swap T(p, p) and T(p+1, p+1)
      const zc t11 = Tp(p, p), t22 = Tp(p + 1, p + 1);
      double c;
      zc sn;
      lartg(Tp(p, p + 1), t22 - t11, c, sn);
      // rows p, p+1 of T's columns p+2..m-1 take (c, s); columns p, p+1 of T's rows 0..p-1 and of V's
      // rows take (c, conj(s)): one code path, the pair and the sine chosen per lane
      const bool isv = lane >= 32, rowop = !isv && lane >= p + 2;
      const int q = isv ? lane - 32 : lane;
      const bool act = isv ? q < m : (rowop ? lane < m : lane < p);
      zc* const base = isv ? V : T;
      const int o1 = rowop ? q * ld + p : p * ld + q, o2 = rowop ? o1 + 1 : o1 + ld;
      const zc sr = rowop ? sn : conj(sn);
      if (act) {
        const zc x = base[o1], y = base[o2];
        base[o1] = c * x + sr * y;
        base[o2] = c * y - conj(sr) * x;
      }
      if (lane == 0) {
        Tp(p, p) = t22;
        Tp(p + 1, p + 1) = t11;
      }
      wave_sync();
    }
    ++ilst;
  }
  for (int j = lane; j < ns; j += 64) w[j] = Tp(j, j);
  wave_sync();
  if (ns == m || ns <= 1 || (s.re == 0.0 && s.im == 0.0)) return ns;

  // reflect the spike s conj(V(0, 0..ns-1)) onto its first entry: T <- P^H T P, V <- V P
  auto reflect = [&](int lo, int hi, zc tau, zc ui) {
    // P = I - tau u u^H on indices lo..hi-1 (u in hv).  Left: T(lo..hi-1, lo..m-1); right: T(0..hi-1,
    // lo..hi-1) and V(0..m-1, lo..hi-1).
    if (lane < hi && lane >= lo) hv[lane] = ui;
    wave_sync();
    const zc tc = conj(tau);
    if (lane < 32 && lane >= lo && lane < m) {
      const int c = lane;
      zc d = mk(0.0);
      for (int r = lo; r < hi; ++r) fma_acc(d, conj(hv[r]), Tp(r, c));
      d = tc * d;
      for (int r = lo; r < hi; ++r) Tp(r, c) -= hv[r] * d;
    }
    wave_sync();
    zc* rowp = nullptr;
    int q = lane;
    if (lane < 32) {
      if (lane < hi) rowp = T;
    } else if (lane - 32 < m) {
      rowp = V;
      q = lane - 32;
    }
    if (rowp) {
      zc d = mk(0.0);
      for (int c = lo; c < hi; ++c) fma_acc(d, rowp[c * ld + q], hv[c]);
      d = tau * d;
      for (int c = lo; c < hi; ++c) rowp[c * ld + q] -= d * conj(hv[c]);
    }
    wave_sync();
  };
  {
    const zc wi = (lane < ns) ? conj(Vp(0, lane)) : mk(0.0);
    const zc alpha = conj(Vp(0, 0));
    zc tau, ui;
    wave_larfg(wi, 0, ns, alpha, tau, ui);
    wave_sync();
    if (tau.re != 0.0 || tau.im != 0.0) reflect(0, ns, tau, ui);
  }
  // Hessenberg reduction of T(0..ns-1, 0..ns-1); the left products also reach columns ns..m-1
  for (int j = 0; j + 2 < ns; ++j) {
    const zc xi = (lane < 32) ? Tp(lane, j) : mk(0.0);
    const zc alpha = Tp(j + 1, j);
    zc tau, ui;
    const double beta = wave_larfg(xi, j + 1, ns, alpha, tau, ui);
    wave_sync();
    if (lane == j + 1) Tp(j + 1, j) = mk(beta);
    if (lane > j + 1 && lane < ns) Tp(lane, j) = mk(0.0);
    wave_sync();
    if (tau.re != 0.0 || tau.im != 0.0) reflect(j + 1, ns, tau, ui);
  }
  return ns;
}

// The trailing window H(kw..I, kw..I) with kw = I - AED_W + 1 (kw - 1 >= L, so the spike lies inside the
// active block).  Called by the whole workgroup; T and V occupy 2 AED_LD AED_W entries of the LDS
// scratch `buf`.  Returns the number of deflated eigenvalues (the window's bottom rows, now with a zero
// subdiagonal above them), or -1 when the window's QR reached its iteration limit (H and Z unchanged).
// When at least MS_NB / 2 undeflated eigenvalues remain, the last min(ns, MS_NB) of them are written to
// shifts and sh[1] is their count, else sh[1] = 0.
__device__ __attribute__((always_inline)) inline int aed_window(zc* H, int ldh, zc* Z, int ldz, int n, int I, zc* buf, zc* hv, zc* shifts, int* sh,
                          double smlnum) {
  const int tid = threadIdx.x, nt = blockDim.x, kw = I - AED_W + 1;
  zc* T = buf;
  zc* V = buf + AED_LD * AED_W;
  auto Hp = [&](int r, int c) -> zc& { return H[(long)c * ldh + r]; };
  for (int e = tid; e < AED_W * AED_W; e += nt) {
    const int r = e % AED_W, c = e / AED_W;
    T[c * AED_LD + r] = (r <= c + 1) ? Hp(kw + r, kw + c) : mk(0.0);
    V[c * AED_LD + r] = mk(r == c ? 1.0 : 0.0);
  }
  const zc s = Hp(kw, kw - 1);
  __syncthreads();
  if (tid < 64) {
    const bool ok = schur_wave(T, V, AED_W, AED_LD, smlnum);
    int ns = -1;
    if (ok) ns = aed_deflate_wave(T, V, s, hv, hv + AED_W, smlnum);
    const int nsh = (ns >= MS_NB / 2) ? min(ns, MS_NB) : 0;
    if (tid == 0) {
      sh[0] = ns;
      sh[1] = nsh;
    }
    if (tid < nsh) shifts[tid] = hv[AED_W + ns - nsh + tid];
  }
  __syncthreads();
  const int ns = sh[0];
  if (ns < 0 || ns == AED_W) return ns < 0 ? -1 : 0;
  // write the window back, the new spike, and apply V to the rest of H and to Z
  for (int e = tid; e < AED_W * AED_W; e += nt) {
    const int r = e % AED_W, c = e / AED_W;
    Hp(kw + r, kw + c) = T[c * AED_LD + r];
  }
  if (tid == 0) Hp(kw, kw - 1) = s * conj(V[0]);
  // lines: rows 0..kw-1 of H (H(r, kw..I) V), columns I+1..n-1 of H (V^H H(kw..I, c), taken as the
  // conjugate row times V), rows of Z.  Two lanes of a wavefront share a line, each forming half of
  // its outputs while streaming the same inputs; the stores follow the last load of both.
  constexpr int HALF = AED_W / 2;
  const int nfa = kw, nfr = n - 1 - I, lines = nfa + nfr + n;
  for (int task = tid; task < 2 * lines; task += nt) {
    const int ln = task >> 1, j0 = (task & 1) * HALF;
    zc* p;
    long st;
    bool cj;
    if (ln < nfa) {
      p = H + (long)kw * ldh + ln;
      st = ldh;
      cj = false;
    } else if (ln < nfa + nfr) {
      p = H + (long)(I + 1 + ln - nfa) * ldh + kw;
      st = 1;
      cj = true;
    } else {
      p = Z + (long)kw * ldz + (ln - nfa - nfr);
      st = ldz;
      cj = false;
    }
    zc acc[HALF];
#pragma unroll
    for (int j = 0; j < HALF; ++j) acc[j] = mk(0.0);
    for (int k = 0; k < AED_W; ++k) {
      zc x = p[k * st];
      if (cj) x.im = -x.im;
#pragma unroll
      for (int j = 0; j < HALF; ++j) fma_acc(acc[j], x, V[(j0 + j) * AED_LD + k]);
    }
#pragma unroll
    for (int j = 0; j < HALF; ++j) {
      if (cj) acc[j].im = -acc[j].im;
      p[(j0 + j) * st] = acc[j];
    }
  }
  __syncthreads();
  return AED_W - ns;
}

// The deferred products read a window's rotations as scalars: every lane of a wave applies the same
// rotation, so its six doubles are loaded once per wave into scalar registers (through the scalar
// data cache, from a copy in global memory) instead of being broadcast from LDS into all 64 lanes,
// which would make the LDS return path the bottleneck of the deferred products.
typedef __attribute__((address_space(4))) const double KDouble;
constexpr int MS_RING = MS_S * MS_NB * 6;  // doubles of one window's rotations, bulge-major

// Hides the pointer from the optimiser so that the scalar loads through it are neither hoisted above
// the barrier that publishes the window's rotations nor merged across windows.
__device__ inline const KDouble* scalar_view(const double* p) {
  asm volatile("" : "+s"(p));
  return (const KDouble*)p;
}
__device__ inline Rot2 scalar_rot(const KDouble* R, int i) {
  Rot2 g;
  g.t1 = mk(R[6 * i], R[6 * i + 1]);
  g.v2 = mk(R[6 * i + 2], R[6 * i + 3]);
  g.t2 = mk(R[6 * i + 4], R[6 * i + 5]);
  return g;
}

// Applies one window's left rotations (rotation (j, s) on elements 30-2j+s, 31-2j+s) to one line,
// bulge by bulge.  Bulge j touches elements 30-2j .. 46-2j, so two elements enter before it and two
// leave after it, and only that live window is held in registers.  Element q lives at p[q * st].
// A column right of the window takes the left rotations as they are; a row above it or a row of Z
// takes the right rotations, and since right-multiplying a row equals left-multiplying its conjugate
// by the same reflector, such a line is conjugated on load and store (sgn flips the sign bit of the
// imaginary parts).  Every lane therefore runs the same code whatever its line.  Elements outside
// [qlo, qhi] lie outside the active block: they enter as zero, see only identity rotations and are
// not stored.  FULL: the whole window lies inside the active block.
template <bool FULL>
__device__ __attribute__((always_inline)) inline void ms_apply_line(zc* p, long st, unsigned long long sgn, int qlo, int qhi,
                                                                  const KDouble* R) {
  constexpr int Q0 = 2 * (MS_NB - 1);
  zc x[MS_LEN];
  auto in = [&](int q) { return FULL || (q >= qlo && q <= qhi); };
  auto flip = [&](zc v) {
    v.im = __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, v.im) ^ sgn);
    return v;
  };
  // element addresses are formed where they are used, from a copy of the line's stride that the
  // optimiser cannot see through: precomputed for all 48 elements they would fill the registers
  // the rotations need
  auto stride = [&]() {
    long sb = st;
    asm volatile("" : "+v"(sb));
    return sb;
  };
  auto ld = [&](int q, long sb) { return in(q) ? p[q * sb] : mk(0.0); };
  auto store = [&](int q, long sb) {
    if (in(q)) p[q * sb] = flip(x[q]);
  };
  // the elements entering at bulges j + 1 and j + 2 are loaded two bulges ahead, so that their
  // latency overlaps the rotations of the bulges in between
  zc pf[2][2];
  {
    const long sb = stride();
#pragma unroll
    for (int q = Q0; q <= Q0 + MS_S; ++q) x[q] = ld(q, sb);
#pragma unroll
    for (int a = 1; a <= 2; ++a) {
      pf[a & 1][0] = ld(Q0 - 2 * a, sb);
      pf[a & 1][1] = ld(Q0 - 2 * a + 1, sb);
    }
#pragma unroll
    for (int q = Q0; q <= Q0 + MS_S; ++q) x[q] = flip(x[q]);
  }
  // the rotations stream through scalar registers one ahead of their use; the scheduling barrier
  // keeps each scalar load above the rotation before it (only ALU instructions may cross)
  Rot2 g = scalar_rot(R, 0);
#pragma unroll
  for (int j = 0; j < MS_NB; ++j) {
    const int qb = Q0 - 2 * j;
    const long sb = stride();
    if (j > 0) {
      x[qb] = flip(pf[j & 1][0]);
      x[qb + 1] = flip(pf[j & 1][1]);
      if (j + 2 < MS_NB) {
        pf[j & 1][0] = ld(qb - 4, sb);
        pf[j & 1][1] = ld(qb - 3, sb);
      }
    }
#pragma unroll
    for (int s = 0; s < MS_S; ++s) {
      const int i = j * MS_S + s;
      const Rot2 gn = scalar_rot(R, i + 1 < MS_S * MS_NB ? i + 1 : i);
      __builtin_amdgcn_sched_barrier(1);
      // steps past nsteps hold zero (identity) rotations
      rot2_left(g, x[qb + s], x[qb + s + 1]);
      g = gn;
    }
    store(qb + MS_S - 1, sb);
    store(qb + MS_S, sb);
  }
  const long sb = stride();
#pragma unroll
  for (int q = 0; q < MS_S - 1; ++q) store(q, sb);
}

// The same left rotations on one column line, spread over the 16 lanes of a row of the wavefront so that
// few lines (the leader's critical columns) do not run one rotation after another on one lane.  Lane j
// applies bulge j's rotations: at step s they take elements (30 - 2j + s, 31 - 2j + s).  Lane j holds
// the second element from step to step, and the first one it releases is the element that bulge j - 1
// released one step earlier (bulges two rows apart), so each step passes one element from lane j - 1 to
// lane j by a row shift.  Lane 0 takes fresh elements 32..46 instead.  An element released by bulge 15,
// or at the last step, and every lane's final element are final for this window.  Each element sees the
// same rotations in the same order as in ms_apply_line.  R: the window's rotations in LDS, step-major.
template <bool FULL>
__device__ __attribute__((always_inline)) inline void ms_apply_column_lanes(zc* p, int qlo, int qhi, const Rot2* R, int j) {
  constexpr int Q0 = 2 * (MS_NB - 1);
  static_assert(MS_NB == 16, "one bulge per lane of a 16-lane row");
  auto in = [&](int q) { return FULL || (q >= qlo && q <= qhi); };
  auto ld = [&](int q) { return in(q) ? p[q] : mk(0.0); };
  auto shr1 = [](double v) {
    const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
    const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)u, 0x111, 0xF, 0xF, false);
    const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), 0x111, 0xF, 0xF, false);
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
  };
  zc fresh[MS_S - 1];
  zc carry = ld(Q0 - 2 * j), nxt = ld(Q0 + 1 - 2 * j);
  if (j == 0) {
#pragma unroll
    for (int s = 1; s < MS_S; ++s) fresh[s - 1] = ld(Q0 + 1 + s);
  }
#pragma unroll
  for (int s = 0; s < MS_S; ++s) {
    const Rot2 g = R[s * MS_NB + j];
    rot2_left(g, carry, nxt);
    const int q = Q0 - 2 * j + s;  // the element this lane releases
    if ((j == MS_NB - 1 || s == MS_S - 1) && in(q)) p[q] = carry;
    if (s + 1 < MS_S) {
      const zc recv = mk(shr1(carry.re), shr1(carry.im));
      carry = nxt;
      nxt = (j == 0) ? fresh[s] : recv;
    } else {
      carry = nxt;
    }
  }
  if (in(Q0 + MS_S - 2 * j)) p[Q0 + MS_S - 2 * j] = carry;
}

// Applies the single-shift rotations 0..wlen-1 (rotation j on elements j, j+1) to one line whose
// element j lives at p[j * stride]; two elements are live at a time.
template <bool LEFT>
__device__ inline void ss_apply_line(zc* p, long stride, const Rot* R, int wlen) {
  zc a = p[0];
  zc b1 = (wlen >= 1) ? p[stride] : mk(0.0);
  zc b2 = (wlen >= 2) ? p[2 * stride] : mk(0.0);
#pragma unroll 2
  for (int j = 0; j < wlen; ++j) {
    zc b = b1;
    b1 = b2;
    b2 = (j + 3 <= wlen) ? p[(long)(j + 3) * stride] : mk(0.0);  // two elements ahead
    if (LEFT) rot_left(R[j], a, b);
    else rot_right(R[j], a, b);
    p[(long)j * stride] = a;
    a = b;
  }
  p[(long)wlen * stride] = a;
}

// ------------------------------------------------------------------------------------------------
// QR teams.  When the matrices do not fill the device, each matrix gets a leader workgroup and G - 1
// helpers (one cooperative launch, so that all of them are resident).  The leader runs the iteration:
// deflation search, deflation windows, the bulge chase of every multishift window, and the window's
// left products on the few columns the next window's block takes in (the window's critical columns).
// Every other deferred product goes to the helpers: the left products on the columns further right, and
// the right products on the rows of H above the window and on every row of Z.  The leader publishes each
// window's rotations and extents in a slot of a per-matrix ring (NSLOT slots) and counts it in ctl[0];
// helper h counts its finished windows in ctl[CTL_DONE + h - 1].
//
// Order.  A line (a row or column of H, a row of Z) belongs to one wavefront of one helper for the whole
// launch and receives the windows in their order.  The helpers start window w only when all of them
// have finished window w - 1, because a row line and a column line of different owners cross in one
// entry, whose left products (earlier windows) precede its right products (later windows).  The
// leader applies window w to its critical columns only after the helpers have finished window w - 1,
// whose share held those columns.  What the helpers write lies above the published windows' rows or
// right of their critical columns, where the leader writes no more in that sweep; what they read from
// the leader was written before the publication (release / acquire at agent scope; one thread of a
// workgroup performs them, next to a workgroup barrier, as the cache operations they stand for act on
// the caches the whole workgroup shares).  Before anything
// else touches H outside the blocks or Z (the end of a sweep: diagonal scaling, deflation window,
// single-shift iterations, the end), the leader waits until the helpers have caught up.  A slot is
// rewritten only after every helper has finished the window it held.  Every wait ends: on the counters,
// on ctl[1] (stop) or ctl[2] (abort), or after `limit` ticks of the wall clock, after which the launch
// reports failure.
// ------------------------------------------------------------------------------------------------
constexpr int NSLOT = 8;
constexpr int DESC_INTS = 8;  // kbase, rlo, rhi, whole window inside the block, last critical column

struct HqrTeam {
  int G = 1;                   // workgroups per matrix (1: the leader does every product itself)
  int batch = 0;
  int* ctl = nullptr;          // per matrix CTL_INTS counters, zero at launch
  double* slots = nullptr;     // per matrix NSLOT x MS_RING doubles of rotations
  int* desc = nullptr;         // per matrix NSLOT x DESC_INTS
  unsigned long long limit = 0;  // wall-clock ticks any wait may take
};

// Thread 0 waits until every helper has finished `target` windows; false on an abort or a timeout.
__device__ inline bool helpers_reached(const HqrTeam& t, int* ctl, int target) {
  return counters_reach(ctl + CTL_DONE, t.G - 1, target, ctl + 2, t.limit);
}

// Leader: the whole workgroup waits until every helper has finished `target` windows, then sees their
// writes.  Returns false on a timeout.
__device__ inline bool team_wait(const HqrTeam& t, int* ctl, int target, int* s_flag) {
  if (threadIdx.x == 0) {
    *s_flag = helpers_reached(t, ctl, target) ? 1 : 0;
    // one acquire per workgroup: its invalidation covers the caches the workgroup's waves share
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
  }
  __syncthreads();
  return __builtin_amdgcn_readfirstlane(*s_flag) != 0;
}

// Leader: window `w`'s rotations and descriptor are in its slot; make them (and every earlier write)
// visible, then count the window.
__device__ inline void team_publish(int* ctl, int w) { wg_release_store(ctl, w + 1); }

// Helper h (1 .. G-1) of matrix b: applies every published window to its lines, window after window,
// until the leader stops.  Line l < n: row l of H (right rotations, rows above the window); n <= l < 2n:
// row l - n of Z; 2n <= l < 3n: column l - 2n of H (left rotations, right of the critical columns).
// Dense enumeration of one chase step's products.  The active bulges own na consecutive runs of
// products whose lengths change by 2 from one bulge to the next: run0 + 2u (dir = +1: left products,
// u counting from the leading bulge) or run0 - 2u (dir = -1: right products).  Returns the run u that
// holds element e, i.e. P(u) <= e < P(u + 1) with P(u) = u run0 + dir u (u - 1), and its offset e - P(u)
// in x.  The float estimate of the quadratic's root is exact to far below one for these sizes, and
// one step each way corrects it.
__device__ inline int chase_run(int e, int na, int run0, int dir, int& x) {
  auto P = [&](int u) { return u * run0 + dir * u * (u - 1); };
  const float b = (float)(run0 - dir);
  const float uf = (dir > 0) ? 0.5f * (sqrtf(b * b + 4.0f * (float)e) - b)
                             : 0.5f * (b - sqrtf(fmaxf(b * b - 4.0f * (float)e, 0.0f)));
  int u = min(max((int)uf, 0), na - 1);
  if (u + 1 < na && P(u + 1) <= e) ++u;
  if (u > 0 && P(u) > e) --u;
  x = e - P(u);
  return u;
}

template <int NT>
__device__ void hqr_help(int n, zc* H, int ldh, zc* Z, int ldz, const HqrTeam& t, int b, int h) {
  __shared__ int s_go, s_desc[DESC_INTS];
  int* ctl = t.ctl + (long)b * CTL_INTS;
  const double* slots = t.slots + (long)b * NSLOT * MS_RING;
  const int* desc = t.desc + (long)b * NSLOT * DESC_INTS;
  const int nh = t.G - 1, nw = NT / 64, lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int nchunks = (3 * n + 63) / 64;
  constexpr unsigned long long kSign = 1ull << 63;
  for (int w = 0;; ++w) {
    if (threadIdx.x == 0) {
      int go = 0;
      const unsigned long long t0 = wall_clock64();
      for (;;) {
        if (ld_relaxed(ctl) > w) { go = 1; break; }
        if (ld_relaxed(ctl + 1) || ld_relaxed(ctl + 2)) break;
        if (wall_clock64() - t0 > t.limit) { st_release(ctl + 2, 1); break; }
        __builtin_amdgcn_s_sleep(4);
      }
      // every helper has finished the window before (the group step)
      if (go && w > 0 && !helpers_reached(t, ctl, w)) go = 0;
      if (go) {
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        for (int i = 0; i < DESC_INTS; ++i) s_desc[i] = desc[(w % NSLOT) * DESC_INTS + i];
      }
      s_go = go;
    }
    __syncthreads();
    if (!__builtin_amdgcn_readfirstlane(s_go)) return;
    // the slot's rotations are read through the scalar cache: drop what it holds of older windows, then
    // fetch the slot's 64-byte lines together (one dword of each, the waves taking turns, all loads in
    // flight), so that the rotations streamed one ahead of their use hit instead of waiting on L2 each
    __builtin_amdgcn_s_dcache_inv();
    __builtin_amdgcn_s_waitcnt(0xC07F);
    {
      constexpr int kLines = MS_RING * 8 / 64, kWaves = NT / 64;
      static_assert(kLines % kWaves == 0, "ring lines split evenly over the waves");
      typedef __attribute__((address_space(4))) const unsigned KU32;
      const unsigned* rp = reinterpret_cast<const unsigned*>(slots + (long)(w % NSLOT) * MS_RING);
      asm volatile("" : "+s"(rp));
      const KU32* ku = (const KU32*)rp;
      unsigned acc = 0;
#pragma unroll
      for (int i = 0; i < kLines / kWaves; ++i) acc ^= ku[(i * kWaves + wid) * 16];
      asm volatile("" ::"s"(acc));
    }
    const int kbase = __builtin_amdgcn_readfirstlane(s_desc[0]), rlo = __builtin_amdgcn_readfirstlane(s_desc[1]);
    const int rhi = __builtin_amdgcn_readfirstlane(s_desc[2]);
    const bool full = __builtin_amdgcn_readfirstlane(s_desc[3]) != 0;
    const int chi = __builtin_amdgcn_readfirstlane(s_desc[4]);
    const KDouble* R = scalar_view(slots + (long)(w % NSLOT) * MS_RING);
    for (int c = (h - 1) + nh * wid; c < nchunks; c += nh * nw) {
      const int l0 = c * 64;
      // chunks with no line of this window: rows of H at or below it, columns of H up to its last
      // critical column
      if ((l0 >= rlo && l0 + 63 < n) || (l0 >= 2 * n && l0 + 63 <= 2 * n + chi)) continue;
      const int l = l0 + lane;
      zc* p = nullptr;
      long st = ldh;
      unsigned long long sg = kSign;
      if (l < rlo) {
        p = H + l + (long)kbase * ldh;
      } else if (l >= n && l < 2 * n) {
        p = Z + (l - n) + (long)kbase * ldz;
        st = ldz;
      } else if (l > 2 * n + chi && l < 3 * n) {
        p = H + (long)(l - 2 * n) * ldh + kbase;
        st = 1;
        sg = 0;
      }
      if (p) {
        if (full) ms_apply_line<true>(p, st, sg, 0, MS_LEN - 1, R);
        else ms_apply_line<false>(p, st, sg, rlo - kbase, rhi - kbase, R);
      }
    }
    wg_release_store(ctl + CTL_DONE + h - 1, w + 1);
  }
}

template <int NT>
__global__ __launch_bounds__(NT, 2) void k_hqr(int n, zc* H, int ldh, long sH, zc* Z, int ldz,
                                                     long sZ, zc* W, long sW, int* info, double* ring,
                                                     HqrTeam team) {
  constexpr int BR = QR_LD, BC = QR_W + 3;  // single-shift block: QR_W + 2 rows (leading dimension QR_LD)
  static_assert(BR * BC <= MS_LD * (MS_LEN + 1), "window buffers share one LDS array");
  __shared__ zc MB[MS_LD * (MS_LEN + 1)];  // window block (single-shift: BR x BC, multishift: LEN x LEN+1)
  __shared__ Rot rots[QR_W];
  __shared__ Rot2 rots2[(MS_S + 1) * MS_NB];  // row s: step s of the window; row MS_S: next window's step 0
  __shared__ zc shifts[MS_NB], betas[MS_NB];
  __shared__ zc aed_hv[2 * AED_W];
  __shared__ int aed_sh[2];
  static_assert(2 * AED_LD * AED_W <= MS_LD * (MS_LEN + 1), "the deflation window fits the window buffer");
  static_assert(AED_W + 1 < MS_NMIN, "the deflation window's spike lies inside the active block");
  __shared__ int s_k, s_m;
  __shared__ zc s_t, s_v1;
  __shared__ double s_v2;
  __shared__ int s_team;
  zc* B = MB;
  // teams: the workgroups of one matrix share blockIdx.x mod 8 (one XCD under round-robin dispatch)
  int b = blockIdx.x, role = 0;
  if (team.G > 1) {
    const int k = blockIdx.x / 8;
    b = (int)(blockIdx.x % 8) + 8 * (k / team.G);
    role = k % team.G;
    if (b >= team.batch) return;
  }
  if (role > 0) {
    hqr_help<NT>(n, H + sH * b, ldh, Z + sZ * b, ldz, team, b, role);
    return;
  }
  const bool teamed = team.G > 1;
  int* const ctl = teamed ? team.ctl + (long)b * CTL_INTS : nullptr;
  int wseq = 0;  // windows published to the helpers
  auto stop_team = [&]() {
    if (teamed && threadIdx.x == 0) st_release(ctl + 1, 1);
  };
  H += sH * b;
  Z += sZ * b;
  W += sW * b;
  Rot2* const rg0 = reinterpret_cast<Rot2*>(ring + (long)MS_RING * b);
  const int tid = threadIdx.x, nt = blockDim.x;
  auto Hp = [&](int r, int c) -> zc& { return H[(long)c * ldh + r]; };
  auto Zp = [&](int r, int c) -> zc& { return Z[(long)c * ldz + r]; };
  if (tid == 0) info[b] = 0;
  if (n == 1) {
    if (tid == 0) W[0] = Hp(0, 0);
    stop_team();
    return;
  }
  const int ilo = 0, ihi = n - 1;
  const double smlnum = kSafmin * ((double)n / kUlp);
  constexpr int kScratch = MS_LD * (MS_LEN + 1);
  if (n <= kScratch) {
    real_subdiag(H, ldh, Z, ldz, n, ilo, ihi, MB);
  } else {
    // the prefix-product scratch does not fit: zlahqr's sequential scan
    for (int i = ilo + 1; i <= ihi; ++i) {
      __syncthreads();
      const zc hi = Hp(i, i - 1);
      if (hi.im != 0.0) {
        zc sc = hi * (1.0 / abs1(hi));
        sc = conj(sc) * (1.0 / absz(sc));
        const double ah = absz(hi);
        __syncthreads();
        for (int c = i + tid; c < n; c += nt) Hp(i, c) = Hp(i, c) * sc;
        __syncthreads();
        for (int r = tid; r <= min(n - 1, i + 1); r += nt) Hp(r, i) = Hp(r, i) * conj(sc);
        for (int r = tid; r < n; r += nt) Zp(r, i) = Zp(r, i) * conj(sc);
        __syncthreads();
        if (tid == 0) Hp(i, i - 1) = mk(ah);
      }
    }
    __syncthreads();
  }

  const int itmax = 30 * max(10, ihi - ilo + 1);
  int kdefl = 0, its = 0, ms_fail = 0, lastL = -1, lastI = -1;
  bool aed_ready = false;  // the shifts of a deflation window taken at I = aed_I await their sweep
  int aed_I = -1, aed_nsh = 0;
  int I = ihi, L = ilo;
  while (I >= ilo) {
    // ---- look for a single small subdiagonal element (largest k in [L+1, I]) ----
    if (tid == 0) s_k = L;
    __syncthreads();
    int found = L;
    for (int k = L + 1 + tid; k <= I; k += nt)
      if (subdiag_negligible(H, ldh, k, ilo, ihi, smlnum)) found = max(found, k);
    if (found > L) atomicMax(&s_k, found);
    __syncthreads();
    L = s_k;
    if (L > ilo && tid == 0) Hp(L, L - 1) = mk(0.0);
    __syncthreads();
    if (L >= I) {
      // H(I, I-1) negligible: one eigenvalue converged
      if (tid == 0) W[I] = Hp(I, I);
      kdefl = 0;
      its = 0;
      I = L - 1;
      L = ilo;
      __syncthreads();
      continue;
    }
    if (++its > itmax) {
      if (tid == 0) info[b] = I + 1;
      stop_team();
      return;
    }
    const int m = I - L + 1;
    if (L != lastL || I != lastI) ms_fail = 0;
    if (m >= MS_NMIN && n <= kScratch && ms_fail < 3) {
      if (!(aed_ready && aed_I == I)) {
        // ================= aggressive early deflation =================
        const int nd = aed_window(H, ldh, Z, ldz, n, I, MB, aed_hv, shifts, aed_sh, smlnum);
        aed_nsh = (nd >= 0) ? aed_sh[1] : 0;
        // when many deflated, the next window is examined before a sweep is spent (zlaqr0's rule:
        // sweep only when at most 14 % of the window deflated)
        aed_ready = nd * 100 <= AED_W * 14;
        aed_I = I - max(nd, 0);
        if (nd > 0) real_subdiag(H, ldh, Z, ldz, n, L, I, MB);
        continue;
      }
      aed_ready = false;
      // ================= multishift sweep =================
      lastL = L;
      lastI = I;
      ++ms_fail;
      // the window's undeflated eigenvalues, else those of the trailing MS_NB x MS_NB block
      const int nb = (aed_nsh > 0) ? aed_nsh : MS_NB;
      if (tid < 64 && aed_nsh == 0) {
        zc* S = MB;
        for (int e = tid; e < nb * nb; e += 64) {
          const int r = e % nb, c = e / nb;
          S[e] = (r <= c + 1) ? Hp(I - nb + 1 + r, I - nb + 1 + c) : mk(0.0);
        }
        wave_sync();
        if (!small_eigvals_wave(S, nb, shifts, smlnum)) {
          wave_sync();
          for (int j = tid; j < nb; j += 64) shifts[j] = Hp(I - nb + 1 + j, I - nb + 1 + j);
        }
      }
      __syncthreads();
      aed_nsh = 0;
      const int T_total = (I - 1 - L) + 2 * (nb - 1) + 1;  // steps t = 0 .. T_total - 1
      for (int t0 = 0; t0 < T_total; t0 += MS_S) {
        const int kbase = L + t0 - 2 * (MS_NB - 1);
        const int rlo = max(kbase, L), rhi = min(kbase + MS_LEN - 1, I);
        const int clo = max(kbase - 1, L);
        auto MBp = [&](int r, int c) -> zc& { return MB[(c - kbase + 1) * MS_LD + (r - kbase)]; };
        const int nr = rhi - rlo + 1, nc = rhi - clo + 1;
        {
          // every lane's loads are issued before their LDS stores (a rolled loop would wait for each load
          // before issuing the next)
          // (four at a time: the kernel's registers are spoken for)
          const int tot = nr * nc;
          for (int e0 = 0; e0 < tot; e0 += 4 * nt) {
            zc wv[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
              const int e = min(e0 + tid + k * nt, tot - 1);
              wv[k] = Hp(rlo + e % nr, clo + e / nr);
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) {
              const int e = e0 + tid + k * nt;
              if (e < tot) MBp(rlo + e % nr, clo + e / nr) = wv[k];
            }
          }
        }
        // reflector of bulge j at step t (zero when the bulge is outside the active block) and the
        // value it leaves in the bulge column
        auto reflector = [&](int j, int t, Rot2& g, zc& beta) {
          const int k = L + t - 2 * j;
          g.t1 = g.v2 = g.t2 = mk(0.0);
          beta = mk(0.0);
          if (j < nb && k >= L && k <= I - 1) {
            zc v1, v2;
            if (k == L) {
              const zc h11s = MBp(L, L) - shifts[j], h21 = MBp(L + 1, L);
              const double rsc = rcp_nr(abs1(h11s) + abs1(h21));
              v1 = h11s * rsc;
              v2 = h21 * rsc;
            } else {
              v1 = MBp(k, k - 1);
              v2 = MBp(k + 1, k - 1);
            }
            g = make_reflector2(v1, v2);
            beta = v1;
          }
        };
        if (t0 == 0) {
          __syncthreads();
          if (tid < MS_NB) reflector(tid, 0, rots2[tid], betas[tid]);
        } else if (tid < MS_NB) {
          rots2[tid] = rots2[MS_S * MS_NB + tid];  // formed in the last step of the previous window
        }
        __syncthreads();
        const int nsteps = min(MS_S, T_total - t0);
        const bool wfull = kbase >= L && kbase + MS_LEN - 1 <= I;
        auto left_product = [&](int s, int t, int j, int x) {
          const int k = L + t - 2 * j, c = k + x;
          rot2_left(rots2[s * MS_NB + j], MBp(k, c), MBp(k + 1, c));
          if (x == 0 && k > L) {
            MBp(k, k - 1) = betas[j];
            MBp(k + 1, k - 1) = mk(0.0);
          }
        };
        auto right_product = [&](int s, int t, int j, int x) {
          const int k = L + t - 2 * j;
          rot2_right(rots2[s * MS_NB + j], MBp(rlo + x, k), MBp(rlo + x, k + 1));
        };
        for (int s = 0; s < nsteps; ++s) {
          const int t = t0 + s;
          // The bulges inside the active block at this step, j = jlo..jhi (k = L + t - 2j in [L, I-1]),
          // own contiguous runs of products: rhi - k + 1 left products each (columns k..rhi), and
          // k - rlo + 1 right products (rows rlo..k).  The runs are enumerated densely over the threads.
          const int jlo = max(0, (L + t - I + 2) >> 1), na = min(nb - 1, t >> 1) - jlo + 1;
          const int k0 = L + t - 2 * jlo;  // the leading active bulge's row
          // With all bulges inside a window that lies inside the active block, bulge j's runs are
          // 18 - s + 2j left and s + 31 - 2j right products, and with eight wavefronts wavefront w takes
          // bulges w and 15 - w, whose runs add up to 66 - 2s and 2s + 32 lines (with four, the dense
          // enumeration needs fewer passes).
          const bool paired = nt >= 512 && wfull && jlo == 0 && na == MS_NB;
          const int lane = tid & 63, nwav = nt >> 6;
          // left products on rows k, k+1 (columns k..rhi) and the bulge column of chase steps
          if (paired) {
            for (int pr = tid >> 6; pr < MS_NB / 2; pr += nwav) {
              const int j2 = MS_NB - 1 - pr, c1 = 18 - s + 2 * pr, c2 = 18 - s + 2 * j2;
              for (int xe = lane; xe < c1 + c2; xe += 64) {
                if (xe < c1) left_product(s, t, pr, xe);
                else left_product(s, t, j2, xe - c1);
              }
            }
          } else if (na > 0) {
            const int run0 = rhi - k0 + 1, total = na * run0 + na * (na - 1);
            for (int e = tid; e < total; e += nt) {
              int x;
              const int j = jlo + chase_run(e, na, run0, 1, x);
              left_product(s, t, j, x);
            }
          }
          __syncthreads();
          // right products on columns k, k+1 (rows rlo..min(k+2, I)).  Rows k+1 and k+2 of column k
          // are what bulge j's next reflector is formed from, and nothing else in this step writes
          // them, so lane j of the last wavefront (which takes no other rows) handles those two rows
          // and then forms that reflector: a bulge that enters at the next step reads H(L, L) and
          // H(L+1, L), final after the left products above.
          const int rl = tid - (nt - 64);
          if (rl >= 0 && rl < MS_NB) {
            // first, as the next step waits for it: rows k+1 and k+2 of bulge j and its next reflector
            // from the two values just formed (k+1 > L), or from H(L, L), H(L+1, L) when it enters
            const int j = rl;
            const int k = L + t - 2 * j;
            zc a1 = mk(0.0), a2 = mk(0.0);
            if (j < nb && k >= L && k <= I - 1) {
              const Rot2 g = rots2[s * MS_NB + j];
              zc b1 = MBp(k + 1, k + 1);
              a1 = MBp(k + 1, k);
              rot2_right(g, a1, b1);
              MBp(k + 1, k) = a1;
              MBp(k + 1, k + 1) = b1;
              if (k + 2 <= I) {
                zc b2 = MBp(k + 2, k + 1);
                a2 = MBp(k + 2, k);
                rot2_right(g, a2, b2);
                MBp(k + 2, k) = a2;
                MBp(k + 2, k + 1) = b2;
              }
            }
            Rot2 gn;
            gn.t1 = gn.v2 = gn.t2 = mk(0.0);
            zc bn = mk(0.0);
            const int kn = k + 1;
            if (j < nb && kn >= L && kn <= I - 1) {
              zc v1 = a1, v2 = a2;
              if (kn == L) {
                const zc h11s = MBp(L, L) - shifts[j], h21 = MBp(L + 1, L);
                const double rsc = rcp_nr(abs1(h11s) + abs1(h21));
                v1 = h11s * rsc;
                v2 = h21 * rsc;
              }
              gn = make_reflector2(v1, v2);
              bn = v1;
            }
            rots2[(s + 1) * MS_NB + j] = gn;
            betas[j] = bn;
          }
          if (paired) {
            for (int pr = tid >> 6; pr < MS_NB / 2; pr += nwav) {
              const int j2 = MS_NB - 1 - pr, c1 = s + 31 - 2 * pr, c2 = s + 31 - 2 * j2;
              for (int xe = lane; xe < c1 + c2; xe += 64) {
                if (xe < c1) right_product(s, t, pr, xe);
                else right_product(s, t, j2, xe - c1);
              }
            }
          } else if (na > 0) {
            const int run0 = k0 - rlo + 1, total = na * run0 - na * (na - 1);
            for (int e = tid; e < total; e += nt) {
              int x;
              const int j = jlo + chase_run(e, na, run0, -1, x);
              right_product(s, t, j, x);
            }
          }
          __syncthreads();
        }
        if (nsteps < MS_S) {
          // the last window of a sweep: its unused steps become identity rotations
          for (int e = tid; e < (MS_S - nsteps) * MS_NB; e += nt) {
            Rot2& g = rots2[nsteps * MS_NB + e];
            g.t1 = g.v2 = g.t2 = mk(0.0);
          }
          __syncthreads();
        }
        const bool full = kbase >= L && kbase + MS_LEN - 1 <= I;
        // last critical column: the next window's block reaches 16 columns further (none after the last)
        const int chi = (t0 + MS_S < T_total) ? min(rhi + MS_S, I) : rhi;
        // the window's rotations: the leader's own ring, or the window's slot in a team (wseq is carried
        // by the iteration loop, whose trip count the compiler cannot prove uniform; the slot is)
        Rot2* const rg = teamed ? reinterpret_cast<Rot2*>(
                                      team.slots +
                                      (long)__builtin_amdgcn_readfirstlane(b * NSLOT + wseq % NSLOT) * MS_RING)
                                : rg0;
        if (teamed) {
          // the slot of this window is free once every helper has finished the window it held
          if (wseq >= NSLOT && !team_wait(team, ctl, wseq - NSLOT + 1, &s_team)) {
            if (tid == 0) info[b] = n + 1;
            stop_team();
            return;
          }
          if (tid == 0) {
            int* d = team.desc + ((long)b * NSLOT + wseq % NSLOT) * DESC_INTS;
            d[0] = kbase;
            d[1] = rlo;
            d[2] = rhi;
            d[3] = full ? 1 : 0;
            d[4] = chi;
          }
        }
        for (int e = tid; e < MS_S * MS_NB; e += nt) rg[(e % MS_NB) * MS_S + e / MS_NB] = rots2[e];
        for (int e = tid; e < nr * nc; e += nt) {
          const int r = rlo + e % nr, c = clo + e / nr;
          Hp(r, c) = MBp(r, c);
        }
        // the ring stores reach L2 before the barrier; the scalar cache then drops what it holds of
        // the previous window's rotations (s_waitcnt encodings of gfx9: 0x0F70 vmcnt(0), 0xC07F lgkmcnt(0))
        __builtin_amdgcn_s_waitcnt(0x0F70);
        __syncthreads();
        if (teamed) {
          team_publish(ctl, wseq++);
          // the critical columns hold the helpers' share of the window before: wait for it
          if (wseq >= 2 && !team_wait(team, ctl, wseq - 1, &s_team)) {
            if (tid == 0) info[b] = n + 1;
            stop_team();
            return;
          }
        }
        if (!teamed) {
          __builtin_amdgcn_s_dcache_inv();
          __builtin_amdgcn_s_waitcnt(0xC07F);
          // the waves fetch the ring's 64-byte lines into the scalar cache together (one dword of
          // each, all loads in flight), so the deferred products start on hits
          constexpr int kLines = MS_RING * 8 / 64, kWaves = NT / 64;
          static_assert(kLines % kWaves == 0, "ring lines split evenly over the waves");
          typedef __attribute__((address_space(4))) const unsigned KU32;
          const unsigned* rp = reinterpret_cast<const unsigned*>(rg);
          asm volatile("" : "+s"(rp));
          const KU32* ku = (const KU32*)rp;
          const int w = __builtin_amdgcn_readfirstlane(tid >> 6);
          unsigned acc = 0;
#pragma unroll
          for (int i = 0; i < kLines / kWaves; ++i) acc ^= ku[(i * kWaves + w) * 16];
          asm volatile("" ::"s"(acc));
        }
        // deferred products: columns right of the window (left rotations on rows kbase + q),
        // rows above it and all rows of Z (right rotations on columns kbase + q; the helpers' share
        // in a team)
        const int nfr = teamed ? chi - rhi : n - 1 - rhi, nfa = rlo, nz = n;
        const int lines = teamed ? 0 : nfr + nfa + nz;
        if (teamed) {
          // the critical columns, 16 lanes each
          for (int ln = tid >> 4; ln < nfr; ln += nt >> 4) {
            zc* const p = H + (long)(rhi + 1 + ln) * ldh + kbase;
            if (full) ms_apply_column_lanes<true>(p, 0, MS_LEN - 1, rots2, tid & 15);
            else ms_apply_column_lanes<false>(p, rlo - kbase, rhi - kbase, rots2, tid & 15);
          }
        }
        constexpr unsigned long long kSign = 1ull << 63;
        for (int ln = tid; ln < lines; ln += nt) {
          const KDouble* R = scalar_view(reinterpret_cast<const double*>(rg));
          zc* p;
          long st;
          unsigned long long sg;
          if (ln < nfr) {
            p = H + (long)(rhi + 1 + ln) * ldh + kbase;
            st = 1;
            sg = 0;
          } else if (ln < nfr + nfa) {
            p = H + (ln - nfr) + (long)kbase * ldh;
            st = ldh;
            sg = kSign;
          } else {
            p = Z + (ln - nfr - nfa) + (long)kbase * ldz;
            st = ldz;
            sg = kSign;
          }
          if (full) ms_apply_line<true>(p, st, sg, 0, MS_LEN - 1, R);
          else ms_apply_line<false>(p, st, sg, rlo - kbase, rhi - kbase, R);
        }
        __syncthreads();
      }
      if (teamed && !team_wait(team, ctl, wseq, &s_team)) {
        if (tid == 0) info[b] = n + 1;
        stop_team();
        return;
      }
      real_subdiag(H, ldh, Z, ldz, n, L, I, MB);
      continue;
    }
    // ================= single-shift iteration (zlahqr) =================
    ++kdefl;
    if (tid == 0) {
      zc t;
      if (kdefl % 20 == 0) {
        t = mk(0.75 * fabs(Hp(I, I - 1).re)) + Hp(I, I);
      } else if (kdefl % 10 == 0) {
        t = mk(0.75 * fabs(Hp(L + 1, L).re)) + Hp(L, L);
      } else {
        t = wilkinson_shift(Hp(I - 1, I - 1), Hp(I - 1, I), Hp(I, I - 1), Hp(I, I));
      }
      s_t = t;
      s_m = L;
    }
    __syncthreads();
    const zc tshift = s_t;
    // ---- look for two consecutive small subdiagonal elements (largest m in [L+1, I-1]) ----
    int mf = L;
    for (int mm = L + 1 + tid; mm <= I - 1; mm += nt)
      if (two_small(H, ldh, mm, tshift)) mf = max(mf, mm);
    if (mf > L) atomicMax(&s_m, mf);
    __syncthreads();
    const int M = s_m;
    if (tid == 0) {
      const zc h11 = Hp(M, M);
      zc h11s = h11 - tshift;
      double h21 = Hp(M + 1, M).re;
      const double s = abs1(h11s) + fabs(h21);
      s_v1 = h11s * (1.0 / s);
      s_v2 = h21 / s;
    }
    __syncthreads();

    // ---- single-shift QR sweep k = M .. I-1, in windows ----
    int ks = M;
    while (ks < I) {
      // the first rotation of a sweep started at M > L is followed by a diagonal scaling;
      // it gets a window of its own
      const int wlen = (ks == M && M > L) ? 1 : min(QR_W, I - ks);
      const int ke = ks + wlen;                     // rotations ks .. ke-1
      const int re = min(ke + 1, I);                // last block row / column
      const int cs = (ks > M) ? ks - 1 : ks;        // first block column
      const int nr = re - ks + 1, nc = re - cs + 1;
      for (int e = tid; e < nr * nc; e += nt) {
        int r = e % nr, c = e / nr;
        B[c * BR + r] = Hp(ks + r, cs + c);
      }
      __syncthreads();
      auto Bp = [&](int r, int c) -> zc& { return B[(c - cs) * BR + (r - ks)]; };
      // the block is at most QR_W + 3 wide: one wavefront chases the bulge through it, with
      // wavefront-level ordering between the phases
      static_assert(QR_W + 3 <= 64, "one lane per block row / column");
      if (tid < 64) {
        for (int k = ks; k < ke; ++k) {
          zc v1;
          zc v2c;
          if (k == M) {
            v1 = s_v1;
            v2c = mk(s_v2);
          } else {
            v1 = Bp(k, k - 1);
            v2c = Bp(k + 1, k - 1);
          }
          // every lane forms the same reflector from the same LDS values; column k-1 is rewritten
          // only in the right phase below
          const Rot g = make_reflector(v1, v2c);
          // left: rows k, k+1, cols k..re
          if (k + tid <= re) rot_left(g, Bp(k, k + tid), Bp(k + 1, k + tid));
          wave_sync();
          // right: cols k, k+1, rows ks..min(k+2, I)
          if (ks + tid <= min(k + 2, I)) rot_right(g, Bp(ks + tid, k), Bp(ks + tid, k + 1));
          if (tid == 0) {
            if (k > M) {
              Bp(k, k - 1) = v1;
              Bp(k + 1, k - 1) = mk(0.0);
            }
            rots[k - ks] = g;
          }
          wave_sync();
        }
      }
      __syncthreads();
      for (int e = tid; e < nr * nc; e += nt) {
        int r = e % nr, c = e / nr;
        Hp(ks + r, cs + c) = B[c * BR + r];
      }
      // deferred: far-right columns (re, n) on rows ks..ke; far-above rows [0, ks) on cols ks..ke;
      // Schur vectors Z(:, ks..ke)
      const int nfr = n - 1 - re, nfa = ks, nz = n;
      const int lines = nfr + nfa + nz;
      for (int ln = tid; ln < lines; ln += nt) {
        if (ln < nfr) {
          ss_apply_line<true>(H + (long)(re + 1 + ln) * ldh + ks, 1, rots, wlen);
        } else {
          const bool isz = ln >= nfr + nfa;
          const int r = isz ? ln - nfr - nfa : ln - nfr;
          ss_apply_line<false>(isz ? Z + (long)ks * ldz + r : H + (long)ks * ldh + r, isz ? ldz : ldh, rots, wlen);
        }
      }
      __syncthreads();
      if (ks == M && M > L) {
        // extra scaling keeping H(M, M-1) real (zlahqr)
        const Rot g = rots[0];
        zc temp = mk(1.0) - g.t1;
        temp = temp * (1.0 / absz(temp));
        if (tid == 0) {
          Hp(M + 1, M) = Hp(M + 1, M) * conj(temp);
          if (M + 2 <= I) Hp(M + 2, M + 1) = Hp(M + 2, M + 1) * temp;
        }
        __syncthreads();
        for (int J = M; J <= I; ++J) {
          if (J == M + 1) continue;
          for (int c = J + 1 + tid; c < n; c += nt) Hp(J, c) = Hp(J, c) * temp;
          __syncthreads();
          for (int r = tid; r < J; r += nt) Hp(r, J) = Hp(r, J) * conj(temp);
          for (int r = tid; r < n; r += nt) Zp(r, J) = Zp(r, J) * conj(temp);
          __syncthreads();
        }
      }
      ks = ke;
    }
    // ---- ensure H(I, I-1) is real ----
    const zc tmp = Hp(I, I - 1);
    __syncthreads();
    if (tmp.im != 0.0) {
      const double rt = absz(tmp);
      const zc tt = tmp * (1.0 / rt);
      for (int c = I + 1 + tid; c < n; c += nt) Hp(I, c) = Hp(I, c) * conj(tt);
      for (int r = tid; r < I; r += nt) Hp(r, I) = Hp(r, I) * tt;
      for (int r = tid; r < n; r += nt) Zp(r, I) = Zp(r, I) * tt;
      __syncthreads();
      if (tid == 0) Hp(I, I - 1) = mk(rt);
    }
    __syncthreads();
  }
  stop_team();
}

// ------------------------------------------------------------------------------------------------
// 4. Eigenvectors of the upper triangular T (ztrevc, side 'R'): column ki of X solves
//    (T(0:ki,0:ki) - T(ki,ki)) x = -T(0:ki, ki), x(ki) = 1, x(ki+1:) = 0.  One wavefront per ki.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_trevc(int n, const zc* T, int ldt, long sT, zc* X, int ldx,
                                               long sX, double smlnum) {
  extern __shared__ zc xs[];  // (blockDim.x / 64) wavefronts x n
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  const int ki = blockIdx.x * (blockDim.x >> 6) + wid;
  const int b = blockIdx.y;
  T += sT * b;
  X += sX * b;
  if (ki >= n) return;
  zc* x = xs + wid * n;
  const zc lam = T[(long)ki * ldt + ki];
  const double smin = fmax(kUlp * abs1(lam), smlnum);
  for (int r = lane; r < ki; r += 64) x[r] = -T[(long)ki * ldt + r];
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  for (int j = ki - 1; j >= 0; --j) {
    zc d = T[(long)j * ldt + j] - lam;
    if (abs1(d) < smin) d = mk(smin);
    const zc xj = zdiv(x[j], d);
    // every lane computed xj identically from the same LDS value; rows 0..j-1 updated in parallel
    const zc* tcol = T + (long)j * ldt;
    for (int r = lane; r < j; r += 64) x[r] = x[r] - xj * tcol[r];
    if (lane == 0) x[j] = xj;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
  }
  zc* xc = X + (long)ki * ldx;
  for (int r = lane; r < n; r += 64) xc[r] = (r < ki) ? x[r] : mk(r == ki ? 1.0 : 0.0);
}

// ------------------------------------------------------------------------------------------------
// 5. zgeev normalisation: unit 2-norm, then rotate so that the largest |v_k|^2 is real.
// ------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_normalize(int n, zc* V, int ldv, long sV) {
  __shared__ zc red[4];
  __shared__ double sv[4];
  __shared__ int si[4];
  const int col = blockIdx.x, b = blockIdx.y;
  zc* v = V + sV * b + (long)col * ldv;
  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  double ss = 0.0;
  for (int r = tid; r < n; r += blockDim.x) ss += norm2(v[r]);
  const double nrm = sqrt(block_sum(mk(ss), red).re);
  const double scl = 1.0 / nrm;
  double best = -1.0;
  int bi = n;
  for (int r = tid; r < n; r += blockDim.x) {
    zc e = v[r] * scl;
    double m2 = e.re * e.re + e.im * e.im;
    if (m2 > best) { best = m2; bi = r; }
  }
  for (int off = 32; off > 0; off >>= 1) {
    double ov = __shfl_xor(best, off);
    int oi = __shfl_xor(bi, off);
    if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
  }
  if (lane == 0) { sv[wid] = best; si[wid] = bi; }
  __syncthreads();
  double bv = sv[0];
  int k = si[0];
  for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
    if (sv[w] > bv || (sv[w] == bv && si[w] < k)) { bv = sv[w]; k = si[w]; }
  const zc vk = v[k] * scl;
  const zc rot = conj(vk) * (1.0 / sqrt(bv));
  __syncthreads();
  for (int r = tid; r < n; r += blockDim.x) {
    zc e = (v[r] * scl) * rot;
    if (r == k) e.im = 0.0;
    v[r] = e;
  }
}

// Workgroups per matrix of the QR launch: a leader plus helpers, as many as the tuning table gives for
// (n, batch) ("qr_team"; -1 or no rule: enough that each helper wavefront takes one 64-line chunk of the
// rows and columns of H and Z per window), while the teams still fit the device once.
int qr_team_size(int n, int batch, int cus) {
  const long want = tune_lookup("qr_team", n, batch, -1);
  if (want == 0 || n < 2 * MS_NMIN || n > MS_LD * (MS_LEN + 1)) return 1;
  const int chunks = (3 * n + 63) / 64;
  int helpers = (chunks + EIG_THREADS / 64 - 1) / (EIG_THREADS / 64);
  if (want > 0) helpers = (int)std::min<long>(want, TEAM_MAX - 1);
  helpers = std::min(helpers, TEAM_MAX - 1);
  helpers = std::min(helpers, cus / std::max(1, batch) - 1);
  return helpers >= 1 ? helpers + 1 : 1;
}

}  // namespace

void EigWork::ensure(int n, int batch) {
  const size_t nn = (size_t)n * n * batch;
  Z.ensure(nn);
  X.ensure(nn);
  tau.ensure((size_t)n * batch);
  info.ensure((size_t)batch);
  const size_t npan = (size_t)std::max(1, n / HB + 1);
  Y.ensure((size_t)n * HB * batch);
  T.ensure(npan * HB * HB * batch);
  W.ensure((size_t)2 * n * HB * batch);
  ring.ensure((size_t)MS_RING * batch);
}

void EigWork::ensure_team(int batch) {
  team_ctl.ensure((size_t)CTL_INTS * batch);
  team_slots.ensure((size_t)NSLOT * MS_RING * batch);
  team_desc.ensure((size_t)NSLOT * DESC_INTS * batch);
}

void zgeev_batched(int n, zc* A, int lda, long sA, zc* W, long sW, zc* V, int ldv, long sV, int batch,
                   EigWork& work, int* info, hipStream_t s, bool teams) {
  if (n <= 0 || batch <= 0) return;
  CWX_CHECK(batch <= 65535, "zgeev: batch too large");
  work.ensure(n, batch);
  const long nn = (long)n * n;
  const size_t sh2 = sizeof(zc) * (2 * (size_t)n + HH_THREADS);
  const size_t lds = lds_limit();
  CWX_CHECK(sh2 + sizeof(zc) * (HH_THREADS / 64 + 2 * HB) <= lds, "zgeev: n too large for the LDS work vectors");
  // ---- Hessenberg reduction: blocked panels, then the last columns unblocked ----
  const long npan_max = n / HB + 1;
  const long sY = (long)n * HB, sT = npan_max * HB * HB, sWk = 2L * n * HB;
  zc* Vd = work.X.p;  // dense V of every panel (n x n per matrix); X is needed only after the QR
  zc* Wk = work.W.p;
  zc* Wk2 = work.W.p + (long)n * HB;
  auto S = [](const zc* p, long st) { return strided<zc>(p, st); };
  int dev = 0, cus = 0;
  CWX_HIP(hipGetDevice(&dev));
  CWX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  // panel teams: when the matrices leave most of the device idle, helper workgroups take rows of each
  // panel column's trailing-matrix product ("lahr2_team" helpers; by default as many as fit)
  Team pteam;
  if (teams && batch < cus && n - 1 - HB_NX > 0) {
    const long want = tune_lookup("lahr2_team", n, batch, -1);
    int G = want >= 0 ? (int)want + 1 : 8;
    G = std::min({G, TEAM_MAX, cus / (8 * ((batch + 7) / 8))});
    int per_cu = 0;
    CWX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_lahr2, HH_THREADS, sh2));
    if (G > 1 && (long)per_cu * cus >= 8L * ((batch + 7) / 8) * G) {
      work.ensure_team(batch);
      int rate_khz = 0;
      CWX_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev));
      pteam.G = G;
      pteam.batch = batch;
      pteam.ctl = work.team_ctl.p;
      pteam.limit = (unsigned long long)std::max(rate_khz, 1) * 1000ull * 20ull;  // 20 s
      hipLaunchKernelGGL(k_fill_i32, dim3(cdiv((size_t)CTL_INTS * batch, 256)), dim3(256), 0, s,
                         (long)CTL_INTS * batch, pteam.ctl);
    }
  }
  int i0 = 0, npan = 0;
  for (; i0 < n - 1 - HB_NX; i0 += HB, ++npan) {
    const int nb = std::min(HB, n - 1 - i0);
    const int m = n - i0 - 1;
    const int nc = n - i0 - nb;
    zc* Vp = Vd + (long)npan * HB * n;
    zc* Tp = work.T.p + (long)npan * HB * HB;
    bool launched = false;
    if (pteam.G > 1) {
      int n_ = n, i0_ = i0, nb_ = nb, lda_ = lda, ldv = n, ldy = n;
      long sA_ = sA, sV = nn, sY_ = sY, sT_ = sT;
      zc *Ap = A, *Vq = Vp, *Yq = work.Y.p + i0 + 1, *Tq = Tp, *tq = work.tau.p;
      void* args[] = {&n_, &i0_, &nb_, &Ap, &lda_, &sA_, &Vq, &ldv, &sV, &Yq, &ldy, &sY_, &Tq, &sT_, &tq, &pteam};
      const unsigned grid = (unsigned)(8 * ((batch + 7) / 8) * pteam.G);
      if (hipLaunchCooperativeKernel((const void*)k_lahr2, dim3(grid), dim3(HH_THREADS), args, sh2, s) == hipSuccess)
        launched = true;
      else {
        (void)hipGetLastError();
        pteam = Team();  // refused (for instance while another process holds the device): one workgroup
      }
    }
    if (!launched)
      hipLaunchKernelGGL(k_lahr2, dim3(batch), dim3(HH_THREADS), sh2, s, n, i0, nb, A, lda, sA, Vp, n, nn,
                         work.Y.p + i0 + 1, n, sY, Tp, sT, work.tau.p, Team());
    // Y(0:i0+1, :) = A(0:i0+1, i0+1:n) V T
    zgemm_set(OpN, OpN, i0 + 1, nb, m, mk(1.0), S(A + (long)(i0 + 1) * lda, sA), lda, S(Vp, nn), n, mk(0.0),
              S(Wk, sWk), n, batch, s);
    zgemm_set(OpN, OpN, i0 + 1, nb, nb, mk(1.0), S(Wk, sWk), n, S(Tp, sT), HB, mk(0.0), S(work.Y.p, sY), n, batch,
              s);
    // right update of the trailing columns: A(:, i0+nb:n) -= Y V(i0+nb:n, :)^H
    zgemm_set(OpN, OpC, n, nc, nb, mk(-1.0), S(work.Y.p, sY), n, S(Vp + (nb - 1), nn), n, mk(1.0),
              S(A + (long)(i0 + nb) * lda, sA), lda, batch, s);
    // rows 0..i0 of the panel columns: A(0:i0+1, i0+1:i0+nb) -= Y(0:i0+1, 0:nb-1) V1^H
    if (nb > 1)
      zgemm_set(OpN, OpC, i0 + 1, nb - 1, nb - 1, mk(-1.0), S(work.Y.p, sY), n, S(Vp, nn), n, mk(1.0),
                S(A + (long)(i0 + 1) * lda, sA), lda, batch, s);
    // left update: A(i0+1:n, i0+nb:n) = (I - V T^H V^H) A(i0+1:n, i0+nb:n)
    zc* Asub = A + (long)(i0 + nb) * lda + i0 + 1;
    zgemm_set(OpC, OpN, nb, nc, m, mk(1.0), S(Vp, nn), n, S(Asub, sA), lda, mk(0.0), S(Wk, sWk), HB, batch, s);
    zgemm_set(OpC, OpN, nb, nc, nb, mk(1.0), S(Tp, sT), HB, S(Wk, sWk), HB, mk(0.0), S(Wk2, sWk), HB, batch, s);
    zgemm_set(OpN, OpN, m, nc, nb, mk(-1.0), S(Vp, nn), n, S(Wk2, sWk), HB, mk(1.0), S(Asub, sA), lda, batch, s);
  }
  if (i0 < n - 2) hipLaunchKernelGGL(k_gehrd, dim3(batch), dim3(HH_THREADS), sh2, s, n, i0, A, lda, sA, work.tau.p);
  // ---- Q: the unblocked reflectors backwards, then the panels' block reflectors backwards ----
  hipLaunchKernelGGL(k_unghr, dim3(batch), dim3(HH_THREADS), sh2, s, n, i0, A, lda, sA, work.tau.p, work.Z.p, n,
                     nn);
  for (int p = npan - 1; p >= 0; --p) {
    const int j0 = p * HB, nb = std::min(HB, n - 1 - j0), m = n - j0 - 1;
    zc* Vp = Vd + (long)p * HB * n;
    zc* Tp = work.T.p + (long)p * HB * HB;
    zc* Zsub = work.Z.p + (long)(j0 + 1) * n + j0 + 1;
    zgemm_set(OpC, OpN, nb, m, m, mk(1.0), S(Vp, nn), n, S(Zsub, nn), n, mk(0.0), S(Wk, sWk), HB, batch, s);
    zgemm_set(OpN, OpN, nb, m, nb, mk(1.0), S(Tp, sT), HB, S(Wk, sWk), HB, mk(0.0), S(Wk2, sWk), HB, batch, s);
    zgemm_set(OpN, OpN, m, m, nb, mk(-1.0), S(Vp, nn), n, S(Wk2, sWk), HB, mk(1.0), S(Zsub, nn), n, batch, s);
  }
  hipLaunchKernelGGL(k_clear_below_sub, dim3(cdiv(n * n, 256 * 4), batch), dim3(256), 0, s, n, A, lda, sA);
  HqrTeam team;
  team.G = (teams && batch < cus) ? qr_team_size(n, batch, cus) : 1;
  if (team.G > 1) {
    // a cooperative launch keeps every workgroup of a team resident; without room for all of them the
    // leaders work alone
    int per_cu = 0;
    CWX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_hqr<EIG_THREADS>, EIG_THREADS, 0));
    // as many helpers as fit beside the leaders
    const long slots = (long)per_cu * cus, lead = 8L * ((batch + 7) / 8);
    if (slots < lead * team.G) team.G = (int)std::max<long>(1, slots / lead);
  }
  if (team.G > 1) {
    work.ensure_team(batch);
    int rate_khz = 0;
    CWX_HIP(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev));
    team.batch = batch;
    team.ctl = work.team_ctl.p;
    team.slots = work.team_slots.p;
    team.desc = work.team_desc.p;
    team.limit = (unsigned long long)std::max(rate_khz, 1) * 1000ull * 20ull;  // 20 s
    hipLaunchKernelGGL(k_fill_i32, dim3(cdiv((size_t)CTL_INTS * batch, 256)), dim3(256), 0, s,
                       (long)CTL_INTS * batch, team.ctl);
    int n_ = n, lda_ = lda, ldz = n;
    long sA_ = sA, sZ = nn, sW_ = sW;
    zc* Ap = A;
    zc* Zp = work.Z.p;
    zc* Wp = W;
    int* infop = info;
    double* ringp = work.ring.p;
    void* args[] = {&n_, &Ap, &lda_, &sA_, &Zp, &ldz, &sZ, &Wp, &sW_, &infop, &ringp, &team};
    const unsigned grid = (unsigned)(8 * ((batch + 7) / 8) * team.G);
    if (hipLaunchCooperativeKernel((const void*)k_hqr<EIG_THREADS>, dim3(grid), dim3(EIG_THREADS), args, 0, s) !=
        hipSuccess) {
      // the runtime may refuse cooperative launches (for instance while another process holds the
      // device); the leaders then work alone
      (void)hipGetLastError();
      team = HqrTeam();
    }
  }
  if (team.G > 1) {
  } else if (tune_lookup("qr_threads", n, batch, batch >= 2 * cus ? EIG_THREADS_SMALL : EIG_THREADS) ==
             EIG_THREADS_SMALL) {
    hipLaunchKernelGGL(k_hqr<EIG_THREADS_SMALL>, dim3(batch), dim3(EIG_THREADS_SMALL), 0, s, n, A, lda, sA,
                       work.Z.p, n, nn, W, sW, info, work.ring.p, team);
  } else {
    hipLaunchKernelGGL(k_hqr<EIG_THREADS>, dim3(batch), dim3(EIG_THREADS), 0, s, n, A, lda, sA, work.Z.p, n,
                       nn, W, sW, info, work.ring.p, team);
  }
  const double smlnum = kSafmin * ((double)n / kUlp);
  // one wavefront per eigenvector, as many per workgroup (up to 4) as fit 64 KiB of LDS
  const int waves = (int)std::max<size_t>(1, std::min<size_t>(4, (64 * 1024) / (sizeof(zc) * n)));
  CWX_CHECK(sizeof(zc) * (size_t)n <= lds, "zgeev: n too large for the LDS eigenvector buffer");
  hipLaunchKernelGGL(k_trevc, dim3(cdiv(n, waves), batch), dim3(64 * waves),
                     sizeof(zc) * waves * (size_t)n, s, n, A, lda, sA, work.X.p, n, nn, smlnum);
  zgemm_batched(OpN, OpN, n, n, n, mk(1.0), work.Z.p, n, nn, work.X.p, n, nn, mk(0.0), V, ldv, sV,
                batch, s);
  hipLaunchKernelGGL(k_normalize, dim3(n, batch), dim3(256), 0, s, n, V, ldv, sV);
  CWX_HIP(hipGetLastError());
}

}  // namespace cwx
