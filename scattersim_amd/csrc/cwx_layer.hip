// scatterlayer on the GPU: plane waves through a stack of homogeneous layers (ScatterLayer.cpp:110-420),
// every incident wave of a focused beam solved at once.
//
// Per incident wave the reference assembles a 6L x 6L system: the incident amplitude in layer 0 and no
// reflected wave in the last layer (ScatterLayer.cpp:152-165), Gauss's law for every other wave
// (167-184) and the continuity of Ex, Ey, Hx, Hy at every interface (186-232).  Unknown 6l + 3d + c is
// component c of the transmitted (d = 0) or reflected (d = 1) wave of layer l.  The solution becomes a
// reflected and a transmitted plane wave per interface, each wound to its boundary (240-266, 390-420).
//
// Stacks of up to 10 layers (6L <= 64) are solved by one wavefront per incident wave: lane r owns row r
// of the augmented system, eliminates with partial pivoting (largest |re| + |im|, lowest row on ties) and
// back-substitutes, then the lanes write the wound records.  Up to 8 layers the rows live in registers
// (k_layer_reg, one instantiation per layer count); 9 and 10 layers keep them in LDS (k_layer_fused).
// Taller stacks assemble the systems in HBM and go through the batched LU of this library.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "cwx_kernels.h"
#include "cwx_layer.h"

namespace cwx {
namespace {

constexpr int LAYER_FUSED_MAXN = 64;  // 6L rows, one per lane

struct LayerArgs {
  int L, n;                // layers and unknowns (6L)
  const zc* ri;            // L complex refractive indices (the first real)
  const double* z;         // L - 1 boundary positions
  double kv;               // vacuum wavenumber 2 pi / lambda
  long P;                  // incident waves
  const zc* inc;           // P records {E0[3], k[3]}
  zc* refl;                // P x (L-1) records: reflected wave of layer l at boundary l
  zc* trans;               // P x (L-1) records: transmitted wave of layer l+1 at boundary l
  int* nsing;              // number of systems with an exactly zero pivot
};

// Unit propagation direction of an incident record (planewave::getDirection of a real k).
__device__ inline void inc_direction(const zc* rec, double d[3]) {
  const double kx = rec[3].re, ky = rec[4].re, kz = rec[5].re;
  const double m = sqrt(kx * kx + ky * ky + kz * kz);
  d[0] = kx / m;
  d[1] = ky / m;
  d[2] = kz / m;
}

// sz of layer l: sqrt(ri^2 - sx^2 - sy^2) with s = d n0 (ScatterLayer.cpp:132-146), in the reference's
// order of operations.
__device__ inline zc layer_sz(const LayerArgs& a, int l, double sx, double sy) {
  const zc r = a.ri[l];
  return zsqrt((r * r - mk(sx * sx)) - mk(sy * sy));
}

// The non-zero entries of row r of one incidence's system, handed to put(col, value); rows 0-2 also
// carry the right-hand side (rhs(value)).
template <class Put, class Rhs>
__device__ void layer_row(const LayerArgs& a, int r, const double d[3], const zc* E0, Put put, Rhs rhs) {
  const int L = a.L;
  const double n0 = a.ri[0].re;
  const double sx = d[0] * n0, sy = d[1] * n0;
  auto col = [](int l, int dir, int c) { return 6 * l + 3 * dir + c; };
  if (r < 3) {                               // layer 0 transmits the incident amplitude
    put(col(0, 0, r), mk(1.0));
    rhs(E0[r]);
    return;
  }
  if (r < 6) {                               // no reflected wave in the last layer
    put(col(L - 1, 1, r - 3), mk(1.0));
    return;
  }
  if (r < 6 + (L - 1)) {                     // Gauss's law, reflected wave of layer l = 0 .. L-2
    const int l = r - 6;
    put(col(l, 1, 0), mk(d[0]));
    put(col(l, 1, 1), mk(d[1]));
    put(col(l, 1, 2), -layer_sz(a, l, sx, sy));
    return;
  }
  if (r < 6 + 2 * (L - 1)) {                 // Gauss's law, transmitted wave of layer l = 1 .. L-1
    const int l = r - (6 + L - 2);
    put(col(l, 0, 0), mk(d[0]));
    put(col(l, 0, 1), mk(d[1]));
    put(col(l, 0, 2), layer_sz(a, l, sx, sy));
    return;
  }
  const int q = r - (6 + 2 * (L - 1));
  const int l = q >> 2, e = q & 3;
  const double zn = (l == 0) ? a.z[0] : a.z[l] - a.z[l - 1];
  const double zp = (l == L - 2) ? 0.0 : a.z[l] - a.z[l + 1];
  const zc s0 = layer_sz(a, l, sx, sy), s1 = layer_sz(a, l + 1, sx, sy);
  const double k = a.kv;
  // exp(i k sz zn) and exp(-i k sz' zp)
  const zc en = zexp(mk(-k * s0.im * zn, k * s0.re * zn));
  const zc ep = zexp(mk(k * s1.im * zp, -k * s1.re * zp));
  if (e == 0 || e == 1) {                    // Ex (equation 8) or Ey (equation 9) continuity
    put(col(l, 0, e), en);
    put(col(l, 1, e), mk(1.0));
    put(col(l + 1, 0, e), mk(-1.0));
    put(col(l + 1, 1, e), -ep);
  } else if (e == 2) {                       // equation 10
    put(col(l, 0, 2), d[1] * en);
    put(col(l, 0, 1), -(s0 * en));
    put(col(l, 1, 2), mk(d[1]));
    put(col(l, 1, 1), s0);
    put(col(l + 1, 0, 2), mk(-d[1]));
    put(col(l + 1, 0, 1), s1);
    put(col(l + 1, 1, 2), -d[1] * ep);
    put(col(l + 1, 1, 1), -(s1 * ep));
  } else {                                   // equation 11
    put(col(l, 0, 0), s0 * en);
    put(col(l, 0, 2), -d[0] * en);
    put(col(l, 1, 0), -s0);
    put(col(l, 1, 2), mk(-d[0]));
    put(col(l + 1, 0, 0), -s1);
    put(col(l + 1, 0, 2), mk(d[0]));
    put(col(l + 1, 1, 0), s1 * ep);
    put(col(l + 1, 1, 2), d[0] * ep);
  }
}

// Record q of incidence p (q < 2(L-1): even q reflected, odd q transmitted, boundary q / 2) from the
// solution x (ScatterLayer.cpp:240-266, wound by planewave::wind(0, 0, -z_l) at 396-400).
__device__ void layer_record(const LayerArgs& a, long p, int q, const double d[3], const zc* x) {
  const int l = q >> 1;
  const bool t = q & 1;
  const double n0 = a.ri[0].re;
  const double sx = d[0] * n0, sy = d[1] * n0;
  const zc sz = layer_sz(a, t ? l + 1 : l, sx, sy);
  const zc kz = t ? a.kv * sz : -(a.kv * sz);
  const int base = t ? 6 * (l + 1) : 6 * l + 3;
  // E0 *= exp(i kz (-z_l))
  const double zl = a.z[l];
  const zc w = zexp(mk(kz.im * zl, -kz.re * zl));
  zc* out = (t ? a.trans : a.refl) + ((size_t)p * (a.L - 1) + l) * 6;
  out[0] = x[base] * w;
  out[1] = x[base + 1] * w;
  out[2] = x[base + 2] * w;
  out[3] = mk(sx * a.kv);
  out[4] = mk(sy * a.kv);
  out[5] = kz;
}

// (value, row) maximum over the wavefront, lowest row on ties.
__device__ inline void wave_argmax(double& v, int& i) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const double ov = __shfl_xor(v, o, 64);
    const int oi = __shfl_xor(i, o, 64);
    if (ov > v || (ov == v && oi < i)) {
      v = ov;
      i = oi;
    }
  }
}

// One wavefront per incident wave; S holds the augmented n x (n+1) system row by row.
__global__ __launch_bounds__(64) void k_layer_fused(LayerArgs a) {
  extern __shared__ zc S[];
  __shared__ zc X[LAYER_FUSED_MAXN];
  __shared__ int ord[LAYER_FUSED_MAXN];
  const int n = a.n, rs = n + 1, lane = threadIdx.x;
  for (long p = blockIdx.x; p < a.P; p += gridDim.x) {
    const zc* rec = a.inc + p * 6;
    double d[3];
    inc_direction(rec, d);
    for (int e = lane; e < n * rs; e += 64) S[e] = mk(0.0);
    __syncthreads();
    if (lane < n) {
      zc* row = S + lane * rs;
      layer_row(a, lane, d, rec, [&](int c, zc v) { row[c] = v; }, [&](zc v) { row[n] = v; });
    }
    __syncthreads();
    // elimination with partial pivoting; the pivot rows stay where they are, ord[k] names them
    bool used = lane >= n;
    int rank = n;
    int sing = 0;
    for (int k = 0; k < n; ++k) {
      double v = used ? -1.0 : abs1(S[lane * rs + k]);
      int pr = lane;
      wave_argmax(v, pr);
      if (lane == pr) {
        used = true;
        rank = k;
      }
      if (lane == 0) ord[k] = pr;
      if (v == 0.0) {
        sing = 1;
      } else if (!used) {
        const zc* prow = S + pr * rs;
        zc* row = S + lane * rs;
        const zc f = zdiv(row[k], prow[k]);
        for (int j = k + 1; j <= n; ++j) row[j] -= f * prow[j];
      }
      __syncthreads();
    }
    // back substitution in pivot order
    for (int k = n - 1; k >= 0; --k) {
      const int pr = ord[k];
      const zc xk = zdiv(S[pr * rs + n], S[pr * rs + k]);
      if (lane == 0) X[k] = xk;
      if (rank < k) S[lane * rs + n] -= S[lane * rs + k] * xk;
      __syncthreads();
    }
    if (sing && lane == 0) atomicAdd(a.nsing, 1);
    if (lane < 2 * (a.L - 1)) layer_record(a, p, lane, d, X);
    __syncthreads();
  }
}

// Broadcast of lane l's value to the whole wavefront (v_readlane: l is wave-uniform).
__device__ inline double lane_d(double v, int l) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)u, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(u >> 32), l);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ inline zc lane_z(zc v, int l) { return zc{lane_d(v.re, l), lane_d(v.im, l)}; }

// Rotation of a 16-lane row by s lanes (DPP row_ror: 0x120 + s).
template <int S>
__device__ inline double ror_d(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const unsigned lo = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)u, 0x120 + S, 0xF, 0xF, false);
  const unsigned hi = (unsigned)__builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), 0x120 + S, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
template <int S>
__device__ inline int ror_i(int v) {
  return __builtin_amdgcn_mov_dpp(v, 0x120 + S, 0xF, 0xF, false);
}
template <int S>
__device__ inline void row_best(double& v, int& i) {
  const double ov = ror_d<S>(v);
  const int oi = ror_i<S>(i);
  if (ov > v || (ov == v && oi < i)) {
    v = ov;
    i = oi;
  }
}

// Stacks of up to 8 layers: one wavefront per incident wave, lane r holding row r of the augmented
// system in registers.  The pivot (largest |re| + |im|, lowest row on ties) is found by DPP rotations
// inside each 16-lane row and lane reads across the rows; the pivot row reaches every lane through
// v_readlane, so an elimination step is register FMAs only.  The rows are assembled through LDS.
template <int L>
__global__ __launch_bounds__(64) void k_layer_reg(LayerArgs a) {
  constexpr int n = 6 * L, rs = n + 1, nrows16 = (n + 15) / 16;
  extern __shared__ zc S[];
  __shared__ zc X[n];
  const int lane = threadIdx.x;
  for (long p = blockIdx.x; p < a.P; p += gridDim.x) {
    const zc* rec = a.inc + p * 6;
    double d[3];
    inc_direction(rec, d);
    for (int e = lane; e < n * rs; e += 64) S[e] = mk(0.0);
    __syncthreads();
    if (lane < n) {
      zc* row = S + lane * rs;
      layer_row(a, lane, d, rec, [&](int c, zc v) { row[c] = v; }, [&](zc v) { row[n] = v; });
    }
    __syncthreads();
    zc row[n + 1];
    const int mine = lane < n ? lane : 0;
#pragma unroll
    for (int c = 0; c <= n; ++c) row[c] = S[mine * rs + c];
    bool used = lane >= n;
    int rank = n;
    int sing = 0;
    int ord[n];
#pragma unroll
    for (int k = 0; k < n; ++k) {
      double v = used ? -1.0 : abs1(row[k]);
      int idx = lane;
      row_best<1>(v, idx);
      row_best<2>(v, idx);
      row_best<4>(v, idx);
      row_best<8>(v, idx);
      double bv = lane_d(v, 0);
      int pr = __builtin_amdgcn_readlane(idx, 0);
#pragma unroll
      for (int r = 1; r < nrows16; ++r) {
        const double ov = lane_d(v, 16 * r);
        const int oi = __builtin_amdgcn_readlane(idx, 16 * r);
        if (ov > bv || (ov == bv && oi < pr)) {
          bv = ov;
          pr = oi;
        }
      }
      pr = __builtin_amdgcn_readfirstlane(pr);
      ord[k] = pr;
      if (lane == pr) {
        used = true;
        rank = k;
      }
      if (bv == 0.0) sing = 1;
      const bool act = !used && bv != 0.0;
      const zc pk = lane_z(row[k], pr);
      const zc f = act ? zdiv(row[k], pk) : mk(0.0);
#pragma unroll
      for (int j = k + 1; j <= n; ++j) {
        const zc pj = lane_z(row[j], pr);
        if (act) row[j] -= f * pj;
      }
    }
#pragma unroll
    for (int k = n - 1; k >= 0; --k) {
      const int pr = ord[k];
      const zc xk = zdiv(lane_z(row[n], pr), lane_z(row[k], pr));
      if (lane == 0) X[k] = xk;
      if (rank < k) row[n] -= row[k] * xk;
    }
    __syncthreads();
    if (sing && lane == 0) atomicAdd(a.nsing, 1);
    if (lane < 2 * (L - 1)) layer_record(a, p, lane, d, X);
    __syncthreads();
  }
}

template <int L>
void launch_layer_reg(const LayerArgs& a, hipStream_t s) {
  const size_t lds = sizeof(zc) * 6 * L * (6 * L + 1);
  const unsigned grid = (unsigned)std::min<long>(a.P, 1L << 20);
  hipLaunchKernelGGL(k_layer_reg<L>, dim3(grid), dim3(64), lds, s, a);
}

// Taller stacks: assemble A (column-major, n x n) and b of every incident wave in HBM.
__global__ void k_layer_assemble(LayerArgs a, long p0, long np, zc* A, zc* B) {
  const int n = a.n;
  for (long pi = blockIdx.x; pi < np; pi += gridDim.x) {
    const long p = p0 + pi;
    zc* Ap = A + pi * (long)n * n;
    zc* bp = B + pi * (long)n;
    for (long e = threadIdx.x; e < (long)n * n; e += blockDim.x) Ap[e] = mk(0.0);
    for (int e = threadIdx.x; e < n; e += blockDim.x) bp[e] = mk(0.0);
    __syncthreads();
    const zc* rec = a.inc + p * 6;
    double d[3];
    inc_direction(rec, d);
    for (int r = threadIdx.x; r < n; r += blockDim.x)
      layer_row(a, r, d, rec, [&](int c, zc v) { Ap[(long)c * n + r] = v; }, [&](zc v) { bp[r] = v; });
    __syncthreads();
  }
}

__global__ void k_layer_records(LayerArgs a, long p0, long np, const zc* X, const int* info) {
  const int nr = 2 * (a.L - 1);
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < np * nr; e += (long)gridDim.x * blockDim.x) {
    const long pi = e / nr;
    const int q = (int)(e % nr);
    double d[3];
    inc_direction(a.inc + (p0 + pi) * 6, d);
    layer_record(a, p0 + pi, q, d, X + pi * a.n);
    if (q == 0 && info[pi] != 0) atomicAdd(a.nsing, 1);
  }
}

}  // namespace

int layer_solve_dev(int L, const zc* d_ri, const double* d_z, double lambda, long P, const zc* d_inc, zc* d_refl,
                    zc* d_trans, hipStream_t s) {
  CWX_CHECK(L >= 2, "a stack needs at least two layers");
  CWX_CHECK(L <= 2000, "at most 2000 layers");
  CWX_CHECK(lambda > 0.0, "--lambda must be positive");
  CWX_CHECK(P >= 0, "negative number of incident waves");
  if (P == 0) return 0;
  LayerArgs a{};
  a.L = L;
  a.n = 6 * L;
  a.ri = d_ri;
  a.z = d_z;
  a.kv = 2.0 * M_PI / lambda;  // ScatterLayer.cpp:351 (M_PI, not the truncated pi of FourierWave.h)
  a.P = P;
  a.inc = d_inc;
  a.refl = d_refl;
  a.trans = d_trans;
  DevBuf<int> nsing(1);
  CWX_HIP(hipMemsetAsync(nsing.p, 0, sizeof(int), s));
  a.nsing = nsing.p;
  if (L <= 8) {
    switch (L) {
      case 2: launch_layer_reg<2>(a, s); break;
      case 3: launch_layer_reg<3>(a, s); break;
      case 4: launch_layer_reg<4>(a, s); break;
      case 5: launch_layer_reg<5>(a, s); break;
      case 6: launch_layer_reg<6>(a, s); break;
      case 7: launch_layer_reg<7>(a, s); break;
      default: launch_layer_reg<8>(a, s); break;
    }
    CWX_HIP(hipGetLastError());
  } else if (a.n <= LAYER_FUSED_MAXN) {
    const size_t lds = sizeof(zc) * a.n * (a.n + 1);
    const unsigned grid = (unsigned)std::min<long>(P, 1L << 20);
    hipLaunchKernelGGL(k_layer_fused, dim3(grid), dim3(64), lds, s, a);
    CWX_HIP(hipGetLastError());
  } else {
    const long n = a.n;
    const long per = n * n + n;  // complex elements per system
    const long chunk = std::max<long>(1, std::min<long>({P, (4L << 30) / (16 * per), 65535}));
    DevBuf<zc> A((size_t)chunk * n * n), B((size_t)chunk * n);
    DevBuf<int> piv((size_t)chunk * n), info(chunk);
    for (long p0 = 0; p0 < P; p0 += chunk) {
      const long np = std::min(chunk, P - p0);
      hipLaunchKernelGGL(k_layer_assemble, dim3((unsigned)std::min<long>(np, 65535)), dim3(256), 0, s, a, p0, np,
                         A.p, B.p);
      CWX_HIP(hipGetLastError());
      zgetrf_batched((int)n, A.p, (int)n, n * n, piv.p, n, info.p, (int)np, s);
      zgetrs_batched((int)n, 1, A.p, (int)n, n * n, piv.p, n, B.p, (int)n, n, (int)np, s);
      hipLaunchKernelGGL(k_layer_records, dim3((unsigned)std::min<long>((np * 2 * (L - 1) + 255) / 256, 65535)),
                         dim3(256), 0, s, a, p0, np, B.p, info.p);
      CWX_HIP(hipGetLastError());
    }
    CWX_HIP(hipStreamSynchronize(s));  // the chunk buffers are released on return
  }
  int h = 0;
  CWX_HIP(hipMemcpyAsync(&h, nsing.p, sizeof(int), hipMemcpyDeviceToHost, s));
  CWX_HIP(hipStreamSynchronize(s));
  return h;
}

}  // namespace cwx
