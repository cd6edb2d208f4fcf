// Separable evaluation of a solved structure's field on a regular grid.
//
// Every plane wave of a solution's .cw records belongs to one Fourier order p and carries that order's
// transverse wavevector (Sx(p) k, Sy(p) k) (ScatterVolume.cpp:910-966); kx depends only on the x-order
// index a = p / M1 and ky only on the y-order index b = p % M1 (the column-major flattening of
// ScatterVolume.cpp:682-683).  In the region holding plane z the field is therefore
//     E(x, y, z) = sum_a sum_b exp(i kx_a x) exp(i ky_b y) C_ab(z),
//     C_p(z) = sum_m A_pm exp(i kz_m (z - z_ref,m)),
// with A the incident/reflected/transmitted amplitudes or [Q_check | Q_hat] of the layer and kz_m the
// half-space kz or gamma_m k.  This is the same sum as gpuEvaluator.cu:53-66 over the same waves,
// regrouped: per plane a (3MF x 4MF) x (4MF) product for C, then per point tile the two small
// products X C and (X C) Y^T in LDS.  All arithmetic is FP64; the output is complex64 or complex128.
#include "cwx_common.h"
#include "cwx_solver.h"

namespace cwx {

namespace {

// grid coordinate origin + step * i with a single rounding (a fused multiply-add, the form a compiler
// with contraction gives x_start + d * xi of gpuEvaluator.cu:78-79), reproducible bit for bit on the host
__device__ inline float coord(float o, float d, int i) { return fmaf(d, (float)i, o); }

struct SepArgs {
  int M0, M1, MF, L, external;
  double k, n0, n1;
  const zc* Sx;   // MF
  const zc* Sy;
  const zc* Sz0;
  const zc* Sz1;
  const zc* EF;   // 3MF
  const zc* x;    // 6MF
  const zc* gam;  // L x 4MF
  const zc* Qch;  // L x (3MF x 4MF)
  const float* zlf;   // L + 1 boundary z as float (the region test of gpuEvaluator.cu:34)
  const double* zl;   // L + 1 boundary z as double (phase references)
  int tol_mode;
};

__device__ inline int plane_region(float z, const float* zl, int nb, int tol_mode) {
  int l = 0;
  for (int li = 0; li < nb; ++li) {
    if (tol_mode == 0) {
      if ((double)z >= (double)zl[li] - 0.001) l = li + 1;
    } else if (z >= zl[li]) {
      l = li + 1;
    }
  }
  return l;
}

__device__ inline zc cexpi(zc kz, double dz) {
  // exp(i kz dz)
  return zexp(mk(-kz.im * dz, kz.re * dz));
}

// C[iz][c][p] (3 x MF per plane) of the region holding plane iz.  One thread per (c, p).
__global__ __launch_bounds__(256) void k_sep_coeffs(SepArgs a, GridSpec g, zc* __restrict__ C) {
  extern __shared__ zc ph[];  // 4MF mode phases of the plane
  const int iz = blockIdx.y;
  const float zf = coord(g.z0, g.dz, iz);
  const double z = (double)zf;
  const int MF = a.MF, n3 = 3 * MF, n4 = 4 * MF;
  const int l = plane_region(zf, a.zlf, a.L + 1, a.tol_mode);
  const bool layer = (l >= 1 && l <= a.L);
  const int i = l - 1;
  if (layer && !a.external) {
    for (int m = threadIdx.x; m < n4; m += blockDim.x) {
      const bool up = m >= 2 * MF;
      const int mode = up ? 2 * (m - 2 * MF) + 1 : 2 * m;
      const zc kz = a.gam[(long)i * n4 + mode] * a.k;
      ph[m] = cexpi(kz, z - (up ? a.zl[i + 1] : a.zl[i]));
    }
  }
  __syncthreads();
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n3) return;
  const int c = e / MF, p = e % MF;
  zc v = mk(0.0);
  if (l == 0) {
    const zc kzi = a.Sz0[p] * a.k;
    const zc kzr = -(a.Sz0[p] * a.k) * a.n0;
    v = a.EF[c * MF + p] * cexpi(kzi, z) + a.x[c * MF + p] * cexpi(kzr, z - a.zl[0]);
  } else if (l == a.L + 1) {
    const zc kzt = (a.Sz1[p] * a.k) * a.n1;
    v = a.x[n3 + c * MF + p] * cexpi(kzt, z - a.zl[a.L]);
  } else if (!a.external) {
    const zc* Q = a.Qch + (long)i * n3 * n4 + e;
    for (int m = 0; m < n4; ++m) fma_acc(v, Q[(long)m * n3], ph[m]);
  }
  C[(long)iz * n3 + e] = v;
}

// exp(i kx_a x) for every grid column and exp(i ky_b y) for every grid row.
__global__ void k_sep_tables(SepArgs a, GridSpec g, zc* __restrict__ X, zc* __restrict__ Y) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int nX = g.nx * a.M0, nY = g.ny * a.M1;
  if (t < nX) {
    const int ix = t / a.M0, ai = t % a.M0;
    const double xv = (double)(coord(g.x0, g.dx, ix));
    X[t] = cexpi(a.Sx[ai * a.M1] * a.k, xv);
  } else if (t < nX + nY) {
    const int u = t - nX;
    const int iy = u / a.M1, bi = u % a.M1;
    const double yv = (double)(coord(g.y0, g.dy, iy));
    Y[u] = cexpi(a.Sy[bi] * a.k, yv);
  }
}

constexpr int TY = 16;

// Synthesis: D = X_tile (TX x M0) C (M0 x M1) per component, then E = D Y_tile^T.
template <typename OutT>
__global__ __launch_bounds__(256) void k_sep_synth(SepArgs a, GridSpec g, const zc* __restrict__ C,
                                                   const zc* __restrict__ X, const zc* __restrict__ Y,
                                                   OutT* __restrict__ out, int TX) {
  extern __shared__ zc sm[];
  const int M0 = a.M0, M1 = a.M1, MF = a.MF;
  zc* sC = sm;                    // 3 MF      [c][p], p = ai * M1 + bi
  zc* sX = sC + 3 * MF;           // TX x M0
  zc* sY = sX + TX * M0;          // TY x M1
  zc* sD = sY + TY * M1;          // 3 x TX x M1
  const int ix0 = blockIdx.x * TX, iy0 = blockIdx.y * TY, iz = blockIdx.z;
  const int tid = threadIdx.x;
  for (int e = tid; e < 3 * MF; e += blockDim.x) sC[e] = C[(long)iz * 3 * MF + e];
  for (int e = tid; e < TX * M0; e += blockDim.x) {
    const int ix = ix0 + e / M0;
    sX[e] = ix < g.nx ? X[(long)ix * M0 + e % M0] : mk(0.0);
  }
  for (int e = tid; e < TY * M1; e += blockDim.x) {
    const int iy = iy0 + e / M1;
    sY[e] = iy < g.ny ? Y[(long)iy * M1 + e % M1] : mk(0.0);
  }
  __syncthreads();
  for (int e = tid; e < 3 * TX * M1; e += blockDim.x) {
    const int c = e / (TX * M1), r = e % (TX * M1);
    const int tx = r / M1, bi = r % M1;
    zc acc = mk(0.0);
    for (int ai = 0; ai < M0; ++ai) fma_acc(acc, sX[tx * M0 + ai], sC[c * MF + ai * M1 + bi]);
    sD[e] = acc;
  }
  __syncthreads();
  const int tx = tid % TX;
  const int ix = ix0 + tx;
  for (int ty = tid / TX; ty < TY; ty += blockDim.x / TX) {
    const int iy = iy0 + ty;
    if (ix >= g.nx || iy >= g.ny) continue;
    zc E[3] = {mk(0.0), mk(0.0), mk(0.0)};
    for (int bi = 0; bi < M1; ++bi) {
      const zc yv = sY[ty * M1 + bi];
#pragma unroll
      for (int c = 0; c < 3; ++c) fma_acc(E[c], sD[(c * TX + tx) * M1 + bi], yv);
    }
    OutT* o = out + (((long)iz * g.ny + iy) * g.nx + ix) * 3;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if constexpr (sizeof(OutT) == 8) o[c] = OutT{(float)E[c].re, (float)E[c].im};
      else o[c] = OutT{E[c].re, E[c].im};
    }
  }
}

}  // namespace

void field_separable(const Solution& sol, const GridSpec& g, int tol_mode, int out_f64, void* out,
                     SepWork& w, hipStream_t s) {
  const Setup& st = sol.st;
  const int MF = st.MF, L = st.L;
  CWX_CHECK(g.nx > 0 && g.ny > 0 && g.nz > 0 && g.nz <= 65535, "bad grid dimensions");
  CWX_CHECK(sol.rec.p != nullptr && sol.tabs.p != nullptr, "solution has no device data");
  std::vector<float> zf(L + 1);
  std::vector<double> zd(L + 1);
  for (int l = 0; l <= L; ++l) {
    zd[l] = st.z[l];
    zf[l] = (float)((sol.opt.external && l > 0 && l < L) ? 0.0 : st.z[l]);
  }
  w.zf.ensure(L + 1);
  w.zd.ensure(L + 1);
  CWX_HIP(hipMemcpyAsync(w.zf.p, zf.data(), sizeof(float) * (L + 1), hipMemcpyHostToDevice, s));
  CWX_HIP(hipMemcpyAsync(w.zd.p, zd.data(), sizeof(double) * (L + 1), hipMemcpyHostToDevice, s));
  w.C.ensure((size_t)g.nz * 3 * MF);
  w.X.ensure((size_t)g.nx * st.M0);
  w.Y.ensure((size_t)g.ny * st.M1);
  SepArgs a{st.M0, st.M1, MF, L, sol.opt.external ? 1 : 0, st.k, sol.opt.n[0], sol.opt.n[1],
            sol.tabs.p, sol.tabs.p + MF, sol.tabs.p + 2 * MF, sol.tabs.p + 3 * MF, sol.tabs.p + 4 * MF,
            sol.x.p, sol.gam.p, sol.Qch.p, w.zf.p, w.zd.p, tol_mode};
  const size_t sh_coef = sizeof(zc) * 4 * (size_t)MF;
  CWX_CHECK(sh_coef <= 64 * 1024, "too many orders for the separable evaluator");
  hipLaunchKernelGGL(k_sep_coeffs, dim3(cdiv(3 * MF, 256), g.nz), dim3(256), sh_coef, s, a, g, w.C.p);
  const int ntab = g.nx * st.M0 + g.ny * st.M1;
  hipLaunchKernelGGL(k_sep_tables, dim3(cdiv(ntab, 256)), dim3(256), 0, s, a, g, w.X.p, w.Y.p);
  // tile width: 64 columns when the LDS footprint allows, else 32 or 16
  int TX = 64;
  auto lds = [&](int tx) { return sizeof(zc) * (3 * (size_t)MF + tx * st.M0 + TY * st.M1 + 3 * tx * st.M1); };
  while (TX > 16 && lds(TX) > 64 * 1024) TX /= 2;
  const size_t sh = lds(TX);
  CWX_CHECK(sh <= 64 * 1024, "too many orders for the separable evaluator tile");
  dim3 grid(cdiv(g.nx, TX), cdiv(g.ny, TY), g.nz);
  if (out_f64)
    hipLaunchKernelGGL((k_sep_synth<double2>), grid, dim3(256), sh, s, a, g, w.C.p, w.X.p, w.Y.p, (double2*)out, TX);
  else
    hipLaunchKernelGGL((k_sep_synth<float2>), grid, dim3(256), sh, s, a, g, w.C.p, w.X.p, w.Y.p, (float2*)out, TX);
  CWX_HIP(hipGetLastError());
}

}  // namespace cwx
