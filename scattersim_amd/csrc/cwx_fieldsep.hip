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

#include <algorithm>
#include <cmath>
#include <cstring>

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

// ------------------------------------------------------------------------------------------------
// Complex64 output path: coefficients of all planes of one layer in one pass over [Q_check | Q_hat],
// then T = X C per plane (FP64, rounded to complex64), then E = T Y^T on the FP32 matrix cores.
// ------------------------------------------------------------------------------------------------
// Entry t of the complex64 path's tables: X[ix][a] = exp(i kx_a x_ix) (FP64) for t < nx M0, then
// Yf[b][iy] = exp(i ky_b y_iy) as complex64, b padded to K4 with zeros.
__device__ inline void sep_table32_entry(const SepArgs& a, const GridSpec& g, int K4, zc* __restrict__ X,
                                         float2* __restrict__ Yf, int t) {
  const int nX = g.nx * a.M0;
  if (t < nX) {
    const int ix = t / a.M0, ai = t % a.M0;
    const double xv = (double)(coord(g.x0, g.dx, ix));
    X[t] = cexpi(a.Sx[ai * a.M1] * a.k, xv);
    return;
  }
  const int u = t - nX;
  if (u >= K4 * g.ny) return;
  const int iy = u % g.ny, b = u / g.ny;
  zc v = mk(0.0);
  if (b < a.M1) v = cexpi(a.Sy[b] * a.k, (double)coord(g.y0, g.dy, iy));
  Yf[u] = make_float2((float)v.re, (float)v.im);
}

struct PlaneGroup {
  int iz0, count, region;  // planes iz0 .. iz0+count-1, all in region `region` (0: above, L+1: below)
};
constexpr int SEP_G = 8;  // planes per group

// 64 coefficients per workgroup over one slice of the 4MF modes (blockIdx.z of gridDim.z slices, so that a few
// layers still give the chip enough workgroups to stream Q at full rate); the SEP_W wavefronts split the slice
// and their partial sums are added in LDS in a fixed order.  Slice s writes C + s cstride; the rows pass adds
// the slices in order.  Workgroups with blockIdx.x >= xb0 (of plane group 0, slice 0) fill the x and y
// phase tables instead, which saves a dependent launch.
constexpr int SEP_W = 8;
__global__ __launch_bounds__(64 * SEP_W) void k_sep_coeffs_g(SepArgs a, GridSpec g, const PlaneGroup* __restrict__ groups,
                                                      zc* __restrict__ C, long cstride, int xb0, int K4,
                                                      zc* __restrict__ X, float2* __restrict__ Yf) {
  extern __shared__ zc ph[];  // count x slice mode phases
  __shared__ zc part[SEP_W][SEP_G][64];
  if ((int)blockIdx.x >= xb0) {
    if (blockIdx.y == 0 && blockIdx.z == 0)
      sep_table32_entry(a, g, K4, X, Yf, ((int)blockIdx.x - xb0) * blockDim.x + threadIdx.x);
    return;
  }
  const PlaneGroup gr = groups[blockIdx.y];
  const int MF = a.MF, n3 = 3 * MF, n4 = 4 * MF;
  const int l = gr.region;
  const bool layer = (l >= 1 && l <= a.L);
  const int i = l - 1;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int e = blockIdx.x * 64 + lane;
  const int S = gridDim.z, sl = blockIdx.z;
  const int msl = (n4 + S - 1) / S, mlo = min(n4, sl * msl), mhi = min(n4, mlo + msl), ns = mhi - mlo;
  C += sl * cstride;
  if (layer && !a.external) {
    for (int t = threadIdx.x; t < gr.count * ns; t += blockDim.x) {
      const int pl = t / ns, m = mlo + t % ns;
      const double z = (double)coord(g.z0, g.dz, gr.iz0 + pl);
      const bool up = m >= 2 * MF;
      const int mode = up ? 2 * (m - 2 * MF) + 1 : 2 * m;
      const zc kz = a.gam[(long)i * n4 + mode] * a.k;
      ph[t] = cexpi(kz, z - (up ? a.zl[i + 1] : a.zl[i]));
    }
    __syncthreads();
    zc acc[SEP_G];
#pragma unroll
    for (int t = 0; t < SEP_G; ++t) acc[t] = mk(0.0);
    if (e < n3) {
      const int mq = (ns + SEP_W - 1) / SEP_W, m0 = wv * mq, m1 = min(ns, m0 + mq);
      const zc* Q = a.Qch + (long)i * n3 * n4 + (long)mlo * n3 + e;
#pragma unroll 4
      for (int m = m0; m < m1; ++m) {
        const zc q = Q[(long)m * n3];
#pragma unroll
        for (int t = 0; t < SEP_G; ++t)
          if (t < gr.count) fma_acc(acc[t], q, ph[t * ns + m]);
      }
    }
#pragma unroll
    for (int t = 0; t < SEP_G; ++t) part[wv][t][lane] = acc[t];
    __syncthreads();
    if (wv == 0 && e < n3) {
      for (int t = 0; t < gr.count; ++t) {
        zc sum = part[0][t][lane];
        for (int w = 1; w < SEP_W; ++w) sum = sum + part[w][t][lane];
        C[(long)(gr.iz0 + t) * n3 + e] = sum;
      }
    }
    return;
  }
  if (wv != 0 || e >= n3) return;
  const int c = e / MF, p = e % MF;
  for (int t = 0; t < gr.count; ++t) {
    const double z = (double)coord(g.z0, g.dz, gr.iz0 + t);
    zc v = mk(0.0);
    if (sl == 0 && l == 0) {
      const zc kzi = a.Sz0[p] * a.k;
      const zc kzr = -(a.Sz0[p] * a.k) * a.n0;
      v = a.EF[c * MF + p] * cexpi(kzi, z) + a.x[c * MF + p] * cexpi(kzr, z - a.zl[0]);
    } else if (sl == 0 && l == a.L + 1) {
      const zc kzt = (a.Sz1[p] * a.k) * a.n1;
      v = a.x[n3 + c * MF + p] * cexpi(kzt, z - a.zl[a.L]);
    }
    C[(long)(gr.iz0 + t) * n3 + e] = v;
  }
}

// T[iz][c][b][ix] = sum_a exp(i kx_a x_ix) C[iz][c][a M1 + b] as complex64, b padded to K4 with zeros:
// one thread per column ix of a plane and group of 4 entries (c, b..b+3), so that its row of X is read
// once for the four.  The plane's coefficients are the S slices of the coefficient pass, added in slice
// order.
__global__ __launch_bounds__(256) void k_sep_rows(SepArgs a, GridSpec g, int S, long cstride,
                                                  const zc* __restrict__ C, const zc* __restrict__ X, int K4,
                                                  float2* __restrict__ T) {
  extern __shared__ zc sCp[];  // the M0 x 4 coefficients (c, a, b0..b0+3) this workgroup uses
  const int iz = blockIdx.y, MF = a.MF, M0 = a.M0, M1 = a.M1;
  const int G = K4 / 4, c = blockIdx.z / G, b0 = 4 * (blockIdx.z % G);
  for (int e = threadIdx.x; e < 4 * M0; e += blockDim.x) {
    const int ai = e >> 2, b = b0 + (e & 3);
    zc v = mk(0.0);
    if (b < M1) {
      const long o = (long)iz * 3 * MF + c * MF + ai * M1 + b;
      v = C[o];
      for (int sl = 1; sl < S; ++sl) v = v + C[sl * cstride + o];
    }
    sCp[e] = v;
  }
  __syncthreads();
  const int ix = blockIdx.x * blockDim.x + threadIdx.x;
  if (ix >= g.nx) return;
  const long per = (long)3 * K4 * g.nx;
  float2* Tp = T + (long)iz * per + ix;
  zc acc[4] = {mk(0.0), mk(0.0), mk(0.0), mk(0.0)};
  for (int ai = 0; ai < M0; ++ai) {
    const zc xa = X[(long)ix * M0 + ai];
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (b0 + u < M1) fma_acc(acc[u], xa, sCp[4 * ai + u]);
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) Tp[(long)(c * K4 + b0 + u) * g.nx] = make_float2((float)acc[u].re, (float)acc[u].im);
}

typedef float f4 __attribute__((ext_vector_type(4)));
constexpr int SY_TILES = 8;  // y tiles of 16 rows per wavefront

// E(x, y) = sum_b Y(y, b) T(b, x) per component with v_mfma_f32_16x16x4f32 (A[i][k] from lane i + 16k,
// B[k][j] from lane j + 16k, D[4 (l / 16) + q][l % 16] in register q of lane l).  A wavefront owns 16
// columns and SY_TILES x 16 rows of one plane; the T fragments of its columns stay in registers.
template <int NK>
__global__ __launch_bounds__(256) void k_sep_synth_mfma(GridSpec g, int K4, const float2* __restrict__ T,
                                                        const float2* __restrict__ Yf, float2* __restrict__ out) {
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int iz = blockIdx.z;
  const int x0 = blockIdx.x * 16;
  const int y0 = (blockIdx.y * 4 + wv) * 16 * SY_TILES;
  if (y0 >= g.ny) return;
  const int lx = lane & 15, lk = lane >> 4;
  const long per = (long)K4 * g.nx;
  const float2* Tz = T + (long)iz * 3 * per;
  // NK = K4 / 4 blocks of 4 orders along y, up to 16 (a template parameter, so that only the fragments
  // in use take registers and more wavefronts keep stores in flight)
  constexpr int KMAX = NK;
  float2 tf[3][KMAX];
  const int nk = NK;
#pragma unroll
  for (int kb = 0; kb < KMAX; ++kb)
    if (kb < nk)
#pragma unroll
      for (int c = 0; c < 3; ++c)
        tf[c][kb] = (x0 + lx < g.nx) ? Tz[c * per + (long)(kb * 4 + lk) * g.nx + x0 + lx] : make_float2(0.f, 0.f);
  for (int yt = 0; yt < SY_TILES; ++yt) {
    const int yb = y0 + 16 * yt;
    if (yb >= g.ny) break;
    f4 P[3], Q[3], R[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) P[c] = Q[c] = R[c] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kb = 0; kb < KMAX; ++kb) {
      if (kb < nk) {
        const float2 y = (yb + lx < g.ny) ? Yf[(long)(kb * 4 + lk) * g.ny + yb + lx] : make_float2(0.f, 0.f);
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          P[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(y.x, tf[c][kb].x, P[c], 0, 0, 0);
          Q[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(y.y, tf[c][kb].y, Q[c], 0, 0, 0);
          R[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(y.x, tf[c][kb].y, R[c], 0, 0, 0);
          R[c] = __builtin_amdgcn_mfma_f32_16x16x4f32(y.y, tf[c][kb].x, R[c], 0, 0, 0);
        }
      }
    }
    const int ix = x0 + lx;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int iy = yb + 4 * lk + q;
      if (ix < g.nx && iy < g.ny) {
        float2* o = out + (((long)iz * g.ny + iy) * g.nx + ix) * 3;
#pragma unroll
        for (int c = 0; c < 3; ++c) o[c] = make_float2(P[c][q] - Q[c][q], R[c][q]);
      }
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
  // the boundary tables (and the plane groups below) go to the device in one copy: each small copy
  // from host memory costs the stream several microseconds
  const size_t zd_off = ((sizeof(float) * (L + 1) + 15) / 16) * 16;
  const size_t gr_off = zd_off + ((sizeof(double) * (L + 1) + 15) / 16) * 16;
  std::vector<PlaneGroup> groups;
  const int n4 = 4 * MF;
  const int gmax = std::max(1, std::min(SEP_G, (int)((32 * 1024) / (sizeof(zc) * n4))));
  if (!out_f64) {
    // plane groups: consecutive planes of one region, at most SEP_G and at most what the phase
    // table in LDS holds; regions computed with the device's coordinate and test
    auto region = [&](int iz) {
      const float z = std::fmaf(g.dz, (float)iz, g.z0);
      int l = 0;
      for (int li = 0; li <= L; ++li) {
        if (tol_mode == 0) {
          if ((double)z >= (double)zf[li] - 0.001) l = li + 1;
        } else if (z >= zf[li]) {
          l = li + 1;
        }
      }
      return l;
    };
    for (int iz = 0; iz < g.nz;) {
      const int r = region(iz);
      int cnt = 1;
      while (iz + cnt < g.nz && cnt < gmax && region(iz + cnt) == r) ++cnt;
      groups.push_back(PlaneGroup{iz, cnt, r});
      iz += cnt;
    }
    CWX_CHECK(groups.size() <= 65535, "too many plane groups");
  }
  std::vector<unsigned char> pack(gr_off + sizeof(PlaneGroup) * groups.size());
  std::memcpy(pack.data(), zf.data(), sizeof(float) * (L + 1));
  std::memcpy(pack.data() + zd_off, zd.data(), sizeof(double) * (L + 1));
  if (!groups.empty()) std::memcpy(pack.data() + gr_off, groups.data(), sizeof(PlaneGroup) * groups.size());
  // repeated evaluations of one solution on one grid reuse the tables already on the device
  if (pack != w.pack_host || !w.pack.p) {
    w.pack.ensure(pack.size());
    CWX_HIP(hipMemcpyAsync(w.pack.p, pack.data(), pack.size(), hipMemcpyHostToDevice, s));
    w.pack_host = pack;
  }
  const float* d_zf = reinterpret_cast<const float*>(w.pack.p);
  const double* d_zd = reinterpret_cast<const double*>(w.pack.p + zd_off);
  const PlaneGroup* d_groups = reinterpret_cast<const PlaneGroup*>(w.pack.p + gr_off);
  w.C.ensure((size_t)g.nz * 3 * MF);
  w.X.ensure((size_t)g.nx * st.M0);
  w.Y.ensure((size_t)g.ny * st.M1);
  SepArgs a{st.M0, st.M1, MF, L, sol.opt.external ? 1 : 0, st.k, sol.opt.n[0], sol.opt.n[1],
            sol.tabs.p, sol.tabs.p + MF, sol.tabs.p + 2 * MF, sol.tabs.p + 3 * MF, sol.tabs.p + 4 * MF,
            sol.x.p, sol.gam.p, sol.Qch.p, d_zf, d_zd, tol_mode};
  if (!out_f64) {
    const int K4 = (st.M1 + 3) / 4 * 4;
    CWX_CHECK(K4 <= 64, "more than 64 orders along y in the complex64 evaluator");
    w.Tf.ensure((size_t)g.nz * 3 * K4 * g.nx);
    w.Yf.ensure((size_t)K4 * g.ny);
    // mode slices: enough workgroups to keep HBM busy when there are few plane groups
    const long base = (long)cdiv(3 * MF, 64) * (long)groups.size();
    const int S = (int)std::max<long>(1, std::min<long>({16L, (4096 + base - 1) / base, (long)n4 / 32}));
    const long cstride = (long)g.nz * 3 * MF;
    w.C.ensure((size_t)S * cstride);
    const int msl = (n4 + S - 1) / S;
    const int xb0 = cdiv(3 * MF, 64), xbt = cdiv(g.nx * st.M0 + K4 * g.ny, 64 * SEP_W);
    hipLaunchKernelGGL(k_sep_coeffs_g, dim3(xb0 + xbt, (unsigned)groups.size(), S), dim3(64 * SEP_W),
                       sizeof(zc) * gmax * msl, s, a, g, d_groups, w.C.p, cstride, xb0, K4,
                       w.X.p, w.Yf.p);
    CWX_CHECK(3 * (K4 / 4) <= 65535, "bad order count");
    hipLaunchKernelGGL(k_sep_rows, dim3(cdiv(g.nx, 256), g.nz, 3 * (K4 / 4)), dim3(256), sizeof(zc) * 4 * st.M0, s,
                       a, g, S, cstride, w.C.p, w.X.p, K4, w.Tf.p);
    dim3 grid(cdiv(g.nx, 16), cdiv(g.ny, 4 * 16 * SY_TILES), g.nz);
    switch (K4 / 4) {
#define CWX_SYNTH(NK) \
  case NK: hipLaunchKernelGGL(k_sep_synth_mfma<NK>, grid, dim3(256), 0, s, g, K4, w.Tf.p, w.Yf.p, (float2*)out); break;
      CWX_SYNTH(1) CWX_SYNTH(2) CWX_SYNTH(3) CWX_SYNTH(4) CWX_SYNTH(5) CWX_SYNTH(6) CWX_SYNTH(7) CWX_SYNTH(8)
      CWX_SYNTH(9) CWX_SYNTH(10) CWX_SYNTH(11) CWX_SYNTH(12) CWX_SYNTH(13) CWX_SYNTH(14) CWX_SYNTH(15)
      CWX_SYNTH(16)
#undef CWX_SYNTH
      default: CWX_CHECK(false, "more than 64 orders along y in the complex64 evaluator");
    }
    CWX_HIP(hipGetLastError());
    return;
  }
  const size_t sh_coef = sizeof(zc) * 4 * (size_t)MF;
  CWX_CHECK(sh_coef <= lds_limit(), "too many orders for the separable evaluator");
  hipLaunchKernelGGL(k_sep_coeffs, dim3(cdiv(3 * MF, 256), g.nz), dim3(256), sh_coef, s, a, g, w.C.p);
  const int ntab = g.nx * st.M0 + g.ny * st.M1;
  hipLaunchKernelGGL(k_sep_tables, dim3(cdiv(ntab, 256)), dim3(256), 0, s, a, g, w.X.p, w.Y.p);
  // tile width: 64 columns when the LDS footprint allows, else 32 or 16
  int TX = 64;
  auto lds = [&](int tx) { return sizeof(zc) * (3 * (size_t)MF + tx * st.M0 + TY * st.M1 + 3 * tx * st.M1); };
  while (TX > 16 && lds(TX) > 64 * 1024) TX /= 2;
  const size_t sh = lds(TX);
  CWX_CHECK(sh <= lds_limit(), "too many orders for the separable evaluator tile");
  dim3 grid(cdiv(g.nx, TX), cdiv(g.ny, TY), g.nz);
  if (out_f64)
    hipLaunchKernelGGL((k_sep_synth<double2>), grid, dim3(256), sh, s, a, g, w.C.p, w.X.p, w.Y.p, (double2*)out, TX);
  else
    hipLaunchKernelGGL((k_sep_synth<float2>), grid, dim3(256), sh, s, a, g, w.C.p, w.X.p, w.Y.p, (float2*)out, TX);
  CWX_HIP(hipGetLastError());
}

}  // namespace cwx
