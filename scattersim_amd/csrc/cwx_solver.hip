// The scattervolume coupled-wave solve on one GPU.
//
// Host part restates ScatterVolume.cpp:574-685 (options, orthogonalize, orders, z, Sx/Sy/Sz, EF);
// the device pipeline replaces FourierWave.h:303-397 (Phi), ScatterVolume.cpp:235-466 (eigen,
// sort, Gd/Gc chain, boundary system), :739-889 (zgesv, beta, Q_check/Q_hat, Pz) and :907-987
// (CoupledWaveStructure assembly).  All per-layer work is batched over layers.
#include "cwx_solver.h"

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <fstream>

namespace cwx {

constexpr double kPiTrunc = 3.141592653;  // FourierWave.h:18
constexpr double kPi = 3.14159265358979323846;
constexpr double kDblEps = 2.220446049250313e-16;

// ------------------------------------------------------------------------------------------------
// host setup
// ------------------------------------------------------------------------------------------------
namespace {

void shift_indices(int m, int nlen, bool rows_parity_odd, std::vector<int>& idx, bool is_rows) {
  // FourierWave.h:82-145: p/q come from the ROW parity for both axes (quirk kept)
  const int M = is_rows ? m : nlen;
  int p;
  if (rows_parity_odd) p = is_rows ? m / 2 : nlen / 2;
  else p = is_rows ? (m + 1) / 2 : (nlen + 1) / 2;
  idx.assign(M, 0);
  if (M % 2 == 1) {
    for (int i = 0; i < M - p - 1; ++i) idx[i] = (M - p) + i;
    for (int i = M - p - 1; i < M; ++i) idx[i] = i - (M - p - 1);
  } else {
    for (int i = 0; i < M - p; ++i) idx[i] = p + i;
    for (int i = M - p; i < M; ++i) idx[i] = i - (M - p);
  }
}

zc csq(zc a) { return a * a; }

}  // namespace

Setup make_setup(const VolumeOptions& o, int L, int s1, int s2) {
  CWX_CHECK(L >= 1 && s1 >= 1 && s2 >= 1, "sample must have at least one layer and one pixel");
  CWX_CHECK(o.lambda > 0.0, "--lambda must be positive");
  Setup st;
  st.L = L;
  st.s1 = s1;
  st.s2 = s2;
  // direction (ScatterVolume.cpp:575-583)
  double d[3] = {o.direction[0], o.direction[1], o.direction[2]};
  const double dn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  CWX_CHECK(dn > 0.0, "--direction must be non-zero");
  for (double& v : d) v /= dn;
  // orthogonalize (FourierWave.h:32-39): no conjugation in the normalisation
  zc E[3] = {mk(o.ex[0], o.ex[1]), mk(o.ey[0], o.ey[1]), mk(o.ez[0], o.ez[1])};
  const zc mag = zsqrt(csq(E[0]) + csq(E[1]) + csq(E[2]));
  CWX_CHECK(!(mag.re == 0.0 && mag.im == 0.0),
            "incident field has Ex^2+Ey^2+Ez^2 = 0 (the reference divides by zero here)");
  for (zc& v : E) v = zdiv(v, mag);
  auto cross = [](const zc* a, const zc* b, zc* out) {
    out[0] = a[1] * b[2] - a[2] * b[1];
    out[1] = a[2] * b[0] - a[0] * b[2];
    out[2] = a[0] * b[1] - a[1] * b[0];
  };
  zc dc[3] = {mk(d[0]), mk(d[1]), mk(d[2])};
  zc sv[3];
  cross(E, dc, sv);
  cross(dc, sv, st.e);
  for (int i = 0; i < 3; ++i) st.dir_n[i] = d[i] * o.n[0];
  st.k = 2.0 * kPiTrunc / o.lambda;
  st.ni[0] = mk(o.n[0], 0.0);
  st.ni[1] = mk(o.n[1], o.kappa);
  // InitLayer_z (ScatterVolume.cpp:162-168)
  st.z.resize(L + 1);
  st.z[0] = o.z - o.size[2] / 2.0;
  for (int i = 1; i <= L; ++i) st.z[i] = st.z[i - 1] + o.size[2] / L;
  // Fourier orders (ScatterVolume.cpp:599-651)
  int M0, M1;
  if (o.ncoef == 0) { M0 = s2; M1 = s1; }
  else if (o.ncoef == 1) { M0 = M1 = (int)std::sqrt((double)o.coef[0]); }
  else { M0 = o.coef[0]; M1 = o.coef[1]; }
  CWX_CHECK(M0 >= 1 && M1 >= 1, "--coef must be positive");
  const double* in_n = o.n;
  const double* sz = o.size;
  const double lam = o.lambda;
  const double* dir = st.dir_n;
  if (!o.no_cutoff) {
    const double a0 = (double)(M0 / 2) / sz[0] * lam + dir[0];
    const double a1 = (double)(M1 / 2) / sz[1] * lam + dir[1];
    if (a0 * a0 + a1 * a1 >= in_n[0] * in_n[0]) {
      st.cutoff = true;
      if (M0 != 1 && M1 != 1) {
        if ((double)(M0 / 2) / sz[0] * lam + dir[0] > 0.71 * in_n[0])
          M0 = (int)(0.707 * in_n[0] * 2 * sz[0] / lam) - 1;
        if ((double)(M1 / 2) / sz[1] * lam + dir[1] > 0.71 * in_n[0])
          M1 = (int)(0.707 * in_n[1] * 2 * sz[1] / lam) - 1;
      }
      if (M0 == 1) {
        if ((double)(M1 / 2) / sz[1] * lam + std::fabs(dir[1]) >= in_n[0]) {
          M1 = (int)(in_n[1] * 2 * sz[1] / lam - 1e-10);
          if (M1 % 2 == 0) M1 -= 1;
        }
      }
      if (M1 == 1) {
        if ((double)(M0 / 2) / sz[1] * lam + std::fabs(dir[0]) >= in_n[0]) {
          M0 = (int)(in_n[0] * 2 * sz[0] / lam - 1e-10);
          if (M0 % 2 == 0) M0 -= 1;
        }
      }
    }
  }
  if (M0 > s2) M0 = s2;
  if (M1 > s1) M1 = s1;
  CWX_CHECK(M0 >= 1 && M1 >= 1, "Fourier-order selection produced no orders");
  st.M0 = M0;
  st.M1 = M1;
  st.MF = M0 * M1;
  // UpWq_Cal (FourierWave.h:319-334)
  st.up.resize(M0);
  st.wq.resize(M1);
  for (int i = 0; i < M0; ++i) st.up[i] = 2 * kPiTrunc * (double)(i - M0 / 2) / sz[0] + dir[0] * st.k;
  for (int i = 0; i < M1; ++i) st.wq[i] = 2 * kPiTrunc * (double)(i - M1 / 2) / sz[1] + dir[1] * st.k;
  const int MF = st.MF;
  st.Sx.resize(MF);
  st.Sy.resize(MF);
  st.Sz0.resize(MF);
  st.Sz1.resize(MF);
  const zc n0sq = csq(st.ni[0]), n1sq = csq(st.ni[1]);
  for (int c = 0; c < M0; ++c)
    for (int r = 0; r < M1; ++r) {
      const int m = r + c * M1;  // column-major flattening of the (M1 x M0) mesh
      const zc sx = mk(st.up[c] / st.k), sy = mk(st.wq[r] / st.k);
      st.Sx[m] = sx;
      st.Sy[m] = sy;
      st.Sz0[m] = zsqrt(n0sq - csq(sx) - csq(sy));
      st.Sz1[m] = zsqrt(n1sq - csq(sx) - csq(sy));
    }
  // frequency indices of the cropped, shifted DFT (FourierWave.h:152-195)
  std::vector<int> indx, indy;
  const bool rodd = (s1 % 2) == 1;
  shift_indices(s1, s2, rodd, indx, true);
  shift_indices(s1, s2, rodd, indy, false);
  const int r0 = s1 / 2 - M1 / 2, c0 = s2 / 2 - M0 / 2;
  st.rows.resize(M1);
  st.cols.resize(M0);
  for (int r = 0; r < M1; ++r) st.rows[r] = indx[r0 + r];
  for (int c = 0; c < M0; ++c) st.cols[c] = indy[c0 + c];
  // EF: DFT of a constant field keeps only the zero frequency (ScatterVolume.cpp:665-679)
  st.EF.assign(3 * MF, mk(0.0));
  for (int comp = 0; comp < 3; ++comp)
    for (int c = 0; c < M0; ++c)
      for (int r = 0; r < M1; ++r) {
        const int m = r + c * M1;
        if (o.psf) st.EF[comp * MF + m] = st.e[comp];
        else if (st.rows[r] == 0 && st.cols[c] == 0) st.EF[comp * MF + m] = st.e[comp];
      }
  return st;
}

// ------------------------------------------------------------------------------------------------
// device kernels of the solve
// ------------------------------------------------------------------------------------------------
namespace {

__global__ void k_square_layers(long n, const zc* __restrict__ smp, zc* __restrict__ out) {
  // out[0:n] = s^2, out[n:2n] = (1/s)^2  (FourierWave.h:341,343)
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const zc v = smp[i];
    out[i] = v * v;
    const zc inv = zdiv(mk(1.0), v);
    out[n + i] = inv * inv;
  }
}

struct PhiArgs {
  int M0, M1, MF;
  double kinv;
  const double* up;  // M0
  const double* wq;  // M1
};

// Element (r, li) coefficient indices of the circulant blocks (FourierWave.h:366-372).
__device__ inline int circ_index(int r, int li, int M0, int M1) {
  const int qi = r / M0, pi = r % M0, qj = li / M0, pj = li % M0;
  const int indR = (qi - qj + M1) % M1, indC = (pi - pj + M0) % M0;
  const int rr = (indR + M1 / 2) % M1, cc = (indC + M0 / 2) % M0;
  return rr + cc * M1;  // column-major (M1 x M0)
}

// P (rows 0..2MF, cols 2MF..4MF of Phi) and Q (rows 2MF..4MF, cols 0..2MF), FourierWave.h:378-392.
__device__ inline zc phi_P(const PhiArgs& a, const zc* NI, int rr, int cc) {
  const int MF = a.MF;
  const int r = rr % MF, li = cc % MF;
  const double upr = a.up[r % a.M0], wqr = a.wq[r / a.M0];
  const double upl = a.up[li % a.M0], wql = a.wq[li / a.M0];
  const zc ni = NI[circ_index(r, li, a.M0, a.M1)];
  const bool diag = (r == li);
  if (rr < MF) {
    if (cc < MF) return (upr * a.kinv * wql) * ni;
    zc v = (-upr * a.kinv * upl) * ni;
    if (diag) v.re += 1.0;
    return v;
  }
  if (cc < MF) {
    zc v = (wqr * a.kinv * wql) * ni;
    if (diag) v.re -= 1.0;
    return v;
  }
  return (-wqr * a.kinv * upl) * ni;
}

__device__ inline zc phi_Q(const PhiArgs& a, const zc* NF, int rr, int cc) {
  const int MF = a.MF;
  const int r = rr % MF, li = cc % MF;
  const double upr = a.up[r % a.M0], wqr = a.wq[r / a.M0];
  const bool diag = (r == li);
  if (rr < MF) {
    if (cc < MF) return diag ? mk(-upr * wqr * a.kinv) : mk(0.0);
    zc v = -NF[circ_index(r, li, a.M0, a.M1)];
    if (diag) v.re += upr * upr * a.kinv;
    return v;
  }
  if (cc < MF) {
    zc v = NF[circ_index(r, li, a.M0, a.M1)];
    if (diag) v.re -= wqr * wqr * a.kinv;
    return v;
  }
  return diag ? mk(upr * wqr * a.kinv) : mk(0.0);
}

// full Phi (4MF x 4MF, column-major) per layer
__global__ void k_phi_full(PhiArgs a, const zc* __restrict__ NF, const zc* __restrict__ NI, zc* Phi) {
  const int n = 4 * a.MF, n2 = 2 * a.MF;
  const int l = blockIdx.y;
  NF += (long)l * a.MF;
  NI += (long)l * a.MF;
  Phi += (long)l * n * n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)n * n; e += (long)gridDim.x * blockDim.x) {
    const int r = (int)(e % n), c = (int)(e / n);
    zc v = mk(0.0);
    if (r < n2 && c >= n2) v = phi_P(a, NI, r, c - n2);
    else if (r >= n2 && c < n2) v = phi_Q(a, NF, r - n2, c);
    Phi[e] = v;
  }
}

// P and Q blocks (2MF x 2MF each) per layer
__global__ void k_phi_blocks(PhiArgs a, const zc* __restrict__ NF, const zc* __restrict__ NI, zc* P,
                             zc* Q) {
  const int n2 = 2 * a.MF;
  const int l = blockIdx.y;
  NF += (long)l * a.MF;
  NI += (long)l * a.MF;
  P += (long)l * n2 * n2;
  Q += (long)l * n2 * n2;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)n2 * n2; e += (long)gridDim.x * blockDim.x) {
    const int r = (int)(e % n2), c = (int)(e / n2);
    P[e] = phi_P(a, NI, r, c);
    Q[e] = phi_Q(a, NF, r, c);
  }
}

// sorter (ScatterVolume.cpp:177-184): true if a < b
__device__ inline bool sorter_less(zc a, zc b) {
  if (fabs(a.im - b.im) > kDblEps) return a.im < b.im;
  if (fabs(a.re - b.re) > kDblEps) return a.re < b.re;
  return a.im < b.im;
}

// Bitonic sort of (value, index) pairs in LDS with the reference comparator; ascending.
// Padding slots hold index -1 and compare greater than everything.
__device__ void bitonic_sort(zc* val, int* idx, int npow2) {
  for (int size = 2; size <= npow2; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      __syncthreads();
      for (int i = threadIdx.x; i < npow2 / 2; i += blockDim.x) {
        const int lo = 2 * i - (i & (stride - 1));
        const int hi = lo + stride;
        const bool up = ((lo & size) == 0);
        const zc a = val[lo], b = val[hi];
        const int ia = idx[lo], ib = idx[hi];
        bool a_gt_b;
        if (ia < 0) a_gt_b = ib >= 0 || ia > ib;
        else if (ib < 0) a_gt_b = false;
        else a_gt_b = sorter_less(b, a);
        if (a_gt_b == up) {
          val[lo] = b; val[hi] = a;
          idx[lo] = ib; idx[hi] = ia;
        }
      }
    }
  }
  __syncthreads();
}

constexpr int SORT_MAX = 2048;

// Full mode: eigenvalues W (n4) of Phi -> source column of each sorted, interleaved position
// (ScatterVolume.cpp:190-230).  Reduced mode: mu (n2) of P Q -> down members gamma = +-sqrt(mu),
// sorted descending; src[2i] = pair, src[2i+1] = pair | sign bit.
__global__ __launch_bounds__(256) void k_sort_modes(int n4, int reduced, const zc* __restrict__ W,
                                                    zc* __restrict__ gam, int* __restrict__ src) {
  __shared__ zc val[SORT_MAX];
  __shared__ int idx[SORT_MAX];
  const int l = blockIdx.x;
  const int n2 = n4 / 2;
  const int cnt = reduced ? n2 : n4;
  W += (long)l * (reduced ? n2 : n4);
  gam += (long)l * n4;
  src += (long)l * n4;
  int np = 1;
  while (np < cnt) np <<= 1;
  for (int i = threadIdx.x; i < np; i += blockDim.x) {
    if (i < cnt) {
      zc v = W[i];
      if (reduced) {
        zc g = zsqrt(v);
        const zc mg = -g;
        // the "down" member of the pair is the larger one under the reference comparator
        if (sorter_less(g, mg)) g = mg;
        v = g;
      }
      val[i] = v;
      idx[i] = i;
    } else {
      val[i] = mk(0.0);
      idx[i] = -1 - i;
    }
  }
  bitonic_sort(val, idx, np);
  for (int i = threadIdx.x; i < n2; i += blockDim.x) {
    if (reduced) {
      // descending order of the down members: position cnt-1-i
      const int j = cnt - 1 - i;
      gam[2 * i] = val[j];
      gam[2 * i + 1] = -val[j];
      src[2 * i] = idx[j];
      src[2 * i + 1] = idx[j] | (1 << 30);
    } else {
      gam[2 * i] = val[n4 - 1 - i];
      gam[2 * i + 1] = val[i];
      src[2 * i] = idx[n4 - 1 - i];
      src[2 * i + 1] = idx[i];
    }
  }
}

// Gd/Gc per layer (ScatterVolume.cpp:297-321) from the sorted modes.
// Full mode: column j of V = Vsrc(:, src[j]).  Reduced: column = [a; +-(Q a)/gamma] with a the
// eigenvector of P Q for the pair and gamma the down member.
__global__ void k_build_gdgc(int n4, int reduced, const zc* __restrict__ V, const zc* __restrict__ QA,
                             const zc* __restrict__ gam, const int* __restrict__ src, double k,
                             const double* __restrict__ dz, zc* __restrict__ Gd, zc* __restrict__ Gc) {
  const int l = blockIdx.y;
  const int n2 = n4 / 2;
  const long nn = (long)n4 * n4;
  gam += (long)l * n4;
  src += (long)l * n4;
  Gd += l * nn;
  Gc += l * nn;
  const zc* Vl = V + (long)l * (reduced ? (long)n2 * n2 : nn);
  const zc* QAl = reduced ? QA + (long)l * n2 * n2 : nullptr;
  const double dzl = dz[l];
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < nn; e += (long)gridDim.x * blockDim.x) {
    const int r = (int)(e % n4), j = (int)(e / n4);
    const zc g = gam[j];
    const int s = src[j];
    zc v;
    if (!reduced) {
      v = Vl[(long)s * n4 + r];
    } else {
      const int pair = s & ((1 << 30) - 1);
      const bool neg = (s >> 30) & 1;
      if (r < n2) {
        v = Vl[(long)pair * n2 + r];
      } else {
        const zc gd = neg ? -g : g;  // the down member of the pair
        v = zdiv(QAl[(long)pair * n2 + (r - n2)], gd);
        if (neg) v = -v;
      }
    }
    const zc ik = mk(-g.im * k * dzl, g.re * k * dzl);  // i k gamma dz
    if (j & 1) {
      Gd[e] = v * zexp(-ik);
      Gc[e] = v;
    } else {
      Gd[e] = v;
      Gc[e] = v * zexp(ik);
    }
  }
}

__global__ void k_copy(long n, const zc* __restrict__ a, zc* __restrict__ b) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}

struct BoundaryArgs {
  int MF;
  double k;
  double z0;
  const zc* Sx;
  const zc* Sy;
  const zc* Sz0;
  const zc* Sz1;
  const zc* EF;
};

// f1/f2/f3 entries (ScatterVolume.cpp:339-395); each is 4MF x 3MF with diagonal MF blocks.
__device__ inline zc f_entry(const BoundaryArgs& a, int which, int r, int c) {
  const int MF = a.MF;
  const int rb = r / MF, cb = c / MF, m = r % MF;
  if ((c % MF) != m) return mk(0.0);
  zc ph = mk(1.0);
  if (which == 1) {
    const zc sz = a.Sz0[m];
    ph = zexp(mk(-a.k * a.z0 * sz.im, a.k * a.z0 * sz.re));  // exp(i k z0 Sz0)
  }
  const zc sx = a.Sx[m], sy = a.Sy[m], s0 = a.Sz0[m], s1 = a.Sz1[m];
  switch (which) {
    case 1:
      if (rb == 0 && cb == 0) return ph;
      if (rb == 1 && cb == 1) return ph;
      if (rb == 2 && cb == 1) return -(ph * s0);
      if (rb == 2 && cb == 2) return ph * sy;
      if (rb == 3 && cb == 0) return ph * s0;
      if (rb == 3 && cb == 2) return -(ph * sx);
      return mk(0.0);
    case 2:
      if (rb == 0 && cb == 0) return mk(1.0);
      if (rb == 1 && cb == 1) return mk(1.0);
      if (rb == 2 && cb == 1) return s0;
      if (rb == 2 && cb == 2) return sy;
      if (rb == 3 && cb == 0) return -s0;
      if (rb == 3 && cb == 2) return -sx;
      return mk(0.0);
    default:
      if (rb == 0 && cb == 0) return mk(-1.0);
      if (rb == 1 && cb == 1) return mk(-1.0);
      if (rb == 2 && cb == 1) return s1;
      if (rb == 2 && cb == 2) return -sy;
      if (rb == 3 && cb == 0) return -s1;
      if (rb == 3 && cb == 2) return sx;
      return mk(0.0);
  }
}

__global__ void k_build_f(BoundaryArgs a, int which, zc* __restrict__ F) {
  const int n4 = 4 * a.MF, n3 = 3 * a.MF;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)n4 * n3; e += (long)gridDim.x * blockDim.x)
    F[e] = f_entry(a, which, (int)(e % n4), (int)(e / n4));
}

// A (6MF x 6MF) and b (ScatterVolume.cpp:126-131, 398-456): Gauss rows, [f2 | Gd0 Gc^-1 f3], -f1 EF.
__global__ void k_build_system(BoundaryArgs a, const zc* __restrict__ Abr, zc* __restrict__ A,
                               zc* __restrict__ b) {
  const int MF = a.MF, n6 = 6 * MF, n4 = 4 * MF, n3 = 3 * MF;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)n6 * n6; e += (long)gridDim.x * blockDim.x) {
    const int r = (int)(e % n6), c = (int)(e / n6);
    zc v = mk(0.0);
    if (r < MF) {
      const int m = r;
      if (c == m) v = a.Sx[m];
      else if (c == MF + m) v = a.Sy[m];
      else if (c == 2 * MF + m) v = -a.Sz0[m];
    } else if (r < 2 * MF) {
      const int m = r - MF;
      if (c == 3 * MF + m) v = a.Sx[m];
      else if (c == 4 * MF + m) v = a.Sy[m];
      else if (c == 5 * MF + m) v = a.Sz1[m];
    } else {
      const int rr = r - 2 * MF;
      if (c < n3) v = f_entry(a, 2, rr, c);
      else v = Abr[(long)(c - n3) * n4 + rr];
    }
    A[e] = v;
  }
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n6; r += gridDim.x * blockDim.x) {
    zc v = mk(0.0);
    if (r >= 2 * MF) {
      const int rr = r - 2 * MF;
      for (int c = 0; c < n3; ++c) {
        const zc f = f_entry(a, 1, rr, c);
        if (f.re != 0.0 || f.im != 0.0) v = v + f * a.EF[c];
      }
      v = -v;
    }
    b[r] = v;
  }
}

// rhs of beta_0: f1 EF + f2 R  (ScatterVolume.cpp:767-772)
__global__ void k_beta0_rhs(BoundaryArgs a, const zc* __restrict__ x, zc* __restrict__ out) {
  const int n4 = 4 * a.MF, n3 = 3 * a.MF;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n4; r += gridDim.x * blockDim.x) {
    zc v = mk(0.0);
    for (int c = 0; c < n3; ++c) {
      const zc f1 = f_entry(a, 1, r, c), f2 = f_entry(a, 2, r, c);
      if (f1.re != 0.0 || f1.im != 0.0) v = v + f1 * a.EF[c];
      if (f2.re != 0.0 || f2.im != 0.0) v = v + f2 * x[c];
    }
    out[r] = v;
  }
}

// Q_check / Q_hat rows 0..2MF and the Ez right-hand side Y (ScatterVolume.cpp:823-881):
// Qch = [V_even diag(beta_even) | V_odd diag(beta_odd)] rows 0..2MF (3MF x 4MF per layer);
// Y(p', c) = -(1/k) (up_p' Hy(p', c) - wq_p' Hx(p', c)), MF x 4MF per layer.
__global__ void k_q_rows(int MF, double k, const zc* __restrict__ Gd, const zc* __restrict__ Gc,
                         const zc* __restrict__ beta, const double* __restrict__ upz,
                         const double* __restrict__ wqz, zc* __restrict__ Qch, zc* __restrict__ Y) {
  const int n4 = 4 * MF, n3 = 3 * MF, n2 = 2 * MF;
  const int l = blockIdx.y;
  const long nn = (long)n4 * n4;
  Gd += l * nn;
  Gc += l * nn;
  beta += (long)l * n4;
  Qch += (long)l * n3 * n4;
  Y += (long)l * MF * n4;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)n4 * n4; e += (long)gridDim.x * blockDim.x) {
    const int r = (int)(e % n4), c = (int)(e / n4);
    // column c of [Q1 | Q2]: even mode 2c (c < 2MF) from Gd, odd mode 2(c-2MF)+1 from Gc
    const int mode = (c < n2) ? 2 * c : 2 * (c - n2) + 1;
    const zc v = ((c < n2) ? Gd : Gc)[(long)mode * n4 + r] * beta[mode];
    if (r < n2) Qch[(long)c * n3 + r] = v;
    else if (r >= n3) {
      const int p = r - n3;  // Hy row p; pair with the Hx row 2MF + p
      const zc hx = ((c < n2) ? Gd : Gc)[(long)mode * n4 + n2 + p] * beta[mode];
      const zc t = upz[p] * v - wqz[p] * hx;
      Y[(long)c * MF + p] = (-1.0 / k) * t;
    }
  }
}

// Ez coefficient matrix: Cn(p, p') = NIf[linear index (indR+M1/2)%M1 * M0 + (indC+M0/2)%M0] read as
// a column-major (M1 x M0) array (ScatterVolume.cpp:864-875; the transposed read is kept).
__global__ void k_cn(int M0, int M1, const zc* __restrict__ NIF, zc* __restrict__ Cn) {
  const int MF = M0 * M1;
  const int l = blockIdx.y;
  NIF += (long)l * MF;
  Cn += (long)l * MF * MF;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)MF * MF; e += (long)gridDim.x * blockDim.x) {
    const int p = (int)(e % MF), pp = (int)(e / MF);
    const int qi = p / M0, pi = p % M0, qj = pp / M0, pj = pp % M0;
    const int indR = (qi - qj + M1) % M1, indC = (pi - pj + M0) % M0;
    const int lin = ((indR + M1 / 2) % M1) * M0 + (indC + M0 / 2) % M0;
    Cn[e] = NIF[lin];
  }
}

__global__ void k_zero_rows(int rows0, int nrows, int ld, int ncols, zc* __restrict__ A, long sA) {
  const int l = blockIdx.y;
  A += l * sA;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)nrows * ncols; e += (long)gridDim.x * blockDim.x)
    A[(e / nrows) * ld + rows0 + e % nrows] = mk(0.0);
}

// .cw records (ScatterVolume.cpp:907-987): one launch covers every list; segment s of kind
//   0: Pi, 1: reflected (boundary 0), 2: transmitted (last boundary),
//   3: down modes of layer `layer` wound at boundary z, 4: up modes of layer `layer` wound at z.
struct Seg {
  long off, cnt;
  int kind, layer;
  double zb;
};

struct RecArgs {
  int MF;
  double k, n0, n1;
  const zc* Sx;
  const zc* Sy;
  const zc* Sz0;
  const zc* Sz1;
  const zc* EF;
  const zc* x;
  const zc* gam;
  const zc* Qch;
};

__global__ void k_records(RecArgs a, const Seg* __restrict__ segs, zc* __restrict__ rec) {
  const Seg sg = segs[blockIdx.y];
  const int MF = a.MF, n2 = 2 * MF, n3 = 3 * MF, n4 = 4 * MF;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < sg.cnt; i += (long)gridDim.x * blockDim.x) {
    zc E[3], K[3];
    int p;
    if (sg.kind <= 2) p = (int)i;
    else p = (int)(i / n2);
    K[0] = a.Sx[p] * a.k;
    K[1] = a.Sy[p] * a.k;
    bool wind = true;
    if (sg.kind == 0) {
      K[2] = a.Sz0[p] * a.k;
      for (int c = 0; c < 3; ++c) E[c] = a.EF[c * MF + p];
      wind = false;
    } else if (sg.kind == 1) {
      K[2] = -(a.Sz0[p] * a.k) * a.n0;
      for (int c = 0; c < 3; ++c) E[c] = a.x[c * MF + p];
    } else if (sg.kind == 2) {
      K[2] = (a.Sz1[p] * a.k) * a.n1;
      for (int c = 0; c < 3; ++c) E[c] = a.x[n3 + c * MF + p];
    } else {
      const int jj = (int)(i % n2);
      const bool up = (sg.kind == 4);
      const zc g = a.gam[(long)sg.layer * n4 + 2 * jj + (up ? 1 : 0)];
      K[2] = g * a.k;
      const zc* Q = a.Qch + (long)sg.layer * n3 * n4 + (long)(jj + (up ? n2 : 0)) * n3;
      for (int c = 0; c < 3; ++c) E[c] = Q[c * MF + p];
    }
    if (wind) {
      // planewave::wind(0, 0, -z): E0 *= exp(i kz (-z))
      const zc arg = mk(K[2].im * sg.zb, -K[2].re * sg.zb);
      const zc ph = zexp(arg);
      for (int c = 0; c < 3; ++c) E[c] = E[c] * ph;
    }
    zc* o = rec + (sg.off + i) * 6;
    for (int c = 0; c < 3; ++c) { o[c] = E[c]; o[3 + c] = K[c]; }
  }
}

template <typename T>
void upload(DevBuf<T>& d, const std::vector<T>& h, hipStream_t s) {
  d.ensure(h.size());
  if (!h.empty()) CWX_HIP(hipMemcpyAsync(d.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, s));
}

unsigned grid_for(long n) { return (unsigned)std::min<long>(65535, std::max<long>(1, (n + 255) / 256)); }


// Fourier coefficients of n^2 (layers 0..L-1) and n^-2 (layers L..2L-1) of every layer, each an
// (M1 x M0) column-major block, by a pruned DFT: F = w1 (M1 x s1) . A (s1 x s2) . w2 (s2 x M0) / (s1 s2)
// with A the column-major layer matrix (FourierWave.h:152-195, 288, 341-343).
void fourier_stage(const Setup& st, const zc* d_sample, zc* F, hipStream_t s) {
  const int L = st.L, s1 = st.s1, s2 = st.s2, M0 = st.M0, M1 = st.M1, MF = st.MF;
  std::vector<zc> w1((size_t)M1 * s1), w2((size_t)s2 * M0);
  for (int i = 0; i < s1; ++i)
    for (int r = 0; r < M1; ++r) {
      const long t = ((long)st.rows[r] * i) % s1;
      const double ang = -2.0 * kPi * (double)t / (double)s1;
      w1[(size_t)i * M1 + r] = mk(std::cos(ang) / ((double)s1 * s2), std::sin(ang) / ((double)s1 * s2));
    }
  for (int c = 0; c < M0; ++c)
    for (int j = 0; j < s2; ++j) {
      const long t = ((long)st.cols[c] * j) % s2;
      const double ang = -2.0 * kPi * (double)t / (double)s2;
      w2[(size_t)c * s2 + j] = mk(std::cos(ang), std::sin(ang));
    }
  DevBuf<zc> d_w1, d_w2;
  upload(d_w1, w1, s);
  upload(d_w2, w2, s);
  const long npix = (long)s1 * s2;
  DevBuf<zc> sq(2 * npix * L), T1((size_t)2 * L * s1 * M0);
  hipLaunchKernelGGL(k_square_layers, dim3(grid_for(npix * L)), dim3(256), 0, s, npix * L, d_sample, sq.p);
  zgemm_batched(OpN, OpN, s1, M0, s2, mk(1.0), sq.p, s1, npix, d_w2.p, s2, 0, mk(0.0), T1.p, s1,
                (long)s1 * M0, 2 * L, s);
  zgemm_batched(OpN, OpN, M1, M0, s1, mk(1.0), d_w1.p, M1, 0, T1.p, s1, (long)s1 * M0, mk(0.0), F, M1,
                MF, 2 * L, s);
  // keep the temporaries alive until the kernels that read them have run
  CWX_HIP(hipStreamSynchronize(s));
}

}  // namespace

// Phi of every layer (blocks = 0: full 4MF x 4MF) or its P and Q blocks (blocks = 1: 2 x L x (2MF)^2,
// all P first) into device memory.  Returns the setup.
Setup build_phi_dev(const VolumeOptions& o, const zc* d_sample, int L, int s1, int s2, int blocks, zc* out,
                    hipStream_t s) {
  Setup st = make_setup(o, L, s1, s2);
  const int MF = st.MF;
  DevBuf<zc> F((size_t)2 * L * MF);
  fourier_stage(st, d_sample, F.p, s);
  DevBuf<double> d_up, d_wq;
  upload(d_up, st.up, s);
  upload(d_wq, st.wq, s);
  PhiArgs pa{st.M0, st.M1, MF, 1.0 / (st.k * st.k), d_up.p, d_wq.p};
  const long n4 = 4L * MF, n2 = 2L * MF;
  if (blocks == 0)
    hipLaunchKernelGGL(k_phi_full, dim3(grid_for(n4 * n4), L), dim3(256), 0, s, pa, F.p, F.p + (long)L * MF, out);
  else
    hipLaunchKernelGGL(k_phi_blocks, dim3(grid_for(n2 * n2), L), dim3(256), 0, s, pa, F.p, F.p + (long)L * MF,
                       out, out + L * n2 * n2);
  CWX_HIP(hipGetLastError());
  CWX_HIP(hipStreamSynchronize(s));
  return st;
}

// ------------------------------------------------------------------------------------------------
// the pipeline
// ------------------------------------------------------------------------------------------------
void solve_volume(Context& ctx, const VolumeOptions& o, const zc* d_sample, int L, int s1, int s2,
                  Solution& sol) {
  hipStream_t s = ctx.stream;
  sol.opt = o;
  sol.st = make_setup(o, L, s1, s2);
  const Setup& st = sol.st;
  const int M0 = st.M0, M1 = st.M1, MF = st.MF;
  const int n2 = 2 * MF, n3 = 3 * MF, n4 = 4 * MF, n6 = 6 * MF;
  const long nn = (long)n4 * n4;
  const double k = st.k;
  sol.n4 = n4;
  sol.info = 0;
  sol.info_msg.clear();
  CWX_CHECK(n4 <= SORT_MAX, "4*M0*M1 exceeds the supported maximum of 2048");
  CWX_CHECK(L <= 65535, "too many layers");

  // small host tables -> device
  DevBuf<double> d_up, d_wq, d_dz, d_upz, d_wqz;
  upload(d_up, st.up, s);
  upload(d_wq, st.wq, s);
  std::vector<double> dz(L);
  for (int i = 0; i < L; ++i) dz[i] = (L == 1) ? (st.z[L] - st.z[0]) : (st.z[i + 1] - st.z[i]);
  upload(d_dz, dz, s);
  // Pz uses M_PI in UP/WQ (ScatterVolume.cpp:813-814), indexed by the column p' = qj*M0 + pj
  std::vector<double> upz(MF), wqz(MF);
  for (int q = 0; q < M1; ++q)
    for (int pp = 0; pp < M0; ++pp) {
      upz[q * M0 + pp] = 2.0 * (double)(pp - M0 / 2) * kPi / o.size[0] + st.dir_n[0] * k;
      wqz[q * M0 + pp] = 2.0 * (double)(q - M1 / 2) * kPi / o.size[1] + st.dir_n[1] * k;
    }
  upload(d_upz, upz, s);
  upload(d_wqz, wqz, s);
  // Sx | Sy | Sz0 | Sz1 | EF (3MF), kept with the solution for the separable field evaluation
  {
    std::vector<zc> tabs;
    tabs.reserve(7 * (size_t)MF);
    for (auto* v : {&st.Sx, &st.Sy, &st.Sz0, &st.Sz1, &st.EF}) tabs.insert(tabs.end(), v->begin(), v->end());
    upload(sol.tabs, tabs, s);
  }
  const zc* dSx = sol.tabs.p;
  const zc* dSy = sol.tabs.p + MF;
  const zc* dSz0 = sol.tabs.p + 2 * MF;
  const zc* dSz1 = sol.tabs.p + 3 * MF;
  const zc* dEF = sol.tabs.p + 4 * MF;
  DevBuf<zc> F((size_t)2 * L * MF);
  fourier_stage(st, d_sample, F.p, s);
  const zc* NF = F.p;
  const zc* NI = F.p + (long)L * MF;
  sol.NIF.alloc((size_t)L * MF);
  CWX_HIP(hipMemcpyAsync(sol.NIF.p, NI, sizeof(zc) * L * MF, hipMemcpyDeviceToDevice, s));

  // ---- 2-3. Phi and its eigen-decomposition ----
  PhiArgs pa{M0, M1, MF, 1.0 / (k * k), d_up.p, d_wq.p};
  DevBuf<int> d_src((size_t)L * n4);
  sol.gam.alloc((size_t)L * n4);
  DevBuf<int> d_info(L);
  DevBuf<zc> W, V, QA;
  ctx.w0.ensure(std::max<size_t>((size_t)L * nn, 1));
  if (o.eig_mode == 0) {
    zc* Phi = ctx.w0.p;
    hipLaunchKernelGGL(k_phi_full, dim3(grid_for(nn), L), dim3(256), 0, s, pa, NF, NI, Phi);
    W.alloc((size_t)L * n4);
    V.alloc((size_t)L * nn);
    zgeev_batched(n4, Phi, n4, nn, W.p, n4, V.p, n4, nn, L, ctx.eig, d_info.p, s);
    hipLaunchKernelGGL(k_sort_modes, dim3(L), dim3(256), 0, s, n4, 0, W.p, sol.gam.p, d_src.p);
  } else {
    const long mm = (long)n2 * n2;
    zc* P = ctx.w0.p;
    zc* Q = ctx.w0.p + L * mm;
    ctx.w1.ensure((size_t)L * mm);
    zc* PQ = ctx.w1.p;
    hipLaunchKernelGGL(k_phi_blocks, dim3(grid_for(mm), L), dim3(256), 0, s, pa, NF, NI, P, Q);
    zgemm_batched(OpN, OpN, n2, n2, n2, mk(1.0), P, n2, mm, Q, n2, mm, mk(0.0), PQ, n2, mm, L, s);
    W.alloc((size_t)L * n2);
    V.alloc((size_t)L * mm);
    zgeev_batched(n2, PQ, n2, mm, W.p, n2, V.p, n2, mm, L, ctx.eig, d_info.p, s);
    QA.alloc((size_t)L * mm);
    zgemm_batched(OpN, OpN, n2, n2, n2, mk(1.0), Q, n2, mm, V.p, n2, mm, mk(0.0), QA.p, n2, mm, L, s);
    hipLaunchKernelGGL(k_sort_modes, dim3(L), dim3(256), 0, s, n4, 1, W.p, sol.gam.p, d_src.p);
  }
  sol.Gd.alloc((size_t)L * nn);
  sol.Gc.alloc((size_t)L * nn);
  hipLaunchKernelGGL(k_build_gdgc, dim3(grid_for(nn), L), dim3(256), 0, s, n4, o.eig_mode, V.p, QA.p,
                     sol.gam.p, d_src.p, k, d_dz.p, sol.Gd.p, sol.Gc.p);
  V.release();
  QA.release();

  // ---- 4. LU of every Gd_i, T_i = Gd_i^-1 Gc_{i-1}, S = Gc_{L-1} T_{L-1} ... T_1 ----
  zc* LUd = ctx.w0.p;  // Phi / P,Q no longer needed
  hipLaunchKernelGGL(k_copy, dim3(grid_for(L * nn)), dim3(256), 0, s, L * nn, sol.Gd.p, LUd);
  ctx.ipiv.ensure((size_t)L * n4 + n4 + n6);
  int* piv = ctx.ipiv.p;
  DevBuf<int> d_info_lu(L), info2(4);
  zgetrf_batched(n4, LUd, n4, nn, piv, n4, d_info_lu.p, L, s);
  ctx.w1.ensure((size_t)std::max(1, L - 1) * nn);
  zc* Tm = ctx.w1.p;
  if (L > 1) {
    hipLaunchKernelGGL(k_copy, dim3(grid_for((L - 1) * nn)), dim3(256), 0, s, (L - 1) * nn, sol.Gc.p, Tm);
    zgetrs_batched(n4, n4, LUd + nn, n4, nn, piv + n4, n4, Tm, n4, nn, L - 1, s);
  }
  ctx.w2.ensure((size_t)nn);
  ctx.w3.ensure((size_t)nn);
  zc* S = ctx.w2.p;
  if (L == 1) {
    hipLaunchKernelGGL(k_copy, dim3(grid_for(nn)), dim3(256), 0, s, nn, sol.Gc.p, S);
  } else {
    // Pm = T_{L-1} ... T_1, accumulated left to right
    zc* Pm = ctx.w3.p;
    hipLaunchKernelGGL(k_copy, dim3(grid_for(nn)), dim3(256), 0, s, nn, Tm, Pm);
    for (int i = 2; i <= L - 1; ++i) {
      zgemm_batched(OpN, OpN, n4, n4, n4, mk(1.0), Tm + (long)(i - 1) * nn, n4, 0, Pm, n4, 0, mk(0.0), S,
                    n4, 0, 1, s);
      std::swap(Pm, S);
    }
    zgemm_batched(OpN, OpN, n4, n4, n4, mk(1.0), sol.Gc.p + (long)(L - 1) * nn, n4, 0, Pm, n4, 0, mk(0.0),
                  S == Pm ? ctx.w3.p : S, n4, 0, 1, s);
    if (S == Pm) S = ctx.w3.p;
  }

  // ---- 5. boundary system: A = [Gauss; f2 | Gd0 S^-1 f3], b = [0; -f1 EF] ----
  BoundaryArgs ba{MF, k, st.z[0], dSx, dSy, dSz0, dSz1, dEF};
  DevBuf<zc> Y((size_t)n4 * n3), Abr((size_t)n4 * n3), A((size_t)n6 * n6);
  sol.x.alloc(n6);
  int* pivS = piv + (long)L * n4;
  int* pivA = pivS + n4;
  zgetrf_batched(n4, S, n4, nn, pivS, n4, info2.p, 1, s);
  hipLaunchKernelGGL(k_build_f, dim3(grid_for((long)n4 * n3)), dim3(256), 0, s, ba, 3, Y.p);
  zgetrs_batched(n4, n3, S, n4, nn, pivS, n4, Y.p, n4, (long)n4 * n3, 1, s);
  zgemm_batched(OpN, OpN, n4, n3, n4, mk(1.0), sol.Gd.p, n4, 0, Y.p, n4, 0, mk(0.0), Abr.p, n4, 0, 1, s);
  hipLaunchKernelGGL(k_build_system, dim3(grid_for((long)n6 * n6)), dim3(256), 0, s, ba, Abr.p, A.p, sol.x.p);
  zgetrf_batched(n6, A.p, n6, (long)n6 * n6, pivA, n6, info2.p + 1, 1, s);
  zgetrs_batched(n6, 1, A.p, n6, (long)n6 * n6, pivA, n6, sol.x.p, n6, n6, 1, s);

  // ---- 6. beta recurrence, Q_check / Q_hat, Ez rows ----
  if (!o.external) {
    sol.beta.alloc((size_t)L * n4);
    hipLaunchKernelGGL(k_beta0_rhs, dim3(grid_for(n4)), dim3(256), 0, s, ba, sol.x.p, sol.beta.p);
    zgetrs_batched(n4, 1, LUd, n4, nn, piv, n4, sol.beta.p, n4, n4, 1, s);
    for (int i = 1; i < L; ++i)
      zgemm_batched(OpN, OpN, n4, 1, n4, mk(1.0), Tm + (long)(i - 1) * nn, n4, 0, sol.beta.p + (long)(i - 1) * n4,
                    n4, 0, mk(0.0), sol.beta.p + (long)i * n4, n4, 0, 1, s);
    sol.Qch.alloc((size_t)L * n3 * n4);
    DevBuf<zc> Yz((size_t)L * MF * n4), Cn((size_t)L * MF * MF);
    hipLaunchKernelGGL(k_q_rows, dim3(grid_for(nn), L), dim3(256), 0, s, MF, k, sol.Gd.p, sol.Gc.p,
                       sol.beta.p, d_upz.p, d_wqz.p, sol.Qch.p, Yz.p);
    if (o.npz) {
      hipLaunchKernelGGL(k_zero_rows, dim3(grid_for((long)MF * n4), L), dim3(256), 0, s, n2, MF, n3, n4,
                         sol.Qch.p, (long)n3 * n4);
    } else {
      hipLaunchKernelGGL(k_cn, dim3(grid_for((long)MF * MF), L), dim3(256), 0, s, M0, M1, sol.NIF.p, Cn.p);
      zgemm_batched(OpN, OpN, MF, n4, MF, mk(1.0), Cn.p, MF, (long)MF * MF, Yz.p, MF, (long)MF * n4, mk(0.0),
                    sol.Qch.p + n2, n3, (long)n3 * n4, L, s);
    }
  }

  // ---- 7. .cw records ----
  const int ZL = L + 1;
  std::vector<Seg> segs;
  long off = 0;
  auto add = [&](long cnt, int kind, int layer, double zb) {
    segs.push_back(Seg{off, cnt, kind, layer, zb});
    sol.seg_off.push_back(off);
    sol.seg_cnt.push_back(cnt);
    off += cnt;
  };
  sol.seg_off.clear();
  sol.seg_cnt.clear();
  const long inner = o.external ? 0 : (long)n2 * MF;
  add(MF, 0, 0, 0.0);  // Pi
  for (int l = 0; l < ZL; ++l) {
    if (l == 0) {
      add(MF, 1, 0, st.z[0]);
      add(inner, 3, 0, st.z[0]);
    } else if (l == ZL - 1) {
      add(inner, 4, l - 1, st.z[l]);
      add(MF, 2, 0, st.z[l]);
    } else {
      add(inner, 4, l - 1, st.z[l]);
      add(inner, 3, l, st.z[l]);
    }
  }
  sol.nrec = off;
  sol.rec.alloc((size_t)off * 6);
  DevBuf<Seg> d_segs;
  upload(d_segs, segs, s);
  long maxcnt = 1;
  for (auto& g : segs) maxcnt = std::max(maxcnt, g.cnt);
  RecArgs ra{MF, k, o.n[0], o.n[1], dSx, dSy, dSz0, dSz1, dEF, sol.x.p, sol.gam.p,
             o.external ? nullptr : sol.Qch.p};
  hipLaunchKernelGGL(k_records, dim3(grid_for(maxcnt), (unsigned)segs.size()), dim3(256), 0, s, ra, d_segs.p,
                     sol.rec.p);
  CWX_HIP(hipGetLastError());

  // ---- status ----
  std::vector<int> hinfo(L), hlu(L), hinfo2(2);
  CWX_HIP(hipMemcpyAsync(hinfo.data(), d_info.p, sizeof(int) * L, hipMemcpyDeviceToHost, s));
  CWX_HIP(hipMemcpyAsync(hlu.data(), d_info_lu.p, sizeof(int) * L, hipMemcpyDeviceToHost, s));
  CWX_HIP(hipMemcpyAsync(hinfo2.data(), info2.p, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
  CWX_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < L; ++i)
    if (hinfo[i] != 0) {
      sol.info = hinfo[i];
      sol.info_msg = "layer " + std::to_string(i) + ": eigensolver did not converge (info " +
                     std::to_string(hinfo[i]) + ")";
      break;
    }
  for (int i = 0; i < L && !sol.info; ++i)
    if (hlu[i] != 0) {
      sol.info = hlu[i];
      sol.info_msg = "layer " + std::to_string(i) + ": singular Gd (zero pivot " + std::to_string(hlu[i]) + ")";
    }
  if (!sol.info && hinfo2[0]) { sol.info = hinfo2[0]; sol.info_msg = "singular transfer product"; }
  if (!sol.info && hinfo2[1]) { sol.info = hinfo2[1]; sol.info_msg = "singular boundary system"; }
}

// ------------------------------------------------------------------------------------------------
// .cw file (CoupledWaveStructure.h:63-92 container behind an is_volume byte, 96-byte records)
// ------------------------------------------------------------------------------------------------
void cw_write_file(const Solution& sol, const std::string& path, hipStream_t s) {
  std::vector<zc> h((size_t)sol.nrec * 6);
  if (sol.nrec)
    CWX_HIP(hipMemcpyAsync(h.data(), sol.rec.p, sizeof(zc) * h.size(), hipMemcpyDeviceToHost, s));
  CWX_HIP(hipStreamSynchronize(s));
  std::ofstream f(path, std::ios::binary);
  CWX_CHECK((bool)f, "cannot open " + path + " for writing");
  const unsigned char is_volume = 0;
  f.write((const char*)&is_volume, 1);
  const uint64_t prec = 8;
  f.write((const char*)&prec, 8);
  const int L = sol.st.L, ZL = L + 1;
  auto put_list = [&](int seg) {
    const uint64_t cnt = (uint64_t)sol.seg_cnt[seg];
    f.write((const char*)&cnt, 8);
    f.write((const char*)(h.data() + sol.seg_off[seg] * 6), (std::streamsize)(cnt * 96));
  };
  put_list(0);
  const uint64_t nb = ZL;
  f.write((const char*)&nb, 8);
  for (int l = 0; l < ZL; ++l) {
    // --external leaves the interior boundaries' z at its value-initialised 0 (ScatterVolume.cpp:953-969)
    const double zb = (sol.opt.external && l > 0 && l < ZL - 1) ? 0.0 : sol.st.z[l];
    f.write((const char*)&zb, 8);
    put_list(1 + 2 * l);
    put_list(2 + 2 * l);
  }
  CWX_CHECK((bool)f, "write to " + path + " failed");
}

}  // namespace cwx
