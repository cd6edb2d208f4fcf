// The scattervolume coupled-wave solve on one GPU.
//
// Host part restates ScatterVolume.cpp:574-685 (options, orthogonalize, orders, z, Sx/Sy/Sz, EF);
// the device pipeline replaces FourierWave.h:303-397 (Phi), ScatterVolume.cpp:235-466 (eigen,
// sort, Gd/Gc chain, boundary system), :739-889 (zgesv, beta, Q_check/Q_hat, Pz) and :907-987
// (CoupledWaveStructure assembly).  All per-layer work is batched over layers.
#include "cwx_solver.h"
#include "cwx_tune.h"

#include <algorithm>
#include <array>
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

// Per-instance tables of a set of samples solved together (one entry per instance unless noted).
// Every per-layer kernel finds its instance through linst.
struct InstArgs {
  int M0, M1, MF;
  const int* linst;    // layer -> instance
  const double* k;     // 2 PI / lambda
  const double* up;    // count x M0
  const double* wq;    // count x M1
  const double* upz;   // count x MF: UP with M_PI, indexed p' = qj*M0 + pj (ScatterVolume.cpp:813-814)
  const double* wqz;   // count x MF
  const double* z0;    // first boundary
  const double* nio;   // count x 2: n0, n1 (the extra factor on the external kz)
  const zc* tabs;      // count x 7MF: Sx | Sy | Sz0 | Sz1 | EF (3MF)
};

struct PhiLocal {
  int M0, M1, MF;
  double kinv;
  const double* up;  // M0
  const double* wq;  // M1
};

__device__ inline PhiLocal phi_local(const InstArgs& a, int layer) {
  const int i = a.linst[layer];
  const double k = a.k[i];
  return PhiLocal{a.M0, a.M1, a.MF, 1.0 / (k * k), a.up + (long)i * a.M0, a.wq + (long)i * a.M1};
}

// Element (r, li) coefficient indices of the circulant blocks (FourierWave.h:366-372).
__device__ inline int circ_index(int r, int li, int M0, int M1) {
  const int qi = r / M0, pi = r % M0, qj = li / M0, pj = li % M0;
  const int indR = (qi - qj + M1) % M1, indC = (pi - pj + M0) % M0;
  const int rr = (indR + M1 / 2) % M1, cc = (indC + M0 / 2) % M0;
  return rr + cc * M1;  // column-major (M1 x M0)
}

// P (rows 0..2MF, cols 2MF..4MF of Phi) and Q (rows 2MF..4MF, cols 0..2MF), FourierWave.h:378-392.
__device__ inline zc phi_P(const PhiLocal& a, const zc* NI, int rr, int cc) {
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

__device__ inline zc phi_Q(const PhiLocal& a, const zc* NF, int rr, int cc) {
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

// full Phi (4MF x 4MF, column-major) per layer; with a list, slot b holds layer list[b]
__global__ void k_phi_full(InstArgs ia, const zc* __restrict__ NF, const zc* __restrict__ NI, zc* Phi,
                           const int* __restrict__ list) {
  const int l = list ? list[blockIdx.y] : blockIdx.y;
  const PhiLocal a = phi_local(ia, l);
  const int n = 4 * a.MF, n2 = 2 * a.MF;
  NF += (long)l * a.MF;
  NI += (long)l * a.MF;
  Phi += (long)blockIdx.y * n * n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)n * n; e += (long)gridDim.x * blockDim.x) {
    const int r = (int)(e % n), c = (int)(e / n);
    zc v = mk(0.0);
    if (r < n2 && c >= n2) v = phi_P(a, NI, r, c - n2);
    else if (r >= n2 && c < n2) v = phi_Q(a, NF, r - n2, c);
    Phi[e] = v;
  }
}

// P and Q blocks (2MF x 2MF each) per layer
__global__ void k_phi_blocks(InstArgs ia, const zc* __restrict__ NF, const zc* __restrict__ NI, zc* P, zc* Q) {
  const int l = blockIdx.y;
  const PhiLocal a = phi_local(ia, l);
  const int n2 = 2 * a.MF;
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

constexpr int SORT_MAX = 4096;      // largest 4 M0 M1 (two sort capacities: 2048 and 4096 entries)
// The reduced (P Q) form takes gamma = +-sqrt(mu).  A pair with |gamma| below this fraction of the
// layer's largest |gamma| is nearly defective (gamma -> 0 merges the pair), and the layer is solved
// through the full 4MF Phi instead, the reference's own formulation.
constexpr double kReducedTau = 1e-4;

// Full mode: eigenvalues W (n4) of Phi -> source column of each sorted, interleaved position
// (ScatterVolume.cpp:190-230).  Reduced mode: mu (n2) of P Q -> down members gamma = +-sqrt(mu),
// sorted descending; src[2i] = pair, src[2i+1] = pair | sign bit; unsafe[l] = 1 when the layer's
// smallest |gamma| is below kReducedTau times its largest, or any gamma is not finite.  Workgroup b
// sorts the eigenvalues at slot b of W into layer list[b] (layer b without a list).
template <int CAP>
__global__ __launch_bounds__(256) void k_sort_modes(int n4, int reduced, const zc* __restrict__ W,
                                                    zc* __restrict__ gam, int* __restrict__ src,
                                                    const int* __restrict__ list, int* __restrict__ unsafe) {
  __shared__ zc val[CAP];
  __shared__ int idx[CAP];
  __shared__ double s_ext[2][4];
  const int l = list ? list[blockIdx.x] : blockIdx.x;
  const int n2 = n4 / 2;
  const int cnt = reduced ? n2 : n4;
  W += (long)blockIdx.x * (reduced ? n2 : n4);
  gam += (long)l * n4;
  src += (long)l * n4;
  int np = 1;
  while (np < cnt) np <<= 1;
  if (reduced && unsafe) {
    double gmin = INFINITY, gmax = 0.0;
    bool bad = false;
    for (int i = threadIdx.x; i < cnt; i += blockDim.x) {
      const double a = absz(zsqrt(W[i]));
      bad |= !isfinite(a);
      gmin = fmin(gmin, a);
      gmax = fmax(gmax, a);
    }
    for (int off = 32; off > 0; off >>= 1) {
      gmin = fmin(gmin, __shfl_xor(gmin, off));
      gmax = fmax(gmax, __shfl_xor(gmax, off));
    }
    const int any_bad = __syncthreads_or(bad ? 1 : 0);
    if ((threadIdx.x & 63) == 0) {
      s_ext[0][threadIdx.x >> 6] = gmin;
      s_ext[1][threadIdx.x >> 6] = gmax;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      double a = s_ext[0][0], b = s_ext[1][0];
      for (int w = 1; w < (int)(blockDim.x >> 6); ++w) {
        a = fmin(a, s_ext[0][w]);
        b = fmax(b, s_ext[1][w]);
      }
      unsafe[l] = (any_bad || !(a > kReducedTau * b)) ? 1 : 0;
    }
  }
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

// zgeev's normalisation of the reduced-form modes v+- = [a; +-(Q a)/gamma] (unit 2-norm, largest
// component real; the lowest index wins ties, as k_normalize).  Both members of a pair have the same
// norm and the same largest component k; for pair p of layer l, ps[l n2 + p] = {1/||v||, v_k of the +
// member, k}.  One wavefront per pair.
struct PairScale {
  double inv_norm;
  zc vk;
  int k;
};

__global__ __launch_bounds__(64) void k_pair_scale(int n2, const zc* __restrict__ V, const zc* __restrict__ QA,
                                                  const zc* __restrict__ W, PairScale* __restrict__ ps) {
  const int p = blockIdx.x, l = blockIdx.y, lane = threadIdx.x;
  const long mm = (long)n2 * n2;
  const zc* a = V + l * mm + (long)p * n2;
  const zc* qa = QA + l * mm + (long)p * n2;
  zc g = zsqrt(W[(long)l * n2 + p]);
  if (sorter_less(g, -g)) g = -g;  // the down member, as k_sort_modes takes it
  const double g2 = norm2(g);
  double sum = 0.0, best = -1.0;
  int bi = 0;
  for (int r = lane; r < 2 * n2; r += 64) {
    const double m2 = r < n2 ? norm2(a[r]) : norm2(qa[r - n2]) / g2;
    sum += m2;
    if (m2 > best) { best = m2; bi = r; }
  }
  for (int off = 32; off > 0; off >>= 1) {
    sum += __shfl_xor(sum, off);
    const double ob = __shfl_xor(best, off);
    const int oi = __shfl_xor(bi, off);
    if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
  }
  if (lane == 0) {
    const zc vk = bi < n2 ? a[bi] : zdiv(qa[bi - n2], g);
    ps[(long)l * n2 + p] = PairScale{1.0 / sqrt(sum), vk, bi};
  }
}

// Gd/Gc per layer (ScatterVolume.cpp:297-321) from the sorted modes.
// Full mode: column j of V = Vsrc(:, src[j]).  Reduced: column = [a; +-(Q a)/gamma] with a the
// eigenvector of P Q for the pair and gamma the down member, scaled by ps as zgeev normalises; a
// layer with fslot[l] >= 0 was solved through the full Phi instead, its vectors at slot fslot[l] of Vf.
__global__ void k_build_gdgc(InstArgs ia, int reduced, const zc* __restrict__ V, const zc* __restrict__ QA,
                             const PairScale* __restrict__ ps, const zc* __restrict__ Vf,
                             const int* __restrict__ fslot, const zc* __restrict__ gam, const int* __restrict__ src,
                             const double* __restrict__ dz, zc* __restrict__ Gd, zc* __restrict__ Gc) {
  const int l = blockIdx.y;
  const int n4 = 4 * ia.MF, n2 = n4 / 2;
  const long nn = (long)n4 * n4;
  const double k = ia.k[ia.linst[l]];
  gam += (long)l * n4;
  src += (long)l * n4;
  Gd += l * nn;
  Gc += l * nn;
  const int slot = (reduced && fslot) ? fslot[l] : -1;
  const bool red = reduced && slot < 0;
  const zc* Vl = red ? V + (long)l * n2 * n2 : (reduced ? Vf + (long)slot * nn : V + (long)l * nn);
  const zc* QAl = red ? QA + (long)l * n2 * n2 : nullptr;
  const PairScale* psl = red ? ps + (long)l * n2 : nullptr;
  const double dzl = dz[l];
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < nn; e += (long)gridDim.x * blockDim.x) {
    const int r = (int)(e % n4), j = (int)(e / n4);
    const zc g = gam[j];
    const int s = src[j];
    zc v;
    if (!red) {
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
      const PairScale sc = psl[pair];
      zc vk = sc.vk;
      if (neg && sc.k >= n2) vk = -vk;
      const zc rot = conj(vk) * (sc.inv_norm / absz(vk));
      v = v * rot;
      if (r == sc.k) v.im = 0.0;
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

// ---- the transfer chain in the reduced mode's block form ----
// In the reduced mode the mode matrix of a layer is V = M Pi R with M = [[A, A], [B, -B]], A = (a_p) the
// eigenvectors of P Q in sorted pair order, B = Q A Gamma^-1 (Gamma the down members), Pi the interleave
// (column 2i + s is block column s n2 + i; s = 0 down, 1 up) and R the diagonal of zgeev's normalisation
// (k_build_gdgc).  M^-1 = 1/2 [[A^-1, B^-1], [A^-1, -B^-1]], so with X = Vr_g^-1 Vr_{g-1} and
// Y = Vr_g^-1 Q_g^-1 QA_{g-1} (Vr the eigenvectors in zgeev's order, QA = Q Vr) the transfer matrix is
//   T_g(2i+s, 2j+t) = Dd_g^-1 R_g^-1 (2i+s) . 1/2 (X(p_i, q_j) +- gam_g(2i) Y(p_i, q_j) / gam_{g-1}(2j))
//                     . R_{g-1} Dc_{g-1} (2j+t)
// (+ for s = t), with p_i, q_j the pairs of sorted index i in layer g and j in layer g - 1, Dd and Dc the
// phase diagonals of Gd = V Dd and Gc = V Dc.  That is T_g = Gd_g^-1 Gc_{g-1} (ScatterVolume.cpp:327-334)
// from two LUs of order 2MF instead of one of order 4MF, to rounding: k_build_gdgc also sets the imaginary
// part of each column's largest entry to exactly zero.
__device__ inline zc mode_rot(const PairScale& sc, bool neg, int n2) {
  zc vk = sc.vk;
  if (neg && sc.k >= n2) vk = -vk;
  return conj(vk) * (sc.inv_norm / absz(vk));
}

// Workgroup (tile, t): rows [64 rt, 64 rt + 64) and columns [32 ct, 32 ct + 32) of T for transfer t, tile =
// rt + ct * (number of row tiles).  The row factors and column factors of the tile are formed once, in LDS; each lane then
// takes one row and eight columns.
constexpr int TB_R = 64, TB_C = 32;

__global__ __launch_bounds__(256) void k_t_block(InstArgs ia, const int* __restrict__ tg, MatSet Xs, MatSet Ys,
                                                 MatSet Ts, const zc* __restrict__ gam, const int* __restrict__ src,
                                                 const PairScale* __restrict__ ps, const double* __restrict__ dz) {
  __shared__ zc s_rf[TB_R], s_cf[TB_C];
  const int t = blockIdx.y, g = tg[t];
  const int n4 = 4 * ia.MF, n2 = n4 / 2;
  const int rtiles = (n4 + TB_R - 1) / TB_R;
  const int r0 = (int)(blockIdx.x % rtiles) * TB_R, c0 = (int)(blockIdx.x / rtiles) * TB_C;
  const double k = ia.k[ia.linst[g]];
  const zc* __restrict__ X = Xs.at(t);
  const zc* __restrict__ Y = Ys.at(t);
  zc* __restrict__ T = Ts.at(t);
  const zc* gg = gam + (long)g * n4;
  const zc* gp = gam + (long)(g - 1) * n4;
  const int* sg = src + (long)g * n4;
  const int* sp = src + (long)(g - 1) * n4;
  const PairScale* psg = ps + (long)g * n2;
  const PairScale* psp = ps + (long)(g - 1) * n2;
  const double dzg = dz[g], dzp = dz[g - 1];
  constexpr int kMask = (1 << 30) - 1;
  const int tid = threadIdx.x;
  if (tid < TB_R) {
    const int row = r0 + tid;
    if (row < n4) {
      // row factor Dd_g^-1 R_g^-1: R = conj(vk) c with |R| = inv_norm, so R^-1 = vk / (|vk| inv_norm)
      const int si = row & 1, p = sg[2 * (row >> 1)] & kMask;
      const PairScale& a = psg[p];
      zc vk = a.vk;
      if (si && a.k >= n2) vk = -vk;
      zc rf = vk * (1.0 / (absz(vk) * a.inv_norm));
      if (si) {
        const zc gr = gg[row];
        rf = rf * zexp(mk(-gr.im * k * dzg, gr.re * k * dzg));  // 1 / exp(-i k gamma dz)
      }
      s_rf[tid] = rf;
    }
  } else if (tid < TB_R + TB_C) {
    const int col = c0 + tid - TB_R;
    if (col < n4) {
      // column factor R_{g-1} Dc_{g-1}
      const int tj = col & 1, q = sp[2 * (col >> 1)] & kMask;
      zc cf = mode_rot(psp[q], tj != 0, n2);
      if (!tj) {
        const zc gc = gp[col];
        cf = cf * zexp(mk(-gc.im * k * dzp, gc.re * k * dzp));  // exp(i k gamma dz)
      }
      s_cf[tid - TB_R] = cf;
    }
  }
  __syncthreads();
  const int row = r0 + (tid & (TB_R - 1));
  if (row >= n4) return;
  const int i = row >> 1, si = row & 1;
  const int p = sg[2 * i] & kMask;
  const zc rf = s_rf[tid & (TB_R - 1)];
  const zc gi = gg[2 * i];
  for (int cc = tid / TB_R; cc < TB_C; cc += 256 / TB_R) {
    const int col = c0 + cc;
    if (col >= n4) break;
    const int j = col >> 1, tj = col & 1;
    const int q = sp[2 * j] & kMask;
    const zc x = X[(long)q * n2 + p];
    const zc y = Y[(long)q * n2 + p] * zdiv(gi, gp[2 * j]);
    zc v = (si == tj) ? x + y : x - y;
    v = 0.5 * v;
    T[(long)col * n4 + row] = rf * v * s_cf[cc];
  }
}

// beta_0 = Gd_0^-1 r in the block form: with r = [u; w] (2MF each), a1 = Vr^-1 u and a2 = Vr^-1 Q^-1 w already
// solved, beta_0(2i+s) = Dd^-1 R^-1 (2i+s) . 1/2 (a1(p_i) +- gam(2i) a2(p_i)) (+ for the down member s = 0).
// Instance blockIdx.y: layer first[i]; a1, a2 (2MF each) and beta_0 (4MF) from the lists.
__global__ void k_beta0_block(InstArgs ia, const int* __restrict__ first, MatSet A1, MatSet A2, MatSet B0,
                              const zc* __restrict__ gam, const int* __restrict__ src,
                              const PairScale* __restrict__ ps, const double* __restrict__ dz) {
  const int ii = blockIdx.y, g = first[ii];
  const int n4 = 4 * ia.MF, n2 = n4 / 2;
  const double k = ia.k[ia.linst[g]];
  const zc* a1 = A1.at(ii);
  const zc* a2 = A2.at(ii);
  zc* b0 = B0.at(ii);
  const zc* gg = gam + (long)g * n4;
  const int* sg = src + (long)g * n4;
  const PairScale* psg = ps + (long)g * n2;
  const double dzg = dz[g];
  constexpr int kMask = (1 << 30) - 1;
  for (int row = blockIdx.x * blockDim.x + threadIdx.x; row < n4; row += gridDim.x * blockDim.x) {
    const int i = row >> 1, si = row & 1;
    const int p = sg[2 * i] & kMask;
    const zc y = gg[2 * i] * a2[p];
    zc v = si ? a1[p] - y : a1[p] + y;
    v = 0.5 * v;
    const PairScale& a = psg[p];
    zc vk = a.vk;
    if (si && a.k >= n2) vk = -vk;
    zc rf = vk * (1.0 / (absz(vk) * a.inv_norm));
    if (si) {
      const zc gr = gg[row];
      rf = rf * zexp(mk(-gr.im * k * dzg, gr.re * k * dzg));
    }
    b0[row] = rf * v;
  }
}

// count[b] = number of non-finite entries of x[b] (n entries per instance); one workgroup per instance
__global__ __launch_bounds__(256) void k_count_nonfinite(int n, const zc* __restrict__ x, int* __restrict__ count) {
  __shared__ int s_c;
  if (threadIdx.x == 0) s_c = 0;
  __syncthreads();
  const zc* xb = x + (long)blockIdx.x * n;
  int c = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) c += (isfinite(xb[i].re) && isfinite(xb[i].im)) ? 0 : 1;
  if (c) atomicAdd(&s_c, c);
  __syncthreads();
  if (threadIdx.x == 0) count[blockIdx.x] = s_c;
}

__global__ void k_fill(long n, zc* __restrict__ a) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) a[i] = mk(0.0);
}

__global__ void k_copy(long n, const zc* __restrict__ a, zc* __restrict__ b) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) b[i] = a[i];
}

// one matrix of n elements per member: dst[b] = src[b]
__global__ void k_copy_set(long n, MatSet src, MatSet dst) {
  const zc* __restrict__ a = src.at(blockIdx.y);
  zc* __restrict__ b = dst.at(blockIdx.y);
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

__device__ inline BoundaryArgs boundary_args(const InstArgs& ia, int i) {
  const zc* t = ia.tabs + (long)i * 7 * ia.MF;
  const int MF = ia.MF;
  return BoundaryArgs{MF, ia.k[i], ia.z0[i], t, t + MF, t + 2 * MF, t + 3 * MF, t + 4 * MF};
}

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

// F[i] = f_which of instance i (4MF x 3MF each)
__global__ void k_build_f(InstArgs ia, int which, zc* __restrict__ F) {
  const int i = blockIdx.y;
  const BoundaryArgs a = boundary_args(ia, i);
  const int n4 = 4 * a.MF, n3 = 3 * a.MF;
  F += (long)i * n4 * n3;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)n4 * n3; e += (long)gridDim.x * blockDim.x)
    F[e] = f_entry(a, which, (int)(e % n4), (int)(e / n4));
}

// ---- the boundary system by elimination of the reflected waves ----
// A x = b (ScatterVolume.cpp:398-456) has, for every order m, five equations in which the reflected
// amplitudes R_m = (Rx, Ry, Rz)_m appear: the Gauss row (Sx, Sy, -Sz0) and rows m of the four f2 blocks
// (identity, identity, (0, Sz0, Sy), (-Sz0, 0, -Sx)), each of the last four also holding row m of one
// block of M = Gd0 S^-1 f3 against the transmitted amplitudes T.  A unitary Q_m^H (three Householder
// reflectors of the 5 x 3 coefficient block C_m) turns them into three rows that give R_m from T and two
// rows W_m (the left null space of C_m) that hold T alone.  Those 2 MF rows and the MF Gauss rows of T
// form a 3MF x 3MF system for T; R_m follows as G_m z_m with G_m = U_m^-1 (Q_m^H)(0:3, :) and z_m the
// five right-hand sides less the T terms.  In exact arithmetic this is the solution of A x = b; the
// elimination is orthogonal and C_m has full column rank (Sx^2 + Sy^2 + Sz0^2 = n0^2), so it is stable.
// Per order and instance: W_m (2 x 5) then G_m (3 x 5), row by row, 25 complex.
constexpr int kBndW = 25;

__global__ void k_bnd_qr(InstArgs ia, zc* __restrict__ Wq) {
  const int i = blockIdx.y;
  const BoundaryArgs a = boundary_args(ia, i);
  const int MF = a.MF;
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= MF) return;
  const zc sx = a.Sx[m], sy = a.Sy[m], s0 = a.Sz0[m], z = mk(0.0), one = mk(1.0);
  // [C_m | I_5]: coefficients of (Rx, Ry, Rz) in the Gauss row and the four f2 rows
  zc g[5][8] = {{sx, sy, -s0, one, z, z, z, z},
                {one, z, z, z, one, z, z, z},
                {z, one, z, z, z, one, z, z},
                {z, s0, sy, z, z, z, one, z},
                {-s0, z, -sx, z, z, z, z, one}};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    // zlarfg on column k, rows k..4; H^H = I - conj(tau) v v^H applied to every column
    double xn2 = 0.0;
#pragma unroll
    for (int r = k + 1; r < 5; ++r) xn2 += norm2(g[r][k]);
    const zc alpha = g[k][k];
    if (xn2 == 0.0 && alpha.im == 0.0) continue;
    const double beta = -copysign(sqrt(norm2(alpha) + xn2), alpha.re);
    const zc tau = mk((beta - alpha.re) / beta, -alpha.im / beta);
    const zc sc = zdiv(one, alpha - mk(beta));
    zc v[5];
#pragma unroll
    for (int r = 0; r < 5; ++r) v[r] = (r < k) ? z : (r == k ? one : g[r][k] * sc);
    const zc tc = conj(tau);
#pragma unroll
    for (int c = k; c < 8; ++c) {
      zc d = z;
#pragma unroll
      for (int r = k; r < 5; ++r) d = d + conj(v[r]) * g[r][c];
      d = tc * d;
#pragma unroll
      for (int r = k; r < 5; ++r) g[r][c] = g[r][c] - v[r] * d;
    }
    g[k][k] = mk(beta);
#pragma unroll
    for (int r = k + 1; r < 5; ++r) g[r][k] = z;
  }
  zc* w = Wq + ((long)i * MF + m) * kBndW;
#pragma unroll
  for (int r = 0; r < 2; ++r)
#pragma unroll
    for (int c = 0; c < 5; ++c) w[r * 5 + c] = g[3 + r][3 + c];
  // G = U^-1 (Q^H)(0:3, :) by back substitution on the upper triangular U = g(0:3, 0:3)
#pragma unroll
  for (int c = 0; c < 5; ++c) {
    zc x2 = zdiv(g[2][3 + c], g[2][2]);
    zc x1 = zdiv(g[1][3 + c] - g[1][2] * x2, g[1][1]);
    zc x0 = zdiv(g[0][3 + c] - g[0][1] * x1 - g[0][2] * x2, g[0][0]);
    w[10 + c] = x0;
    w[15 + c] = x1;
    w[20 + c] = x2;
  }
}

// The reduced 3MF x 3MF matrix per instance (column-major, leading dimension n3, instances sA apart):
// rows 0..MF-1 the Gauss rows of T (ScatterVolume.cpp:398-413), row MF + a MF + m the null-space row a
// of order m against the rows of M = Gd0 S^-1 f3 (Abr, 4MF x 3MF column-major).
__global__ void k_bnd_matrix(InstArgs ia, const zc* __restrict__ Abr, const zc* __restrict__ Wq, zc* __restrict__ Ar,
                             long sA) {
  const int i = blockIdx.y;
  const BoundaryArgs a = boundary_args(ia, i);
  const int MF = a.MF, n4 = 4 * MF, n3 = 3 * MF;
  Abr += (long)i * n4 * n3;
  Ar += (long)i * sA;
  Wq += (long)i * MF * kBndW;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)n3 * n3; e += (long)gridDim.x * blockDim.x) {
    const int r = (int)(e % n3), c = (int)(e / n3);
    zc v = mk(0.0);
    if (r < MF) {
      if (c == r) v = a.Sx[r];
      else if (c == MF + r) v = a.Sy[r];
      else if (c == 2 * MF + r) v = a.Sz1[r];
    } else {
      const int rr = r - MF, row = rr / MF, m = rr % MF;
      const zc* w = Wq + (long)m * kBndW + row * 5;
      const zc* col = Abr + (long)c * n4 + m;
#pragma unroll
      for (int j = 1; j < 5; ++j) v = v + w[j] * col[(long)(j - 1) * MF];
    }
    Ar[(long)c * n3 + r] = v;
  }
}

// b = [0; -f1 EF] (ScatterVolume.cpp:454) into x, one instance per blockIdx.y.
__global__ void k_bnd_b(InstArgs ia, zc* __restrict__ b) {
  const int i = blockIdx.y;
  const BoundaryArgs a = boundary_args(ia, i);
  const int MF = a.MF, n6 = 6 * MF, n3 = 3 * MF;
  b += (long)i * n6;
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n6; r += gridDim.x * blockDim.x) {
    zc v = mk(0.0);
    if (r >= 2 * MF) {
      const int rr = r - 2 * MF;
      // f1 is zero outside the columns c = rr (mod MF): the three that remain, in increasing c
      for (int c = rr % MF; c < n3; c += MF) {
        const zc f = f_entry(a, 1, rr, c);
        if (f.re != 0.0 || f.im != 0.0) v = v + f * a.EF[c];
      }
      v = -v;
    }
    b[r] = v;
  }
}

// Right-hand sides of the reduced system from full right-hand sides b (rows 2MF..6MF hold -f1 EF, the Gauss
// rows are zero): t(m) = 0 for the Gauss rows of T, t(MF + a MF + m) = sum_j W_m(a, j) b_m(j).  Instance
// blockIdx.z, right-hand side blockIdx.y: b at xs + (z nrhs + y) 6MF, t at t + (z nrhs + y) 3MF.
__global__ void k_bnd_rhs(InstArgs ia, const zc* __restrict__ Wq, const zc* __restrict__ xs, int nrhs,
                          zc* __restrict__ t) {
  const int i = blockIdx.z, k = blockIdx.y;
  const int MF = ia.MF, n6 = 6 * MF, n3 = 3 * MF;
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= MF) return;
  const zc* b = xs + ((long)i * nrhs + k) * n6;
  zc* tt = t + ((long)i * nrhs + k) * n3;
  const zc* w = Wq + ((long)i * MF + m) * kBndW;
  tt[m] = mk(0.0);
#pragma unroll
  for (int row = 0; row < 2; ++row) {
    zc v = mk(0.0);
#pragma unroll
    for (int j = 1; j < 5; ++j) v = v + w[row * 5 + j] * b[(j + 1) * MF + m];
    tt[MF + row * MF + m] = v;
  }
}

// x = [R; T] from the reduced solution T (t) and y = M T (Y, 4MF per right-hand side): R_m = G_m z_m with
// z_m = (0, b_m(j) - y_m(j)).  Each thread reads its order's right-hand side rows before it overwrites them.
__global__ void k_bnd_back(InstArgs ia, const zc* __restrict__ Wq, const zc* __restrict__ Y, const zc* __restrict__ t,
                           zc* __restrict__ xs, int nrhs) {
  const int i = blockIdx.z, k = blockIdx.y;
  const int MF = ia.MF, n6 = 6 * MF, n4 = 4 * MF, n3 = 3 * MF;
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= MF) return;
  zc* x = xs + ((long)i * nrhs + k) * n6;
  const zc* y = Y + ((long)i * nrhs + k) * n4;
  const zc* tt = t + ((long)i * nrhs + k) * n3;
  const zc* w = Wq + ((long)i * MF + m) * kBndW + 10;
  zc zv[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) zv[j] = x[(j + 2) * MF + m] - y[j * MF + m];
  zc R[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    zc v = mk(0.0);
#pragma unroll
    for (int j = 0; j < 4; ++j) v = v + w[c * 5 + 1 + j] * zv[j];
    R[c] = v;
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    x[c * MF + m] = R[c];
    x[(3 + c) * MF + m] = tt[c * MF + m];
  }
}

// rhs of beta_0: f1 EF + f2 R  (ScatterVolume.cpp:767-772), into the first layer's beta of instance i
__global__ void k_beta0_rhs(InstArgs ia, const zc* __restrict__ x, MatSet out) {
  const int i = blockIdx.y;
  const BoundaryArgs a = boundary_args(ia, i);
  const int n4 = 4 * a.MF, n3 = 3 * a.MF, n6 = 6 * a.MF;
  x += (long)i * n6;
  zc* o = out.at(i);
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n4; r += gridDim.x * blockDim.x) {
    zc v = mk(0.0);
    // f1 and f2 are zero outside the columns c = r (mod MF): the three that remain, in increasing c
    for (int c = r % a.MF; c < n3; c += a.MF) {
      const zc f1 = f_entry(a, 1, r, c), f2 = f_entry(a, 2, r, c);
      if (f1.re != 0.0 || f1.im != 0.0) v = v + f1 * a.EF[c];
      if (f2.re != 0.0 || f2.im != 0.0) v = v + f2 * x[c];
    }
    o[r] = v;
  }
}

// Q_check / Q_hat rows 0..2MF and the Ez right-hand side Y (ScatterVolume.cpp:823-881):
// Qch = [V_even diag(beta_even) | V_odd diag(beta_odd)] rows 0..2MF (3MF x 4MF per layer);
// Y(p', c) = -(1/k) (up_p' Hy(p', c) - wq_p' Hx(p', c)), MF x 4MF per layer.
__global__ void k_q_rows(InstArgs ia, const zc* __restrict__ Gd, const zc* __restrict__ Gc,
                         const zc* __restrict__ beta, zc* __restrict__ Qch, zc* __restrict__ Y) {
  const int MF = ia.MF;
  const int n4 = 4 * MF, n3 = 3 * MF, n2 = 2 * MF;
  const int l = blockIdx.y;
  const int inst = ia.linst[l];
  const double k = ia.k[inst];
  const double* upz = ia.upz + (long)inst * MF;
  const double* wqz = ia.wqz + (long)inst * MF;
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
  int kind, layer, inst;
  double zb;
};

__global__ void k_records(InstArgs ia, const zc* __restrict__ x, const zc* __restrict__ gam,
                          const zc* __restrict__ Qch, const Seg* __restrict__ segs, zc* __restrict__ rec) {
  const Seg sg = segs[blockIdx.y];
  const int MF = ia.MF, n2 = 2 * MF, n3 = 3 * MF, n4 = 4 * MF, n6 = 6 * MF;
  const double k = ia.k[sg.inst];
  const double n0 = ia.nio[2 * sg.inst], n1 = ia.nio[2 * sg.inst + 1];
  const zc* t = ia.tabs + (long)sg.inst * 7 * MF;
  const zc *Sx = t, *Sy = t + MF, *Sz0 = t + 2 * MF, *Sz1 = t + 3 * MF, *EF = t + 4 * MF;
  x += (long)sg.inst * n6;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < sg.cnt; i += (long)gridDim.x * blockDim.x) {
    zc E[3], K[3];
    int p;
    if (sg.kind <= 2) p = (int)i;
    else p = (int)(i / n2);
    K[0] = Sx[p] * k;
    K[1] = Sy[p] * k;
    bool wind = true;
    if (sg.kind == 0) {
      K[2] = Sz0[p] * k;
      for (int c = 0; c < 3; ++c) E[c] = EF[c * MF + p];
      wind = false;
    } else if (sg.kind == 1) {
      K[2] = -(Sz0[p] * k) * n0;
      for (int c = 0; c < 3; ++c) E[c] = x[c * MF + p];
    } else if (sg.kind == 2) {
      K[2] = (Sz1[p] * k) * n1;
      for (int c = 0; c < 3; ++c) E[c] = x[n3 + c * MF + p];
    } else {
      const int jj = (int)(i % n2);
      const bool up = (sg.kind == 4);
      const zc g = gam[(long)sg.layer * n4 + 2 * jj + (up ? 1 : 0)];
      K[2] = g * k;
      const zc* Q = Qch + (long)sg.layer * n3 * n4 + (long)(jj + (up ? n2 : 0)) * n3;
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
// with A the column-major layer matrix (FourierWave.h:152-195, 288, 341-343).  The kept frequencies
// depend only on the grid and the orders, so one pair of DFT matrices serves every layer.
void fourier_stage(const Setup& st, int L, const zc* d_sample, zc* F, hipStream_t s) {
  const int s1 = st.s1, s2 = st.s2, M0 = st.M0, M1 = st.M1, MF = st.MF;
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
  for (int b0 = 0; b0 < 2 * L; b0 += 65535) {
    const int nb = std::min(65535, 2 * L - b0);
    zgemm_batched(OpN, OpN, s1, M0, s2, mk(1.0), sq.p + b0 * npix, s1, npix, d_w2.p, s2, 0, mk(0.0),
                  T1.p + (long)b0 * s1 * M0, s1, (long)s1 * M0, nb, s);
    zgemm_batched(OpN, OpN, M1, M0, s1, mk(1.0), d_w1.p, M1, 0, T1.p + (long)b0 * s1 * M0, s1, (long)s1 * M0,
                  mk(0.0), F + (long)b0 * MF, M1, MF, nb, s);
  }
  // keep the temporaries alive until the kernels that read them have run
  CWX_HIP(hipStreamSynchronize(s));
}

// Host tables of a set of instances that share the grid and the orders.
struct InstSet {
  int count = 0, Ltot = 0, maxL = 0;
  std::vector<Setup> st;
  std::vector<int> L, first;  // layers and first global layer of every instance
  DevBuf<int> linst;
  DevBuf<double> k, up, wq, upz, wqz, z0, nio, dz;
  DevBuf<zc> tabs;
  InstArgs args() const {
    const Setup& s0 = st[0];
    return InstArgs{s0.M0, s0.M1, s0.MF, linst.p, k.p, up.p, wq.p, upz.p, wqz.p, z0.p, nio.p, tabs.p};
  }
};

void make_instances(const std::vector<VolumeOptions>& opts, const std::vector<int>& Ls, int s1, int s2,
                    InstSet& is, hipStream_t s) {
  const int count = (int)opts.size();
  CWX_CHECK(count >= 1 && (int)Ls.size() == count, "instance set is empty");
  is.count = count;
  is.st.clear();
  is.L = Ls;
  is.first.assign(count, 0);
  is.Ltot = 0;
  is.maxL = 0;
  for (int i = 0; i < count; ++i) {
    is.st.push_back(make_setup(opts[i], Ls[i], s1, s2));
    is.first[i] = is.Ltot;
    is.Ltot += Ls[i];
    is.maxL = std::max(is.maxL, Ls[i]);
    CWX_CHECK(is.st[i].M0 == is.st[0].M0 && is.st[i].M1 == is.st[0].M1,
              "instances solved together must have the same Fourier orders (instance " + std::to_string(i) +
                  " has " + std::to_string(is.st[i].M0) + "x" + std::to_string(is.st[i].M1) + ")");
    CWX_CHECK(opts[i].eig_mode == opts[0].eig_mode, "instances solved together must use one eigen mode");
  }
  CWX_CHECK(is.Ltot <= 65535, "too many layers in one set");
  const int M0 = is.st[0].M0, M1 = is.st[0].M1, MF = M0 * M1;
  std::vector<int> linst(is.Ltot);
  std::vector<double> k(count), up, wq, upz, wqz, z0(count), nio(2 * count), dz(is.Ltot);
  std::vector<zc> tabs;
  tabs.reserve((size_t)7 * MF * count);
  for (int i = 0; i < count; ++i) {
    const Setup& st = is.st[i];
    const VolumeOptions& o = opts[i];
    const int L = Ls[i];
    for (int l = 0; l < L; ++l) {
      linst[is.first[i] + l] = i;
      dz[is.first[i] + l] = (L == 1) ? (st.z[L] - st.z[0]) : (st.z[l + 1] - st.z[l]);
    }
    k[i] = st.k;
    up.insert(up.end(), st.up.begin(), st.up.end());
    wq.insert(wq.end(), st.wq.begin(), st.wq.end());
    // Pz uses M_PI in UP/WQ (ScatterVolume.cpp:813-814), indexed by the column p' = qj*M0 + pj
    for (int q = 0; q < M1; ++q)
      for (int pp = 0; pp < M0; ++pp) {
        upz.push_back(2.0 * (double)(pp - M0 / 2) * kPi / o.size[0] + st.dir_n[0] * st.k);
        wqz.push_back(2.0 * (double)(q - M1 / 2) * kPi / o.size[1] + st.dir_n[1] * st.k);
      }
    z0[i] = st.z[0];
    nio[2 * i] = o.n[0];
    nio[2 * i + 1] = o.n[1];
    for (auto* v : {&st.Sx, &st.Sy, &st.Sz0, &st.Sz1, &st.EF}) tabs.insert(tabs.end(), v->begin(), v->end());
  }
  upload(is.linst, linst, s);
  upload(is.k, k, s);
  upload(is.up, up, s);
  upload(is.wq, wq, s);
  upload(is.upz, upz, s);
  upload(is.wqz, wqz, s);
  upload(is.z0, z0, s);
  upload(is.nio, nio, s);
  upload(is.dz, dz, s);
  upload(is.tabs, tabs, s);
}

// Pointer lists for the set-based linear algebra, gathered on the host and uploaded once.
struct PtrLists {
  std::vector<void*> h;
  DevBuf<void*> d;
  size_t add(const std::vector<void*>& v) {
    const size_t o = h.size();
    h.insert(h.end(), v.begin(), v.end());
    return o;
  }
  void commit(hipStream_t s) { upload(d, h, s); }
  MatSet mats(size_t o) const { return listed<zc>(reinterpret_cast<zc* const*>(d.p + o)); }
  IntSet ints(size_t o) const { return listed<int>(reinterpret_cast<int* const*>(d.p + o)); }
};

}  // namespace

// Phi of every layer (blocks = 0: full 4MF x 4MF) or its P and Q blocks (blocks = 1: 2 x L x (2MF)^2,
// all P first) into device memory.  Returns the setup.
Setup build_phi_dev(const VolumeOptions& o, const zc* d_sample, int L, int s1, int s2, int blocks, zc* out,
                    hipStream_t s) {
  InstSet is;
  make_instances({o}, {L}, s1, s2, is, s);
  const Setup& st = is.st[0];
  const int MF = st.MF;
  DevBuf<zc> F((size_t)2 * L * MF);
  fourier_stage(st, L, d_sample, F.p, s);
  const long n4 = 4L * MF, n2 = 2L * MF;
  if (blocks == 0)
    hipLaunchKernelGGL(k_phi_full, dim3(grid_for(n4 * n4), L), dim3(256), 0, s, is.args(), F.p, F.p + (long)L * MF,
                       out, nullptr);
  else
    hipLaunchKernelGGL(k_phi_blocks, dim3(grid_for(n2 * n2), L), dim3(256), 0, s, is.args(), F.p,
                       F.p + (long)L * MF, out, out + L * n2 * n2);
  CWX_HIP(hipGetLastError());
  CWX_HIP(hipStreamSynchronize(s));
  return st;
}

// ------------------------------------------------------------------------------------------------
// the pipeline, for a set of instances that share the grid and the Fourier orders
// ------------------------------------------------------------------------------------------------
namespace {

struct SetResult {
  DevBuf<zc>* x = nullptr;      // count x 6MF (caller-owned buffer)
  std::vector<int> info;        // per instance
  std::vector<std::string> msg;
  Solution* full = nullptr;     // count == 1: keep every array and build the .cw records here
  int K = 0;                    // count == 1: further incident fields solved with the same factors
  const zc* EFs = nullptr;      // K x 3MF incident Fourier coefficients
  zc* xs = nullptr;             // K x 6MF boundary solutions
};

// The arrays a full solution keeps besides gam, Gd, Gc, beta and x (already in sol): the n^-2
// coefficients, the order tables, the Q_check / Q_hat rows and the .cw records
// (ScatterVolume.cpp:802-987).  is holds the one instance; NI its L x MF n^-2 coefficients.
void finish_solution(Solution& sol, const VolumeOptions& o, const InstSet& is, const zc* NI, hipStream_t s) {
  const InstArgs ia = is.args();
  const int Ltot = is.Ltot;
  const int M0 = is.st[0].M0, M1 = is.st[0].M1, MF = M0 * M1;
  const int n2 = 2 * MF, n3 = 3 * MF, n4 = 4 * MF;
  const long nn = (long)n4 * n4;
  DevBuf<zc>& Gd = sol.Gd;
  DevBuf<zc>& Gc = sol.Gc;
  DevBuf<zc>& beta = sol.beta;
  DevBuf<zc>& gam = sol.gam;
  DevBuf<zc>& x = sol.x;
  sol.opt = o;
  sol.st = is.st[0];
  sol.n4 = n4;
  sol.NIF.alloc((size_t)Ltot * MF);
  CWX_HIP(hipMemcpyAsync(sol.NIF.p, NI, sizeof(zc) * Ltot * MF, hipMemcpyDeviceToDevice, s));
  sol.tabs.alloc((size_t)7 * MF);
  CWX_HIP(hipMemcpyAsync(sol.tabs.p, is.tabs.p, sizeof(zc) * 7 * MF, hipMemcpyDeviceToDevice, s));
  if (!sol.opt.external) {
    sol.Qch.alloc((size_t)Ltot * n3 * n4);
    DevBuf<zc> Yz((size_t)Ltot * MF * n4), Cn((size_t)Ltot * MF * MF);
    hipLaunchKernelGGL(k_q_rows, dim3(grid_for(nn), Ltot), dim3(256), 0, s, ia, Gd.p, Gc.p, beta.p, sol.Qch.p,
                       Yz.p);
    if (sol.opt.npz) {
      hipLaunchKernelGGL(k_zero_rows, dim3(grid_for((long)MF * n4), Ltot), dim3(256), 0, s, n2, MF, n3, n4,
                         sol.Qch.p, (long)n3 * n4);
    } else {
      hipLaunchKernelGGL(k_cn, dim3(grid_for((long)MF * MF), Ltot), dim3(256), 0, s, M0, M1, sol.NIF.p, Cn.p);
      zgemm_batched(OpN, OpN, MF, n4, MF, mk(1.0), Cn.p, MF, (long)MF * MF, Yz.p, MF, (long)MF * n4, mk(0.0),
                    sol.Qch.p + n2, n3, (long)n3 * n4, Ltot, s);
    }
    CWX_HIP(hipStreamSynchronize(s));  // Yz and Cn are released on return
  }

  // ---- 7. .cw records ----
  const Setup& st = sol.st;
  const int L = Ltot, ZL = L + 1;
  std::vector<Seg> segs;
  long off = 0;
  auto add = [&](long cnt, int kind, int layer, double zb) {
    segs.push_back(Seg{off, cnt, kind, layer, 0, zb});
    sol.seg_off.push_back(off);
    sol.seg_cnt.push_back(cnt);
    off += cnt;
  };
  sol.seg_off.clear();
  sol.seg_cnt.clear();
  const long inner = sol.opt.external ? 0 : (long)n2 * MF;
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
  hipLaunchKernelGGL(k_records, dim3(grid_for(maxcnt), (unsigned)segs.size()), dim3(256), 0, s, ia, x.p, gam.p,
                     sol.opt.external ? nullptr : sol.Qch.p, d_segs.p, sol.rec.p);
  CWX_HIP(hipGetLastError());
  CWX_HIP(hipStreamSynchronize(s));  // d_segs is released on return
  }

// Buffers of one set solve.  An ordinary call keeps them until its status read-back; a captured graph
// keeps them, with work areas of its own, for as long as it is replayed.
struct Level {
  size_t oA, oB, oC;
  int batch;
};

struct SetBufs {
  InstSet is;
  int count = 0, Ltot = 0, M0 = 0, M1 = 0, MF = 0, n2 = 0, n3 = 0, n4 = 0, n6 = 0, eig_mode = 0;
  long nn = 0;
  bool capture = false, keep_beta = false;
  Solution* full = nullptr;
  // Fourier stage
  DevBuf<zc> dft1, dft2, sq, T1, F;
  // eigen stage
  DevBuf<int> d_src, d_info, d_unsafe;
  DevBuf<zc> gam_local, W, V, QA;
  DevBuf<PairScale> pscale;
  DevBuf<zc> Vf;                       // reduced mode: layers redone through the full Phi
  DevBuf<int> d_fslot, d_flist, d_finfo;
  std::vector<int> flist;
  // chain, boundary system, beta
  DevBuf<zc> Gd_local, Gc_local, beta_local, Y, Abr, A, Wq;
  DevBuf<int> d_info_lu, infoS, infoA, d_nonfinite;
  DevBuf<zc>* gam = nullptr;
  DevBuf<zc>* Gd = nullptr;
  DevBuf<zc>* Gc = nullptr;
  DevBuf<zc>* beta = nullptr;
  DevBuf<zc>* x = nullptr;
  SetWork own;
  SetWork* wk = nullptr;
  PtrLists pl;
  std::vector<Level> levels, bsteps;
  size_t oTLU = 0, oTP = 0, oTB = 0, oCS = 0, oCD = 0, oGd0 = 0, oLU0 = 0, oPiv0 = 0, oB0 = 0;
  int nT = 0, ncS = 0;
  // the chain's block form (reduced mode): lists over the T items and over the instances' first layers
  size_t oKV = 0, oKPV = 0, oKQ = 0, oKPQ = 0, oKX = 0, oKZ = 0, oKVp = 0, oKQAp = 0, oKT = 0, oKDV = 0, oKDQ = 0;
  // the two independent solves of the chain (V_g X = Vr_{g-1}, Q_g Z = QA_{g-1}) as one batch: V's then Q's
  size_t oK2LU = 0, oK2P = 0, oK2D = 0, oK2B = 0;
  size_t oK0V = 0, oK0PV = 0, oK0Q = 0, oK0PQ = 0, oK0B = 0, oK0A1 = 0, oK0A2 = 0;
  size_t oK0VQ = 0, oK0PVQ = 0, oK0A12 = 0;  // the lists above, Vr then Q (a1 then a2)
  DevBuf<int> d_tg, d_first;
  bool used_block = false;
};

// Host work of a set solve: instance tables, DFT matrices, every buffer and the pointer lists, all on the
// device before the first kernel.  With `capture`, the launch sequence that follows may be recorded in a
// graph: the buffers (and work areas) are the plan's own and the launches need no host decision.
void set_prepare(Context& ctx, const std::vector<VolumeOptions>& opts, const std::vector<int>& Ls, int s1, int s2,
                 SetResult& res, SetBufs& b, bool capture) {
  hipStream_t s = ctx.stream;
  make_instances(opts, Ls, s1, s2, b.is, s);
  InstSet& is = b.is;
  b.capture = capture;
  b.count = is.count;
  b.Ltot = is.Ltot;
  const Setup& st0 = is.st[0];
  b.M0 = st0.M0;
  b.M1 = st0.M1;
  b.MF = st0.MF;
  b.n2 = 2 * b.MF;
  b.n3 = 3 * b.MF;
  b.n4 = 4 * b.MF;
  b.n6 = 6 * b.MF;
  b.nn = (long)b.n4 * b.n4;
  b.eig_mode = opts[0].eig_mode;
  b.full = res.full;
  const int count = b.count, Ltot = b.Ltot, MF = b.MF, n2 = b.n2, n3 = b.n3, n4 = b.n4, n6 = b.n6;
  const long nn = b.nn, mm = (long)n2 * n2;
  CWX_CHECK(n4 <= SORT_MAX, "4*M0*M1 exceeds the supported maximum of 4096");
  CWX_CHECK(res.full == nullptr || count == 1, "full solutions are kept for single instances only");
  CWX_CHECK(!capture || (res.full == nullptr && res.K == 0), "a captured solve keeps only the boundary solution");
  b.wk = capture ? &b.own : &ctx.work;
  SetWork& wk = *b.wk;

  // ---- Fourier stage: a pruned DFT F = w1 (M1 x s1) . A (s1 x s2) . w2 (s2 x M0) / (s1 s2) per layer,
  //      with A the column-major layer matrix (FourierWave.h:152-195, 288, 341-343) ----
  {
    const int M0 = b.M0, M1 = b.M1;
    std::vector<zc> w1((size_t)M1 * s1), w2((size_t)s2 * M0);
    for (int i = 0; i < s1; ++i)
      for (int r = 0; r < M1; ++r) {
        const long t = ((long)st0.rows[r] * i) % s1;
        const double ang = -2.0 * kPi * (double)t / (double)s1;
        w1[(size_t)i * M1 + r] = mk(std::cos(ang) / ((double)s1 * s2), std::sin(ang) / ((double)s1 * s2));
      }
    for (int c = 0; c < M0; ++c)
      for (int j = 0; j < s2; ++j) {
        const long t = ((long)st0.cols[c] * j) % s2;
        const double ang = -2.0 * kPi * (double)t / (double)s2;
        w2[(size_t)c * s2 + j] = mk(std::cos(ang), std::sin(ang));
      }
    upload(b.dft1, w1, s);
    upload(b.dft2, w2, s);
    const long npix = (long)s1 * s2;
    b.sq.ensure((size_t)2 * npix * Ltot);
    b.T1.ensure((size_t)2 * Ltot * s1 * M0);
    b.F.ensure((size_t)2 * Ltot * MF);
  }
  // ---- eigen stage ----
  b.d_src.ensure((size_t)Ltot * n4);
  b.gam = res.full ? &res.full->gam : &b.gam_local;
  b.gam->alloc((size_t)Ltot * n4);
  b.d_info.ensure(Ltot);
  b.d_unsafe.ensure(Ltot);
  wk.w0.ensure(std::max<size_t>((size_t)Ltot * nn, 1));
  wk.w1.ensure((size_t)Ltot * nn);
  if (b.eig_mode == 0) {
    b.W.ensure((size_t)Ltot * n4);
    b.V.ensure((size_t)Ltot * nn);
    wk.eig.ensure(n4, Ltot);
  } else {
    b.W.ensure((size_t)Ltot * n2);
    b.V.ensure((size_t)Ltot * mm);
    b.QA.ensure((size_t)Ltot * mm);
    b.pscale.ensure((size_t)Ltot * n2);
    wk.eig.ensure(n2, Ltot);
  }
  // ---- chain, boundary system, beta ----
  b.Gd = res.full ? &res.full->Gd : &b.Gd_local;
  b.Gc = res.full ? &res.full->Gc : &b.Gc_local;
  b.Gd->alloc((size_t)Ltot * nn);
  b.Gc->alloc((size_t)Ltot * nn);
  wk.ipiv.ensure((size_t)Ltot * n4 + (size_t)count * (n4 + n6));
  wk.w2.ensure((size_t)count * nn);
  {
    // the product tree's pool; before the tree, in the reduced mode, the inverses of the diagonal blocks of
    // Q and V and the solves' scratch; after it, those of S and the scratch of S^-1 f3
    size_t w3 = (size_t)std::max(1, Ltot) * nn;
    if (b.eig_mode == 1) w3 = std::max(w3, (size_t)Ltot * (2 * (size_t)trtri_diag_size(n2) + 2 * (size_t)mm));
    w3 = std::max(w3, (size_t)count * ((size_t)trtri_diag_size(n4) + (size_t)n4 * n3));
    wk.w3.ensure(w3);
  }
  b.d_info_lu.ensure(2 * (size_t)Ltot);
  b.infoS.ensure(count);
  b.infoA.ensure(count);
  b.d_nonfinite.ensure(count);
  b.Y.ensure((size_t)count * n4 * n3);
  b.Abr.ensure((size_t)count * n4 * n3);
  b.A.ensure((size_t)count * n6 * n6);
  b.Wq.ensure((size_t)count * MF * kBndW);
  b.x = res.x;
  b.x->alloc((size_t)count * n6);
  b.beta = res.full ? &res.full->beta : &b.beta_local;
  const bool keep_beta = res.full && !opts[0].external;
  b.keep_beta = keep_beta;
  if (keep_beta) b.beta->alloc((size_t)Ltot * n4);
  // pointer lists
  zc* LUd = wk.w0.p;
  zc* Tm = wk.w1.p;
  int* piv = wk.ipiv.p;
  zc* Sm = wk.w2.p;
  zc* pool = wk.w3.p;
  std::vector<void*> tLU, tP, tB;
  for (int i = 0; i < count; ++i)
    for (int l = 1; l < is.L[i]; ++l) {
      const long g = is.first[i] + l;
      tLU.push_back(LUd + g * nn);
      tP.push_back(piv + g * n4);
      tB.push_back(Tm + g * nn);
    }
  b.nT = (int)tB.size();
  b.oTLU = b.pl.add(tLU);
  b.oTP = b.pl.add(tP);
  b.oTB = b.pl.add(tB);
  // product tree: level items are pointers; pairs (X_{2m+1}, X_{2m}) -> X_{2m+1} X_{2m}
  long pool_used = 0;
  std::vector<std::vector<const zc*>> seq(count);
  for (int i = 0; i < count; ++i) {
    const int f = is.first[i], L = is.L[i];
    for (int l = 1; l < L; ++l) seq[i].push_back(Tm + (long)(f + l) * nn);
    seq[i].push_back(b.Gc->p + (long)(f + L - 1) * nn);
  }
  for (;;) {
    std::vector<void*> Av, Bv, Cv;
    for (int i = 0; i < count; ++i) {
      auto& q = seq[i];
      if (q.size() < 2) continue;
      std::vector<const zc*> next;
      for (size_t m = 0; m + 1 < q.size(); m += 2) {
        zc* out = (q.size() == 2) ? Sm + (long)i * nn : pool + (pool_used++) * nn;
        Av.push_back((void*)q[m + 1]);
        Bv.push_back((void*)q[m]);
        Cv.push_back(out);
        next.push_back(out);
      }
      if (q.size() % 2) next.push_back(q.back());
      q = next;
    }
    if (Av.empty()) break;
    b.levels.push_back(Level{b.pl.add(Av), b.pl.add(Bv), b.pl.add(Cv), (int)Av.size()});
  }
  CWX_CHECK(pool_used <= std::max(1, Ltot), "product tree exceeded its pool");
  std::vector<void*> cS, cD;
  for (int i = 0; i < count; ++i)
    if (is.L[i] == 1) {
      cS.push_back(b.Gc->p + (long)is.first[i] * nn);
      cD.push_back(Sm + (long)i * nn);
    }
  b.ncS = (int)cS.size();
  b.oCS = b.pl.add(cS);
  b.oCD = b.pl.add(cD);
  std::vector<void*> gd0, lu0, piv0, b0;
  for (int i = 0; i < count; ++i) {
    const long f = is.first[i];
    gd0.push_back(b.Gd->p + f * nn);
    lu0.push_back(LUd + f * nn);
    piv0.push_back(piv + f * n4);
    b0.push_back(keep_beta ? b.beta->p + f * n4 : nullptr);
  }
  b.oGd0 = b.pl.add(gd0);
  b.oLU0 = b.pl.add(lu0);
  b.oPiv0 = b.pl.add(piv0);
  b.oB0 = b.pl.add(b0);
  if (keep_beta) {
    // beta_j = T_j beta_{j-1} for every instance with more than j layers
    for (int j = 1; j < is.maxL; ++j) {
      std::vector<void*> Av, Bv, Cv;
      for (int i = 0; i < count; ++i)
        if (is.L[i] > j) {
          const long g = is.first[i] + j;
          Av.push_back(Tm + g * nn);
          Bv.push_back(b.beta->p + (g - 1) * n4);
          Cv.push_back(b.beta->p + g * n4);
        }
      b.bsteps.push_back(Level{b.pl.add(Av), b.pl.add(Bv), b.pl.add(Cv), (int)Av.size()});
    }
  }
  if (b.eig_mode == 1) {
    // the block form: in wk.w0, X at [0, nT mm) (P's place), Q at [Ltot mm, 2 Ltot mm) factored in place, the
    // factors of V at [2 Ltot mm, 3 Ltot mm), Y at [3 Ltot mm, ...); pivots of Q then V at the start of ipiv
    // (Q and V are factored as one batch of 2 Ltot matrices)
    const long mm = (long)n2 * n2;
    zc* w0 = wk.w0.p;
    zc* Qb = w0 + (long)Ltot * mm;
    zc* LUV = w0 + 2L * Ltot * mm;
    zc* Zb = w0 + 3L * Ltot * mm;
    int* pivQ = piv;
    int* pivV = piv + (long)Ltot * n2;
    // inverses of the diagonal blocks of the factors of Q then V (for the solves by products), in wk.w3
    const long dsz = trtri_diag_size(n2);
    zc* Dinv = wk.w3.p;
    std::vector<void*> kV, kPV, kQ, kPQ, kX, kZ, kVp, kQAp, kT, kDV, kDQ;
    std::vector<int> tg;
    for (int i = 0; i < count; ++i)
      for (int l = 1; l < is.L[i]; ++l) {
        const long g = is.first[i] + l, t = (long)tg.size();
        kDQ.push_back(Dinv + g * dsz);
        kDV.push_back(Dinv + (Ltot + g) * dsz);
        kV.push_back(LUV + g * mm);
        kPV.push_back(pivV + g * n2);
        kQ.push_back(Qb + g * mm);
        kPQ.push_back(pivQ + g * n2);
        kX.push_back(w0 + t * mm);
        kZ.push_back(Zb + t * mm);
        kVp.push_back(b.V.p + (g - 1) * mm);
        kQAp.push_back(b.QA.p + (g - 1) * mm);
        kT.push_back(Tm + g * nn);
        tg.push_back((int)g);
      }
    b.oKV = b.pl.add(kV);
    b.oKPV = b.pl.add(kPV);
    b.oKQ = b.pl.add(kQ);
    b.oKPQ = b.pl.add(kPQ);
    b.oKX = b.pl.add(kX);
    b.oKZ = b.pl.add(kZ);
    b.oKVp = b.pl.add(kVp);
    b.oKQAp = b.pl.add(kQAp);
    b.oKT = b.pl.add(kT);
    b.oKDV = b.pl.add(kDV);
    b.oKDQ = b.pl.add(kDQ);
    auto cat = [](std::vector<void*> a, const std::vector<void*>& c) {
      a.insert(a.end(), c.begin(), c.end());
      return a;
    };
    b.oK2LU = b.pl.add(cat(kV, kQ));
    b.oK2P = b.pl.add(cat(kPV, kPQ));
    b.oK2D = b.pl.add(cat(kDV, kDQ));
    b.oK2B = b.pl.add(cat(kX, kZ));
    upload(b.d_tg, tg, s);
    std::vector<void*> k0V, k0PV, k0Q, k0PQ, k0B, k0A1, k0A2;
    std::vector<int> first;
    for (int i = 0; i < count; ++i) {
      const long f = is.first[i];
      k0V.push_back(LUV + f * mm);
      k0PV.push_back(pivV + f * n2);
      k0Q.push_back(Qb + f * mm);
      k0PQ.push_back(pivQ + f * n2);
      k0B.push_back(keep_beta ? b.beta->p + f * n4 : nullptr);
      k0A1.push_back(w0 + (long)i * n4);       // beta_0's right-hand side, then a1 and a2, in X's place
      k0A2.push_back(w0 + (long)i * n4 + n2);
      first.push_back((int)f);
    }
    b.oK0V = b.pl.add(k0V);
    b.oK0PV = b.pl.add(k0PV);
    b.oK0Q = b.pl.add(k0Q);
    b.oK0PQ = b.pl.add(k0PQ);
    b.oK0B = b.pl.add(k0B);
    b.oK0A1 = b.pl.add(k0A1);
    b.oK0A2 = b.pl.add(k0A2);
    // the solves a1 = Vr^-1 u and w <- Q^-1 w are independent: one batch of 2 count
    b.oK0VQ = b.pl.add(cat(k0V, k0Q));
    b.oK0PVQ = b.pl.add(cat(k0PV, k0PQ));
    b.oK0A12 = b.pl.add(cat(k0A1, k0A2));
    upload(b.d_first, first, s);
  }
  b.pl.commit(s);
  CWX_HIP(hipGetLastError());
}

// The kernels of a set solve, in stream order, on the buffers set_prepare made.  Without `capture` the
// reduced mode may redo nearly defective layers through the full Phi (a host decision after the sort),
// and a full solution gets its Q rows and .cw records; with `capture` nothing but launches is issued.
void set_launch(Context& ctx, SetBufs& b, const zc* d_samples, int s1, int s2, SetResult& res) {
  hipStream_t s = ctx.stream;
  const InstSet& is = b.is;
  const InstArgs ia = is.args();
  SetWork& wk = *b.wk;
  const bool capture = b.capture;
  const int count = b.count, Ltot = b.Ltot, MF = b.MF, n2 = b.n2, n3 = b.n3, n4 = b.n4, n6 = b.n6;
  const int M0 = b.M0;
  const long nn = b.nn, mm = (long)n2 * n2;
  const int eig_mode = b.eig_mode;
  Solution* full = b.full;

  // ---- 1. Fourier coefficients of n^2 and n^-2 of every layer ----
  {
    const long npix = (long)s1 * s2;
    hipLaunchKernelGGL(k_square_layers, dim3(grid_for(npix * Ltot)), dim3(256), 0, s, npix * Ltot, d_samples, b.sq.p);
    for (int b0 = 0; b0 < 2 * Ltot; b0 += 65535) {
      const int nb = std::min(65535, 2 * Ltot - b0);
      zgemm_batched(OpN, OpN, s1, M0, s2, mk(1.0), b.sq.p + b0 * npix, s1, npix, b.dft2.p, s2, 0, mk(0.0),
                    b.T1.p + (long)b0 * s1 * M0, s1, (long)s1 * M0, nb, s);
      zgemm_batched(OpN, OpN, b.M1, M0, s1, mk(1.0), b.dft1.p, b.M1, 0, b.T1.p + (long)b0 * s1 * M0, s1,
                    (long)s1 * M0, mk(0.0), b.F.p + (long)b0 * MF, b.M1, MF, nb, s);
    }
  }
  const zc* NF = b.F.p;
  const zc* NI = b.F.p + (long)Ltot * MF;

  // ---- 2-3. Phi and its eigen-decomposition ----
  DevBuf<zc>& gam = *b.gam;
  auto sort_modes = [&](int reduced, const zc* Wv, int nl, const int* list, int* unsafe) {
    if (n4 <= 2048)
      hipLaunchKernelGGL(k_sort_modes<2048>, dim3(nl), dim3(256), 0, s, n4, reduced, Wv, gam.p, b.d_src.p, list,
                         unsafe);
    else
      hipLaunchKernelGGL(k_sort_modes<SORT_MAX>, dim3(nl), dim3(256), 0, s, n4, reduced, Wv, gam.p, b.d_src.p, list,
                         unsafe);
  };
  if (eig_mode == 0) {
    zc* Phi = wk.w0.p;
    hipLaunchKernelGGL(k_phi_full, dim3(grid_for(nn), Ltot), dim3(256), 0, s, ia, NF, NI, Phi, nullptr);
    zgeev_batched(n4, Phi, n4, nn, b.W.p, n4, b.V.p, n4, nn, Ltot, wk.eig, b.d_info.p, s, !capture);
    sort_modes(0, b.W.p, Ltot, nullptr, nullptr);
  } else {
    zc* P = wk.w0.p;
    zc* Q = wk.w0.p + Ltot * mm;
    zc* PQ = wk.w1.p;
    hipLaunchKernelGGL(k_phi_blocks, dim3(grid_for(mm), Ltot), dim3(256), 0, s, ia, NF, NI, P, Q);
    zgemm_batched(OpN, OpN, n2, n2, n2, mk(1.0), P, n2, mm, Q, n2, mm, mk(0.0), PQ, n2, mm, Ltot, s);
    zgeev_batched(n2, PQ, n2, mm, b.W.p, n2, b.V.p, n2, mm, Ltot, wk.eig, b.d_info.p, s, !capture);
    zgemm_batched(OpN, OpN, n2, n2, n2, mk(1.0), Q, n2, mm, b.V.p, n2, mm, mk(0.0), b.QA.p, n2, mm, Ltot, s);
    sort_modes(1, b.W.p, Ltot, nullptr, b.d_unsafe.p);
    hipLaunchKernelGGL(k_pair_scale, dim3(n2, Ltot), dim3(64), 0, s, n2, b.V.p, b.QA.p, b.W.p, b.pscale.p);
    if (!capture) {
      std::vector<int> hunsafe(Ltot);
      CWX_HIP(hipMemcpyAsync(hunsafe.data(), b.d_unsafe.p, sizeof(int) * Ltot, hipMemcpyDeviceToHost, s));
      CWX_HIP(hipStreamSynchronize(s));
      std::vector<int> fslot(Ltot, -1);
      for (int l = 0; l < Ltot; ++l)
        if (hunsafe[l]) {
          fslot[l] = (int)b.flist.size();
          b.flist.push_back(l);
        }
      if (!b.flist.empty()) {
        // nearly defective pairs: those layers' modes come from zgeev of the full Phi (P, Q are no longer needed)
        const int nf = (int)b.flist.size();
        upload(b.d_flist, b.flist, s);
        upload(b.d_fslot, fslot, s);
        zc* Phi = wk.w0.p;
        hipLaunchKernelGGL(k_phi_full, dim3(grid_for(nn), nf), dim3(256), 0, s, ia, NF, NI, Phi, b.d_flist.p);
        DevBuf<zc> Wf((size_t)nf * n4);
        b.Vf.alloc((size_t)nf * nn);
        b.d_finfo.alloc(nf);
        zgeev_batched(n4, Phi, n4, nn, Wf.p, n4, b.Vf.p, n4, nn, nf, wk.eig, b.d_finfo.p, s);
        sort_modes(0, Wf.p, nf, b.d_flist.p, nullptr);
        CWX_HIP(hipStreamSynchronize(s));  // Wf is released on return
      }
    }
  }
  DevBuf<zc>& Gd = *b.Gd;
  DevBuf<zc>& Gc = *b.Gc;
  hipLaunchKernelGGL(k_build_gdgc, dim3(grid_for(nn), Ltot), dim3(256), 0, s, ia, eig_mode, b.V.p, b.QA.p,
                     b.pscale.p, b.Vf.p, b.d_fslot.p, gam.p, b.d_src.p, is.dz.p, Gd.p, Gc.p);
  // the reduced mode's chain in block form, unless layers were redone through the full Phi (their modes lack
  // the block structure, and that Phi has overwritten P and Q)
  const bool block = eig_mode == 1 && b.flist.empty() && tune_lookup("chain_block", n4, Ltot, 1) != 0;
  // solves with many right-hand sides apply the factors' diagonal blocks by their inverses
  const bool inv = tune_lookup("getrs_inv", n4, Ltot, 1) != 0;
  b.used_block = block;
  if (!capture && !block) {
    b.V.release();
    b.QA.release();
    b.W.release();
    b.Vf.release();
    b.pscale.release();
  }

  // ---- 4. LU of every Gd, T_g = Gd_g^-1 Gc_{g-1} (g not first in its instance), and per instance
  //         S = Gc_last T_last ... T_first+1 by a pairwise product tree (ScatterVolume.cpp:327-336) ----
  zc* LUd = wk.w0.p;  // Phi / P,Q no longer needed
  int* piv = wk.ipiv.p;
  int* pivS = piv + (long)Ltot * n4;
  int* pivA = pivS + (long)count * n4;
  zc* Tm = wk.w1.p;  // T_g at slot g
  zc* Sm = wk.w2.p;
  const PtrLists& pl = b.pl;
  if (block) {
    // LU of every layer's eigenvectors Vr and of its Q; X = Vr_g^-1 Vr_{g-1}, Y = Vr_g^-1 Q_g^-1 QA_{g-1}, then T_g
    zc* Qb = wk.w0.p + (long)Ltot * mm;
    zc* LUV = wk.w0.p + 2L * Ltot * mm;
    hipLaunchKernelGGL(k_copy, dim3(grid_for((long)Ltot * mm)), dim3(256), 0, s, (long)Ltot * mm, b.V.p, LUV);
    // Q_g at Qb + g mm and V_g at Qb + (Ltot + g) mm: one batch (pivots and infos in the same order)
    const long dsz = trtri_diag_size(n2);
    if (2L * Ltot <= 65535) {
      zgetrf_batched(n2, Qb, n2, mm, piv, n2, b.d_info_lu.p, 2 * Ltot, s);
      if (inv) ztrtri_diag_set(n2, strided(Qb, mm), n2, strided(wk.w3.p, dsz), 2 * Ltot, s);
    } else {
      zgetrf_batched(n2, Qb, n2, mm, piv, n2, b.d_info_lu.p, Ltot, s);
      zgetrf_batched(n2, LUV, n2, mm, piv + (long)Ltot * n2, n2, b.d_info_lu.p + Ltot, Ltot, s);
      if (inv) {
        ztrtri_diag_set(n2, strided(Qb, mm), n2, strided(wk.w3.p, dsz), Ltot, s);
        ztrtri_diag_set(n2, strided(LUV, mm), n2, strided(wk.w3.p + Ltot * dsz, dsz), Ltot, s);
      }
    }
    if (b.nT > 0) {
      hipLaunchKernelGGL(k_copy_set, dim3(grid_for(mm), (unsigned)b.nT), dim3(256), 0, s, mm, pl.mats(b.oKVp),
                         pl.mats(b.oKX));
      hipLaunchKernelGGL(k_copy_set, dim3(grid_for(mm), (unsigned)b.nT), dim3(256), 0, s, mm, pl.mats(b.oKQAp),
                         pl.mats(b.oKZ));
      // X = V_g^-1 Vr_{g-1} and Z = Q_g^-1 QA_{g-1} as one batch of 2 nT solves (twice the workgroups per
      // launch), then Y = V_g^-1 Z
      const bool pair = 2L * b.nT <= 65535;
      if (inv) {
        // the diagonal blocks applied as products by their inverses; scratch behind the inverses in w3
        const MatSet Yt = strided(wk.w3.p + 2L * Ltot * dsz, mm);
        if (pair) {
          zgetrs_inv_set(n2, n2, pl.mats(b.oK2LU), n2, pl.ints(b.oK2P), pl.mats(b.oK2D), pl.mats(b.oK2B), n2, Yt, n2,
                         2 * b.nT, s);
        } else {
          zgetrs_inv_set(n2, n2, pl.mats(b.oKV), n2, pl.ints(b.oKPV), pl.mats(b.oKDV), pl.mats(b.oKX), n2, Yt, n2,
                         b.nT, s);
          zgetrs_inv_set(n2, n2, pl.mats(b.oKQ), n2, pl.ints(b.oKPQ), pl.mats(b.oKDQ), pl.mats(b.oKZ), n2, Yt, n2,
                         b.nT, s);
        }
        zgetrs_inv_set(n2, n2, pl.mats(b.oKV), n2, pl.ints(b.oKPV), pl.mats(b.oKDV), pl.mats(b.oKZ), n2, Yt, n2,
                       b.nT, s);
      } else {
        if (pair) {
          zgetrs_set(n2, n2, pl.mats(b.oK2LU), n2, pl.ints(b.oK2P), pl.mats(b.oK2B), n2, 2 * b.nT, s);
        } else {
          zgetrs_set(n2, n2, pl.mats(b.oKV), n2, pl.ints(b.oKPV), pl.mats(b.oKX), n2, b.nT, s);
          zgetrs_set(n2, n2, pl.mats(b.oKQ), n2, pl.ints(b.oKPQ), pl.mats(b.oKZ), n2, b.nT, s);
        }
        zgetrs_set(n2, n2, pl.mats(b.oKV), n2, pl.ints(b.oKPV), pl.mats(b.oKZ), n2, b.nT, s);
      }
      hipLaunchKernelGGL(k_t_block, dim3((unsigned)(cdiv(n4, TB_R) * cdiv(n4, TB_C)), (unsigned)b.nT), dim3(256), 0, s, ia, b.d_tg.p,
                         pl.mats(b.oKX), pl.mats(b.oKZ), pl.mats(b.oKT), gam.p, b.d_src.p, b.pscale.p, is.dz.p);
    }
  } else {
    hipLaunchKernelGGL(k_copy, dim3(grid_for((long)Ltot * nn)), dim3(256), 0, s, (long)Ltot * nn, Gd.p, LUd);
    zgetrf_batched(n4, LUd, n4, nn, piv, n4, b.d_info_lu.p, Ltot, s);
    if (b.nT > 0) {
      hipLaunchKernelGGL(k_copy, dim3(grid_for((long)(Ltot - 1) * nn)), dim3(256), 0, s, (long)(Ltot - 1) * nn, Gc.p,
                         Tm + nn);
      zgetrs_set(n4, n4, pl.mats(b.oTLU), n4, pl.ints(b.oTP), pl.mats(b.oTB), n4, b.nT, s);
    }
  }
  for (const Level& lv : b.levels)
    zgemm_set(OpN, OpN, n4, n4, n4, mk(1.0), pl.mats(lv.oA), n4, pl.mats(lv.oB), n4, mk(0.0), pl.mats(lv.oC), n4,
              lv.batch, s);
  if (b.ncS > 0)
    hipLaunchKernelGGL(k_copy_set, dim3(grid_for(nn), (unsigned)b.ncS), dim3(256), 0, s, nn, pl.mats(b.oCS),
                       pl.mats(b.oCD));

  // ---- 5. boundary system per instance: A = [Gauss; f2 | Gd0 S^-1 f3], b = [0; -f1 EF] ----
  zc* xp = b.x->p;
  zgetrf_batched(n4, Sm, n4, nn, pivS, n4, b.infoS.p, count, s);
  hipLaunchKernelGGL(k_build_f, dim3(grid_for((long)n4 * n3), count), dim3(256), 0, s, ia, 3, b.Y.p);
  if (inv) {
    // the pool is free once the tree is done: inverses of S's diagonal blocks, then the solve's scratch
    const long dsz4 = trtri_diag_size(n4);
    ztrtri_diag_set(n4, strided(Sm, nn), n4, strided(wk.w3.p, dsz4), count, s);
    zgetrs_inv_set(n4, n3, strided(Sm, nn), n4, strided(pivS, n4), strided(wk.w3.p, dsz4),
                   strided(b.Y.p, (long)n4 * n3), n4, strided(wk.w3.p + count * dsz4, (long)n4 * n3), n4, count, s);
  } else {
    zgetrs_batched(n4, n3, Sm, n4, nn, pivS, n4, b.Y.p, n4, (long)n4 * n3, count, s);
  }
  zgemm_set(OpN, OpN, n4, n3, n4, mk(1.0), pl.mats(b.oGd0), n4, strided<zc>(b.Y.p, (long)n4 * n3), n4, mk(0.0),
            strided<zc>(b.Abr.p, (long)n4 * n3), n4, count, s);
  // the reflected amplitudes eliminated order by order: a 3MF x 3MF system for T (A's storage holds it)
  const unsigned gmf = cdiv(MF, 128);
  hipLaunchKernelGGL(k_bnd_qr, dim3(gmf, count), dim3(128), 0, s, ia, b.Wq.p);
  hipLaunchKernelGGL(k_bnd_matrix, dim3(grid_for((long)n3 * n3), count), dim3(256), 0, s, ia, b.Abr.p, b.Wq.p, b.A.p,
                     (long)n6 * n6);
  hipLaunchKernelGGL(k_bnd_b, dim3(grid_for(n6), count), dim3(256), 0, s, ia, xp);
  zgetrf_batched(n3, b.A.p, n3, (long)n6 * n6, pivA, n6, b.infoA.p, count, s);
  {
    // S^-1 f3 (b.Y) is spent once Abr holds Gd0 S^-1 f3: the reduced right-hand sides and M T go there
    zc* tr = b.Y.p;
    zc* ym = tr + (long)count * n3;
    hipLaunchKernelGGL(k_bnd_rhs, dim3(gmf, 1, count), dim3(128), 0, s, ia, b.Wq.p, xp, 1, tr);
    zgetrs_batched(n3, 1, b.A.p, n3, (long)n6 * n6, pivA, n6, tr, n3, n3, count, s);
    zgemm_batched(OpN, OpN, n4, 1, n3, mk(1.0), b.Abr.p, n4, (long)n4 * n3, tr, n3, n3, mk(0.0), ym, n4, n4, count, s);
    hipLaunchKernelGGL(k_bnd_back, dim3(gmf, 1, count), dim3(128), 0, s, ia, b.Wq.p, ym, tr, xp, 1);
  }
  if (res.K > 0) {
    // b_k = [0; -f1 EF_k] for every further field, then the same factors (ScatterVolume.cpp:739-740 per field)
    CWX_CHECK(count == 1, "several incident fields are solved for one instance at a time");
    const int K = res.K;
    const int KB = std::min(K, 4096);
    DevBuf<zc> F1((size_t)n4 * n3), tr((size_t)n3 * KB), ym((size_t)n4 * KB);
    hipLaunchKernelGGL(k_build_f, dim3(grid_for((long)n4 * n3), 1), dim3(256), 0, s, ia, 1, F1.p);
    hipLaunchKernelGGL(k_fill, dim3(grid_for((long)n6 * K)), dim3(256), 0, s, (long)n6 * K, res.xs);
    for (int k0 = 0; k0 < K; k0 += KB) {
      const int kb = std::min(KB, K - k0);
      zc* xk = res.xs + (long)k0 * n6;
      zgemm_batched(OpN, OpN, n4, kb, n3, mk(-1.0), F1.p, n4, 0, res.EFs + (long)k0 * n3, n3, 0, mk(0.0), xk + n2,
                    n6, 0, 1, s);
      hipLaunchKernelGGL(k_bnd_rhs, dim3(gmf, kb, 1), dim3(128), 0, s, ia, b.Wq.p, xk, kb, tr.p);
      zgetrs_batched(n3, kb, b.A.p, n3, 0, pivA, 0, tr.p, n3, 0, 1, s);
      zgemm_batched(OpN, OpN, n4, kb, n3, mk(1.0), b.Abr.p, n4, 0, tr.p, n3, 0, mk(0.0), ym.p, n4, 0, 1, s);
      hipLaunchKernelGGL(k_bnd_back, dim3(gmf, kb, 1), dim3(128), 0, s, ia, b.Wq.p, ym.p, tr.p, xk, kb);
    }
    CWX_HIP(hipStreamSynchronize(s));  // F1, tr and ym are released on return
  }

  // ---- 6. beta recurrence, Q_check / Q_hat, Ez rows ----
  if (b.keep_beta) {
    hipLaunchKernelGGL(k_beta0_rhs, dim3(grid_for(n4), count), dim3(256), 0, s, ia, xp, pl.mats(b.oB0));
    if (block) {
      // beta_0 = Gd_0^-1 r from the factors of Vr_0 and Q_0
      hipLaunchKernelGGL(k_copy_set, dim3(grid_for(n4), count), dim3(256), 0, s, (long)n4, pl.mats(b.oK0B),
                         pl.mats(b.oK0A1));
      if (2L * count <= 65535) {
        zgetrs_set(n2, 1, pl.mats(b.oK0VQ), n2, pl.ints(b.oK0PVQ), pl.mats(b.oK0A12), n2, 2 * count, s);
      } else {
        zgetrs_set(n2, 1, pl.mats(b.oK0V), n2, pl.ints(b.oK0PV), pl.mats(b.oK0A1), n2, count, s);
        zgetrs_set(n2, 1, pl.mats(b.oK0Q), n2, pl.ints(b.oK0PQ), pl.mats(b.oK0A2), n2, count, s);
      }
      zgetrs_set(n2, 1, pl.mats(b.oK0V), n2, pl.ints(b.oK0PV), pl.mats(b.oK0A2), n2, count, s);
      hipLaunchKernelGGL(k_beta0_block, dim3(grid_for(n4), count), dim3(256), 0, s, ia, b.d_first.p,
                         pl.mats(b.oK0A1), pl.mats(b.oK0A2), pl.mats(b.oK0B), gam.p, b.d_src.p, b.pscale.p, is.dz.p);
    } else {
      zgetrs_set(n4, 1, pl.mats(b.oLU0), n4, pl.ints(b.oPiv0), pl.mats(b.oB0), n4, count, s);
    }
    for (const Level& lv : b.bsteps)
      zgemm_set(OpN, OpN, n4, 1, n4, mk(1.0), pl.mats(lv.oA), n4, pl.mats(lv.oB), n4, mk(0.0), pl.mats(lv.oC), n4,
                lv.batch, s);
  }
  // a solution with a value that is not finite is a failure, whatever the pivots said
  hipLaunchKernelGGL(k_count_nonfinite, dim3(count), dim3(256), 0, s, n6, xp, b.d_nonfinite.p);
  CWX_HIP(hipGetLastError());
}

// Status of every instance after set_launch (one synchronisation).  Solutions kept in full get their Q rows
// and records first.  In a captured solve an instance whose reduced eigenproblem had a nearly defective
// pair reports kUnsafeReduced: it is to be solved again without the graph.
constexpr int kUnsafeReduced = 1 << 29;

void set_status(Context& ctx, SetBufs& b, const std::vector<VolumeOptions>& opts, SetResult& res) {
  hipStream_t s = ctx.stream;
  const InstSet& is = b.is;
  const int count = b.count, Ltot = b.Ltot, n6 = b.n6;
  if (b.full) finish_solution(*b.full, opts[0], is, b.F.p + (long)Ltot * b.MF, s);
  // hlu: the LU of every Gd, or in the block form that of every Q (then every Vr behind it)
  const int nlu = b.used_block ? 2 * Ltot : Ltot;
  std::vector<int> hinfo(Ltot), hlu(nlu), hS(count), hA(count), hnf(count), hun(Ltot, 0);
  CWX_HIP(hipMemcpyAsync(hinfo.data(), b.d_info.p, sizeof(int) * Ltot, hipMemcpyDeviceToHost, s));
  CWX_HIP(hipMemcpyAsync(hlu.data(), b.d_info_lu.p, sizeof(int) * nlu, hipMemcpyDeviceToHost, s));
  CWX_HIP(hipMemcpyAsync(hS.data(), b.infoS.p, sizeof(int) * count, hipMemcpyDeviceToHost, s));
  CWX_HIP(hipMemcpyAsync(hA.data(), b.infoA.p, sizeof(int) * count, hipMemcpyDeviceToHost, s));
  CWX_HIP(hipMemcpyAsync(hnf.data(), b.d_nonfinite.p, sizeof(int) * count, hipMemcpyDeviceToHost, s));
  if (b.capture && b.eig_mode == 1)
    CWX_HIP(hipMemcpyAsync(hun.data(), b.d_unsafe.p, sizeof(int) * Ltot, hipMemcpyDeviceToHost, s));
  std::vector<int> hfinfo(b.flist.size());
  if (!b.flist.empty())
    CWX_HIP(hipMemcpyAsync(hfinfo.data(), b.d_finfo.p, sizeof(int) * b.flist.size(), hipMemcpyDeviceToHost, s));
  CWX_HIP(hipStreamSynchronize(s));
  for (size_t f = 0; f < b.flist.size(); ++f) hinfo[b.flist[f]] = hfinfo[f];
  res.info.assign(count, 0);
  res.msg.assign(count, std::string());
  for (int i = 0; i < count; ++i) {
    int& inf = res.info[i];
    std::string& m = res.msg[i];
    for (int l = 0; l < is.L[i] && !inf; ++l)
      if (hun[is.first[i] + l]) {
        inf = kUnsafeReduced;
        m = "layer " + std::to_string(l) + ": nearly defective pair in the reduced eigenproblem";
      }
    for (int l = 0; l < is.L[i] && !inf; ++l) {
      const int g = is.first[i] + l;
      if (hinfo[g] != 0) {
        inf = hinfo[g];
        m = "layer " + std::to_string(l) + ": eigensolver did not converge (info " + std::to_string(hinfo[g]) + ")";
      }
    }
    for (int l = 0; l < is.L[i] && !inf; ++l) {
      const int g = is.first[i] + l;
      const int ilu = b.used_block ? Ltot + g : g;  // Gd, or Vr in the block form
      if (hlu[ilu] != 0) {
        inf = hlu[ilu];
        m = "layer " + std::to_string(l) + (b.used_block ? ": singular mode matrix" : ": singular Gd") +
            " (zero pivot " + std::to_string(hlu[ilu]) + ")";
      } else if (b.used_block && hlu[g] != 0) {
        inf = hlu[g];
        m = "layer " + std::to_string(l) + ": singular Q block (zero pivot " + std::to_string(hlu[g]) + ")";
      }
    }
    if (!inf && hS[i]) { inf = hS[i]; m = "singular transfer product"; }
    if (!inf && hA[i]) { inf = hA[i]; m = "singular boundary system"; }
    if (!inf && hnf[i]) {
      inf = n6 + 1;
      m = std::to_string(hnf[i]) + " entries of the boundary solution are not finite";
    }
  }
}

void solve_set(Context& ctx, const std::vector<VolumeOptions>& opts, const zc* d_samples, const std::vector<int>& Ls,
               int s1, int s2, SetResult& res) {
  SetBufs b;
  set_prepare(ctx, opts, Ls, s1, s2, res, b, false);
  set_launch(ctx, b, d_samples, s1, s2, res);
  set_status(ctx, b, opts, res);
}

}  // namespace

void solve_volume(Context& ctx, const VolumeOptions& o, const zc* d_sample, int L, int s1, int s2,
                  Solution& sol) {
  CWX_CHECK(L <= 65535, "too many layers");
  SetResult res;
  res.x = &sol.x;
  res.full = &sol;
  solve_set(ctx, {o}, d_sample, {L}, s1, s2, res);
  sol.info = res.info[0];
  sol.info_msg = res.msg[0];
}

int solve_batch(Context& ctx, const std::vector<VolumeOptions>& opts, const zc* d_samples, const std::vector<int>& Ls,
                int s1, int s2, size_t max_bytes, zc* d_x, std::vector<int>& info, std::string& first_msg) {
  const int count = (int)opts.size();
  CWX_CHECK(count >= 1 && (int)Ls.size() == count, "batch is empty");
  // orders of the first instance fix the per-layer working set
  const Setup st0 = make_setup(opts[0], Ls[0], s1, s2);
  const long n4 = 4L * st0.MF, n2 = 2L * st0.MF, n6 = 6L * st0.MF;
  const long nn = n4 * n4;
  const long n_eig = opts[0].eig_mode ? n2 : n4;
  // bytes per layer: Phi or P,Q + LU (w0), PQ + T (w1), pool (w3, or the inverses of the diagonal blocks of Q
  // and V and the solves' scratch where that is more), Gd, Gc, eigen work (Z, X), V, QA
  const long w3_layer = std::max(nn, opts[0].eig_mode ? 2 * trtri_diag_size((int)n2) + 2 * n2 * n2 : 0L);
  const size_t per_layer = sizeof(zc) * (size_t)(nn * 4 + w3_layer + n_eig * n_eig * 4 + (long)s1 * s2 * 2);
  const size_t per_inst =
      sizeof(zc) * (size_t)(nn + 2 * n4 * 3 * st0.MF + n6 * n6 + trtri_diag_size((int)n4) + n4 * 3 * st0.MF);
  if (max_bytes == 0) max_bytes = (size_t)tune_lookup("batch_chunk_mib", 4 * opts[0].coef[0] * opts[0].coef[1], count,
                                                      64 << 10) << 20;
  info.assign(count, 0);
  first_msg.clear();
  int worst = 0;
  const long npix = (long)s1 * s2;
  long layer0 = 0;
  for (int i0 = 0; i0 < count;) {
    // grow the chunk while it fits the budget and the launch limits
    int i1 = i0;
    size_t bytes = 0;
    long layers = 0;
    while (i1 < count) {
      const size_t add = per_layer * Ls[i1] + per_inst;
      if (i1 > i0 && (bytes + add > max_bytes || layers + Ls[i1] > 60000)) break;
      bytes += add;
      layers += Ls[i1];
      ++i1;
    }
    std::vector<VolumeOptions> o(opts.begin() + i0, opts.begin() + i1);
    std::vector<int> l(Ls.begin() + i0, Ls.begin() + i1);
    DevBuf<zc> xs;
    SetResult res;
    res.x = &xs;
    solve_set(ctx, o, d_samples + layer0 * npix, l, s1, s2, res);
    CWX_HIP(hipMemcpyAsync(d_x + (long)i0 * n6, xs.p, sizeof(zc) * n6 * (i1 - i0), hipMemcpyDeviceToDevice,
                           ctx.stream));
    CWX_HIP(hipStreamSynchronize(ctx.stream));
    for (int i = i0; i < i1; ++i) {
      info[i] = res.info[i - i0];
      if (info[i] && !worst) {
        worst = info[i];
        first_msg = "instance " + std::to_string(i) + ": " + res.msg[i - i0];
      }
    }
    layer0 += layers;
    i0 = i1;
  }
  return worst;
}

namespace {

// A captured set solve: its plan (buffers, tables, pointer lists), its sample and result areas and the
// instantiated graph of its launches, kept in the context and replayed for every chunk whose instances
// have the same layer counts, orders and eigen mode.
struct SetGraph {
  std::vector<int> Ls;
  int s1 = 0, s2 = 0, M0 = 0, M1 = 0, eig_mode = 0;
  SetBufs b;
  DevBuf<zc> samples, x;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  ~SetGraph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

}  // namespace

int solve_batch_graph(Context& ctx, const std::vector<VolumeOptions>& opts, const zc* d_samples,
                      const std::vector<int>& Ls, int s1, int s2, int chunk, zc* d_x, std::vector<int>& info,
                      std::string& first_msg, int* captured) {
  const int count = (int)opts.size();
  CWX_CHECK(count >= 1 && (int)Ls.size() == count, "batch is empty");
  CWX_CHECK(chunk >= 1, "chunk must be at least one instance");
  hipStream_t s = ctx.stream;
  const Setup st0 = make_setup(opts[0], Ls[0], s1, s2);
  const long n6 = 6L * st0.MF, npix = (long)s1 * s2;
  info.assign(count, 0);
  first_msg.clear();
  int worst = 0, ncap = 0;
  long layer0 = 0;
  for (int i0 = 0; i0 < count; i0 += chunk) {
    const int i1 = std::min(count, i0 + chunk);
    std::vector<VolumeOptions> o(opts.begin() + i0, opts.begin() + i1);
    std::vector<int> l(Ls.begin() + i0, Ls.begin() + i1);
    long layers = 0;
    for (int v : l) layers += v;
    const Setup sc = make_setup(o[0], l[0], s1, s2);
    SetGraph* g = nullptr;
    for (auto& e : ctx.graphs) {
      SetGraph* c = static_cast<SetGraph*>(e.get());
      if (c->Ls == l && c->s1 == s1 && c->s2 == s2 && c->M0 == sc.M0 && c->M1 == sc.M1 &&
          c->eig_mode == o[0].eig_mode) {
        g = c;
        break;
      }
    }
    SetResult res;
    if (!g) {
      auto ng = std::make_shared<SetGraph>();
      ng->Ls = l;
      ng->s1 = s1;
      ng->s2 = s2;
      ng->M0 = sc.M0;
      ng->M1 = sc.M1;
      ng->eig_mode = o[0].eig_mode;
      res.x = &ng->x;
      set_prepare(ctx, o, l, s1, s2, res, ng->b, true);
      ng->samples.alloc((size_t)layers * npix);
      CWX_HIP(hipStreamSynchronize(s));
      CWX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        capture_scope cs;  // no device synchronisation by the buffer cache while the stream is captured
        set_launch(ctx, ng->b, ng->samples.p, s1, s2, res);
      } catch (...) {
        hipGraph_t dead = nullptr;
        (void)hipStreamEndCapture(s, &dead);
        if (dead) (void)hipGraphDestroy(dead);
        throw;
      }
      CWX_HIP(hipStreamEndCapture(s, &ng->graph));
      CWX_HIP(hipGraphInstantiate(&ng->exec, ng->graph, nullptr, nullptr, 0));
      ctx.graphs.push_back(ng);
      g = ng.get();
      ++ncap;
    } else {
      // the chunk's own wavelengths, incidences and tables into the plan's buffers (same sizes)
      make_instances(o, l, s1, s2, g->b.is, s);
    }
    CWX_HIP(hipMemcpyAsync(g->samples.p, d_samples + layer0 * npix, sizeof(zc) * layers * npix,
                           hipMemcpyDeviceToDevice, s));
    CWX_HIP(hipGraphLaunch(g->exec, s));
    CWX_HIP(hipMemcpyAsync(d_x + (long)i0 * n6, g->x.p, sizeof(zc) * n6 * (i1 - i0), hipMemcpyDeviceToDevice, s));
    set_status(ctx, g->b, o, res);
    long lay = layer0;
    for (int i = i0; i < i1; ++i) {
      int inf = res.info[i - i0];
      std::string msg = res.msg[i - i0];
      if (inf == kUnsafeReduced) {
        // a nearly defective reduced pair: this instance again, outside the graph, where its layers can
        // take the full Phi
        DevBuf<zc> xi;
        SetResult r1;
        r1.x = &xi;
        solve_set(ctx, {opts[i]}, d_samples + lay * npix, {Ls[i]}, s1, s2, r1);
        CWX_HIP(hipMemcpyAsync(d_x + (long)i * n6, xi.p, sizeof(zc) * n6, hipMemcpyDeviceToDevice, s));
        CWX_HIP(hipStreamSynchronize(s));
        inf = r1.info[0];
        msg = r1.msg[0];
      }
      info[i] = inf;
      if (inf && !worst) {
        worst = inf;
        first_msg = "instance " + std::to_string(i) + ": " + msg;
      }
      lay += Ls[i];
    }
    layer0 += layers;
  }
  if (captured) *captured = ncap;
  return worst;
}

int solve_batch_mixed(Context& ctx, const std::vector<VolumeOptions>& opts, const zc* d_samples,
                      const std::vector<int>& Ls, int s1, int s2, size_t max_bytes, zc* d_x,
                      const std::vector<long>& x_off, std::vector<int>& info, std::string& first_msg) {
  const int count = (int)opts.size();
  CWX_CHECK(count >= 1 && (int)Ls.size() == count && (int)x_off.size() == count + 1, "batch is empty");
  // group the instances by their orders and eigen mode; each group is one batched solve
  std::vector<long> lay0(count + 1, 0);
  for (int i = 0; i < count; ++i) lay0[i + 1] = lay0[i] + Ls[i];
  std::vector<std::array<int, 3>> key(count);
  for (int i = 0; i < count; ++i) {
    const Setup st = make_setup(opts[i], Ls[i], s1, s2);
    key[i] = {st.M0, st.M1, opts[i].eig_mode};
  }
  std::vector<int> order(count);
  for (int i = 0; i < count; ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
  info.assign(count, 0);
  first_msg.clear();
  int worst = 0;
  const long npix = (long)s1 * s2;
  for (size_t g0 = 0; g0 < order.size();) {
    size_t g1 = g0;
    while (g1 < order.size() && key[order[g1]] == key[order[g0]]) ++g1;
    std::vector<VolumeOptions> o;
    std::vector<int> l;
    long layers = 0;
    for (size_t t = g0; t < g1; ++t) {
      o.push_back(opts[order[t]]);
      l.push_back(Ls[order[t]]);
      layers += Ls[order[t]];
    }
    // the group's layers back to back, and its [R; T] rows
    DevBuf<zc> samples((size_t)layers * npix);
    long at = 0;
    for (size_t t = g0; t < g1; ++t) {
      const int i = order[t];
      CWX_HIP(hipMemcpyAsync(samples.p + at * npix, d_samples + lay0[i] * npix, sizeof(zc) * Ls[i] * npix,
                             hipMemcpyDeviceToDevice, ctx.stream));
      at += Ls[i];
    }
    const long n6 = x_off[order[g0] + 1] - x_off[order[g0]];
    DevBuf<zc> xs((size_t)n6 * (g1 - g0));
    std::vector<int> ginfo;
    std::string msg;
    solve_batch(ctx, o, samples.p, l, s1, s2, max_bytes, xs.p, ginfo, msg);
    for (size_t t = g0; t < g1; ++t) {
      const int i = order[t];
      CWX_CHECK(x_off[i + 1] - x_off[i] == n6, "x offsets do not match the orders");
      CWX_HIP(hipMemcpyAsync(d_x + x_off[i], xs.p + (long)(t - g0) * n6, sizeof(zc) * n6, hipMemcpyDeviceToDevice,
                             ctx.stream));
      info[i] = ginfo[t - g0];
    }
    CWX_HIP(hipStreamSynchronize(ctx.stream));  // the group buffers are released on return
    for (int i = 0; i < count; ++i)
      if (info[i] && !worst) {
        worst = info[i];
        first_msg = "instance " + std::to_string(i) + " failed";
      }
    g0 = g1;
  }
  return worst;
}

int solve_fields(Context& ctx, const VolumeOptions& o, const zc* d_sample, int L, int s1, int s2, int K,
                 const zc* d_EF, zc* d_x, std::string& msg) {
  CWX_CHECK(L >= 1 && L <= 65535, "bad layer count");
  CWX_CHECK(K >= 1 && d_EF && d_x, "need at least one incident field");
  DevBuf<zc> x0;
  SetResult res;
  res.x = &x0;
  res.K = K;
  res.EFs = d_EF;
  res.xs = d_x;
  solve_set(ctx, {o}, d_sample, {L}, s1, s2, res);
  msg = res.msg[0];
  return res.info[0];
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


// ------------------------------------------------------------------------------------------------
// Compact heterogeneous .cw (docs/cw_file_format.pptx slide 1; reader ScatterView.py:68-178):
// is_volume = 1, precision 8, the incident waves, every boundary's z with only the external lists
// (reflected at the first boundary, transmitted at the last; the interior waves follow from the mode
// data), then M0 M1 (int32), the sample size (3 doubles), the layer count (uint32) and per layer
// beta (4MF), gamma (4MF), gg (the 4MF x 4MF mode matrix row by row, gg[r 4MF + m] = V(r, m): rows
// Ex, Ey and the two transverse magnetic blocks, columns the interleaved down / up modes, as
// cpuEvaluator.cpp:171-205 reads it) and ft (the n^-2 coefficients in the solver's storage order:
// the (M1 x M0) matrix column by column, which ScatterVolume.cpp:872 and cpuEvaluator.cpp:225 index
// alike; as the sample is transposed on load (FourierWave.h:288), ScatterView.py:174's row-major
// reshape gives square layers back as the .npy file holds them).  After the
// layers comes a trailer this library adds (the reference reader stops before it): the magic
// "CWXTRL01", then lambda, direction[3], ex[2], ey[2], ez[2], n[2], kappa, z (doubles), the solve's
// grid and remaining options (s1, s2, ncoef, coef[2], psf, external, npz, no_cutoff, eig_mode: int32
// each) and the boundary solution x (6MF complex): what rebuilding the solution needs beyond the
// reference's layout.
// ------------------------------------------------------------------------------------------------
static const char kCw2Magic[8] = {'C', 'W', 'X', 'T', 'R', 'L', '0', '1'};

void cw2_write_file(const Solution& sol, const std::string& path, hipStream_t s) {
  const Setup& st = sol.st;
  const int L = st.L, ZL = L + 1, MF = st.MF, M0 = st.M0, M1 = st.M1, n4 = 4 * MF, n6 = 6 * MF;
  CWX_CHECK(sol.Gd.p && sol.Gc.p && sol.beta.p && sol.gam.p && sol.NIF.p && sol.x.p, "solution has no device data");
  const size_t nn = (size_t)n4 * n4;
  std::vector<zc> rec((size_t)sol.nrec * 6), gam((size_t)L * n4), beta((size_t)L * n4), Gd(nn * L), Gc(nn * L),
      nif((size_t)L * MF), x(n6);
  if (sol.nrec) CWX_HIP(hipMemcpyAsync(rec.data(), sol.rec.p, sizeof(zc) * rec.size(), hipMemcpyDeviceToHost, s));
  CWX_HIP(hipMemcpyAsync(gam.data(), sol.gam.p, sizeof(zc) * gam.size(), hipMemcpyDeviceToHost, s));
  CWX_HIP(hipMemcpyAsync(beta.data(), sol.beta.p, sizeof(zc) * beta.size(), hipMemcpyDeviceToHost, s));
  CWX_HIP(hipMemcpyAsync(Gd.data(), sol.Gd.p, sizeof(zc) * Gd.size(), hipMemcpyDeviceToHost, s));
  CWX_HIP(hipMemcpyAsync(Gc.data(), sol.Gc.p, sizeof(zc) * Gc.size(), hipMemcpyDeviceToHost, s));
  CWX_HIP(hipMemcpyAsync(nif.data(), sol.NIF.p, sizeof(zc) * nif.size(), hipMemcpyDeviceToHost, s));
  CWX_HIP(hipMemcpyAsync(x.data(), sol.x.p, sizeof(zc) * x.size(), hipMemcpyDeviceToHost, s));
  CWX_HIP(hipStreamSynchronize(s));
  std::ofstream f(path, std::ios::binary);
  CWX_CHECK((bool)f, "cannot open " + path + " for writing");
  auto put = [&](const void* p, size_t n) { f.write((const char*)p, (std::streamsize)n); };
  auto put_u64 = [&](uint64_t v) { put(&v, 8); };
  auto put_list = [&](int seg) {
    const uint64_t cnt = (uint64_t)sol.seg_cnt[seg];
    put_u64(cnt);
    put(rec.data() + sol.seg_off[seg] * 6, cnt * 96);
  };
  const unsigned char is_volume = 1;
  put(&is_volume, 1);
  put_u64(8);
  put_list(0);
  put_u64((uint64_t)ZL);
  for (int l = 0; l < ZL; ++l) {
    put(&st.z[l], 8);
    if (l == 0) put_list(1);
    else put_u64(0);
    if (l == ZL - 1) put_list(2 + 2 * l);
    else put_u64(0);
  }
  const int32_t m[2] = {M0, M1};
  put(m, 8);
  put(sol.opt.size, 24);
  const uint32_t nl = (uint32_t)L;
  put(&nl, 4);
  std::vector<zc> gg(nn);
  for (int l = 0; l < L; ++l) {
    put(beta.data() + (size_t)l * n4, sizeof(zc) * n4);
    put(gam.data() + (size_t)l * n4, sizeof(zc) * n4);
    // even modes are the columns of Gd, odd modes those of Gc (k_build_gdgc)
    for (int mo = 0; mo < n4; ++mo) {
      const zc* col = ((mo & 1) ? Gc.data() : Gd.data()) + nn * l + (size_t)mo * n4;
      for (int r = 0; r < n4; ++r) gg[(size_t)r * n4 + mo] = col[r];
    }
    put(gg.data(), sizeof(zc) * nn);
    put(nif.data() + (size_t)l * MF, sizeof(zc) * MF);
  }
  put(kCw2Magic, 8);
  const VolumeOptions& o = sol.opt;
  const double d[13] = {o.lambda, o.direction[0], o.direction[1], o.direction[2], o.ex[0], o.ex[1], o.ey[0],
                        o.ey[1], o.ez[0], o.ez[1], o.n[0], o.n[1], o.kappa};
  put(d, sizeof(d));
  put(&o.z, 8);
  const int32_t iv[10] = {st.s1, st.s2, o.ncoef, o.coef[0], o.coef[1], o.psf ? 1 : 0, o.external ? 1 : 0,
                          o.npz ? 1 : 0, o.no_cutoff ? 1 : 0, o.eig_mode};
  put(iv, sizeof(iv));
  put(x.data(), sizeof(zc) * n6);
  CWX_CHECK((bool)f, "write to " + path + " failed");
}

void cw2_read_file(Context& ctx, const std::string& path, Solution& sol) {
  hipStream_t s = ctx.stream;
  std::ifstream f(path, std::ios::binary);
  CWX_CHECK((bool)f, "cannot open " + path);
  auto get = [&](void* p, size_t n) {
    f.read((char*)p, (std::streamsize)n);
    CWX_CHECK((bool)f, path + ": truncated compact .cw file");
  };
  auto get_u64 = [&]() {
    uint64_t v;
    get(&v, 8);
    return v;
  };
  auto skip_list = [&]() {
    const uint64_t cnt = get_u64();
    CWX_CHECK(cnt < (1ull << 40), path + ": implausible wave count");
    f.seekg((std::streamoff)(cnt * 96), std::ios::cur);
  };
  unsigned char is_volume = 0;
  get(&is_volume, 1);
  CWX_CHECK(is_volume == 1, path + ": not a heterogeneous (is_volume) .cw file");
  CWX_CHECK(get_u64() == 8, path + ": only double-precision .cw files are supported");
  skip_list();
  const uint64_t ZL = get_u64();
  CWX_CHECK(ZL >= 2 && ZL <= 65536, path + ": implausible boundary count");
  for (uint64_t l = 0; l < ZL; ++l) {
    double z;
    get(&z, 8);
    skip_list();
    skip_list();
  }
  int32_t m[2];
  get(m, 8);
  double size[3];
  get(size, 24);
  uint32_t L;
  get(&L, 4);
  CWX_CHECK(m[0] >= 1 && m[1] >= 1 && (long)m[0] * m[1] * 4 <= SORT_MAX, path + ": unsupported Fourier orders");
  CWX_CHECK(L >= 1 && L + 1 == ZL, path + ": layer count does not match the boundaries");
  const int M0 = m[0], M1 = m[1], MF = M0 * M1, n4 = 4 * MF, n6 = 6 * MF;
  const size_t nn = (size_t)n4 * n4;
  CWX_CHECK(sizeof(zc) * nn * L <= ((size_t)64 << 30), path + ": mode data larger than 64 GiB");
  std::vector<zc> beta((size_t)L * n4), gam((size_t)L * n4), V(nn * L), nif((size_t)L * MF), gg(nn);
  for (uint32_t l = 0; l < L; ++l) {
    get(beta.data() + (size_t)l * n4, sizeof(zc) * n4);
    get(gam.data() + (size_t)l * n4, sizeof(zc) * n4);
    get(gg.data(), sizeof(zc) * nn);
    for (int r = 0; r < n4; ++r)
      for (int mo = 0; mo < n4; ++mo) V[nn * l + (size_t)mo * n4 + r] = gg[(size_t)r * n4 + mo];
    get(nif.data() + (size_t)l * MF, sizeof(zc) * MF);
  }
  char magic[8];
  f.read(magic, 8);
  CWX_CHECK((bool)f && std::memcmp(magic, kCw2Magic, 8) == 0,
            path + ": no solver trailer (the reference's layout holds neither the wavelength nor the incident "
                   "field, so the solution cannot be rebuilt from it)");
  double d[13];
  get(d, sizeof(d));
  VolumeOptions o;
  o.lambda = d[0];
  for (int i = 0; i < 3; ++i) o.direction[i] = d[1 + i];
  o.ex[0] = d[4]; o.ex[1] = d[5]; o.ey[0] = d[6]; o.ey[1] = d[7]; o.ez[0] = d[8]; o.ez[1] = d[9];
  o.n[0] = d[10]; o.n[1] = d[11]; o.kappa = d[12];
  get(&o.z, 8);
  int32_t iv[10];
  get(iv, sizeof(iv));
  const int s1 = iv[0], s2 = iv[1];
  CWX_CHECK(s1 >= 1 && s2 >= 1 && iv[2] >= 0 && iv[2] <= 2, path + ": bad solver trailer");
  o.ncoef = iv[2];
  o.coef[0] = iv[3];
  o.coef[1] = iv[4];
  o.psf = iv[5] != 0;
  o.external = iv[6] != 0;
  o.npz = iv[7] != 0;
  o.no_cutoff = iv[8] != 0;
  o.eig_mode = iv[9];
  for (int i = 0; i < 3; ++i) o.size[i] = size[i];
  std::vector<zc> x(n6);
  get(x.data(), sizeof(zc) * n6);
  // the solve's own options and grid give back its setup (orders, tables, cut-off flag) exactly
  InstSet is;
  make_instances({o}, {(int)L}, s1, s2, is, s);
  CWX_CHECK(is.st[0].M0 == M0 && is.st[0].M1 == M1, path + ": orders do not match the solver trailer");
  sol.gam.alloc(gam.size());
  sol.beta.alloc(beta.size());
  sol.Gd.alloc(V.size());
  sol.Gc.alloc(V.size());
  sol.x.alloc(x.size());
  DevBuf<zc> NI(nif.size()), dV(V.size());
  std::vector<int> ident((size_t)L * n4);
  for (size_t i = 0; i < ident.size(); ++i) ident[i] = (int)(i % n4);
  DevBuf<int> src;
  upload(src, ident, s);
  CWX_HIP(hipMemcpyAsync(sol.gam.p, gam.data(), sizeof(zc) * gam.size(), hipMemcpyHostToDevice, s));
  CWX_HIP(hipMemcpyAsync(sol.beta.p, beta.data(), sizeof(zc) * beta.size(), hipMemcpyHostToDevice, s));
  CWX_HIP(hipMemcpyAsync(dV.p, V.data(), sizeof(zc) * V.size(), hipMemcpyHostToDevice, s));
  CWX_HIP(hipMemcpyAsync(sol.x.p, x.data(), sizeof(zc) * x.size(), hipMemcpyHostToDevice, s));
  CWX_HIP(hipMemcpyAsync(NI.p, nif.data(), sizeof(zc) * nif.size(), hipMemcpyHostToDevice, s));
  // Gd / Gc: the mode matrix with the layer phases, by the solve's own kernel (the modes are sorted)
  hipLaunchKernelGGL(k_build_gdgc, dim3(grid_for((long)nn), L), dim3(256), 0, s, is.args(), 0, dV.p, nullptr,
                     nullptr, nullptr, nullptr, sol.gam.p, src.p, is.dz.p, sol.Gd.p, sol.Gc.p);
  CWX_HIP(hipGetLastError());
  finish_solution(sol, o, is, NI.p, s);
  CWX_HIP(hipStreamSynchronize(s));
  sol.info = 0;
  sol.info_msg.clear();
}

}  // namespace cwx
