// Plane-wave field evaluation: E(r) = sum_w E0_w exp(i k_w . r) over the waves of the region holding r.
//
// Replaces gpuEvaluator.cu:28-165 (kernel_xy / kernel_xz / kernel_yz + evaluate).  Semantics kept:
//   * region of a point = (number of boundaries z_l with z >= z_l - 0.001)   (gpuEvaluator.cu:32-37);
//     the CPU evaluator's tolerance-free test (cpuEvaluator.cpp:303,349,386) is selectable;
//   * waves of region l are the contiguous range [waves_begin[l], waves_end[l]) of the unpacked table
//     (cpuEvaluator.cpp:34-53);
//   * slice coordinates are float: x = x_start + d * xi (gpuEvaluator.cu:78-79, 99-100, 119-120).
//
// Layout and launch (MI355X): one workgroup = 256 lanes x PTS points that share one z (hence one
// region), so the wave loop is uniform across the workgroup; waves are staged through LDS in chunks of
// 256 records and read back as broadcasts.  Two arithmetic forms:
//   * F32  : the reference's arithmetic (FP32 table, FP32 phase, FP32 accumulation);
//   * F64R : FP64 table, FP64 phase argument reduced modulo 2*pi (two-constant Cody-Waite), FP32
//            sincos/exp of the reduced argument, FP64 accumulation -> ~1e-7 relative per term.
#include "cwx_common.h"
#include "cwx_kernels.h"

namespace cwx {

namespace {

// grid coordinate origin + step * i with a single rounding (a fused multiply-add, the form a compiler
// with contraction gives x_start + d * xi of gpuEvaluator.cu:78-79), reproducible bit for bit on the host
__device__ inline float coord(float o, float d, int i) { return fmaf(d, (float)i, o); }

constexpr int kFieldThreads = 256;

__device__ inline int region_of(float z, const float* __restrict__ zl, int layers, int tol_mode) {
  int l = 0;
  if (tol_mode == 0) {
    for (int li = 0; li < layers; ++li)
      if ((double)z >= (double)zl[li] - 0.001) l = li + 1;
  } else {
    for (int li = 0; li < layers; ++li)
      if (z >= zl[li]) l = li + 1;
  }
  return l;
}

struct WaveF32 {
  float2 e[3];
  float2 k[3];
};
struct WaveF64 {
  double2 e[3];
  double2 k[3];
};

template <int PTS>
__global__ __launch_bounds__(kFieldThreads) void k_field_f32(
    const WaveF32* __restrict__ W, const float* __restrict__ zl, const int* __restrict__ wb,
    const int* __restrict__ we, int layers, int tol_mode, GridSpec g, float2* __restrict__ out) {
  __shared__ WaveF32 sw[kFieldThreads];
  const int iz = blockIdx.y;
  const long plane = (long)g.nx * g.ny;
  const float z = coord(g.z0, g.dz, iz);
  const int l = region_of(z, zl, layers, tol_mode);
  const int b = wb[l], e = we[l];

  float px[PTS], py[PTS];
  float2 E[PTS][3];
  long idx[PTS];
#pragma unroll
  for (int p = 0; p < PTS; ++p) {
    idx[p] = (long)blockIdx.x * kFieldThreads * PTS + p * kFieldThreads + threadIdx.x;
    long q = idx[p] < plane ? idx[p] : 0;
    int ix = (int)(q % g.nx), iy = (int)(q / g.nx);
    px[p] = coord(g.x0, g.dx, ix);
    py[p] = coord(g.y0, g.dy, iy);
#pragma unroll
    for (int c = 0; c < 3; ++c) E[p][c] = make_float2(0.f, 0.f);
  }
  for (int base = b; base < e; base += kFieldThreads) {
    const int cnt = min(kFieldThreads, e - base);
    __syncthreads();
    if ((int)threadIdx.x < cnt) sw[threadIdx.x] = W[base + threadIdx.x];
    __syncthreads();
    for (int w = 0; w < cnt; ++w) {
      const WaveF32 wv = sw[w];
#pragma unroll
      for (int p = 0; p < PTS; ++p) {
        // k . r with complex k and real r, in the reference's order: kx*x + ky*y + kz*z
        float kr_re = wv.k[0].x * px[p] + wv.k[1].x * py[p];
        float kr_im = wv.k[0].y * px[p] + wv.k[1].y * py[p];
        kr_re = kr_re + wv.k[2].x * z;
        kr_im = kr_im + wv.k[2].y * z;
        float s, c;
        sincosf(kr_re, &s, &c);
        const float m = expf(-kr_im);
        const float pr = m * c, pi = m * s;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          E[p][q].x += wv.e[q].x * pr - wv.e[q].y * pi;
          E[p][q].y += wv.e[q].x * pi + wv.e[q].y * pr;
        }
      }
    }
  }
#pragma unroll
  for (int p = 0; p < PTS; ++p) {
    if (idx[p] < plane) {
      float2* o = out + ((long)iz * plane + idx[p]) * 3;
      o[0] = E[p][0];
      o[1] = E[p][1];
      o[2] = E[p][2];
    }
  }
}

__device__ inline void reduced_phase(double arg_re, double arg_im, float* pr, float* pi) {
  // arg = k . r; returns exp(i arg) with the real part of arg reduced modulo 2*pi in FP64.
  const double inv2pi = 0.15915494309189535;
  const double twopi_hi = 6.283185307179586;       // 2*pi rounded to double
  const double twopi_lo = 2.4492935982947064e-16;  // 2*pi - twopi_hi
  const double n = rint(arg_re * inv2pi);
  double r = fma(-n, twopi_hi, arg_re);
  r = fma(-n, twopi_lo, r);
  float s, c;
  sincosf((float)r, &s, &c);
  const float m = expf((float)(-arg_im));
  *pr = m * c;
  *pi = m * s;
}

template <int PTS, typename OutT>
__global__ __launch_bounds__(kFieldThreads) void k_field_f64r(
    const WaveF64* __restrict__ W, const float* __restrict__ zl, const int* __restrict__ wb,
    const int* __restrict__ we, int layers, int tol_mode, GridSpec g, OutT* __restrict__ out) {
  __shared__ WaveF64 sw[kFieldThreads];
  const int iz = blockIdx.y;
  const long plane = (long)g.nx * g.ny;
  const float zf = coord(g.z0, g.dz, iz);
  const double z = (double)zf;
  const int l = region_of(zf, zl, layers, tol_mode);
  const int b = wb[l], e = we[l];

  double px[PTS], py[PTS];
  double2 E[PTS][3];
  long idx[PTS];
#pragma unroll
  for (int p = 0; p < PTS; ++p) {
    idx[p] = (long)blockIdx.x * kFieldThreads * PTS + p * kFieldThreads + threadIdx.x;
    long q = idx[p] < plane ? idx[p] : 0;
    int ix = (int)(q % g.nx), iy = (int)(q / g.nx);
    px[p] = (double)(coord(g.x0, g.dx, ix));
    py[p] = (double)(coord(g.y0, g.dy, iy));
#pragma unroll
    for (int c = 0; c < 3; ++c) E[p][c] = make_double2(0.0, 0.0);
  }
  for (int base = b; base < e; base += kFieldThreads) {
    const int cnt = min(kFieldThreads, e - base);
    __syncthreads();
    if ((int)threadIdx.x < cnt) sw[threadIdx.x] = W[base + threadIdx.x];
    __syncthreads();
    for (int w = 0; w < cnt; ++w) {
      const WaveF64 wv = sw[w];
      const double zr = wv.k[2].x * z, zi = wv.k[2].y * z;
#pragma unroll
      for (int p = 0; p < PTS; ++p) {
        const double ar = fma(wv.k[0].x, px[p], fma(wv.k[1].x, py[p], zr));
        const double ai = fma(wv.k[0].y, px[p], fma(wv.k[1].y, py[p], zi));
        float fr, fi;
        reduced_phase(ar, ai, &fr, &fi);
        const double pr = fr, pi = fi;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          E[p][q].x = fma(wv.e[q].x, pr, fma(-wv.e[q].y, pi, E[p][q].x));
          E[p][q].y = fma(wv.e[q].x, pi, fma(wv.e[q].y, pr, E[p][q].y));
        }
      }
    }
  }
#pragma unroll
  for (int p = 0; p < PTS; ++p) {
    if (idx[p] < plane) {
      OutT* o = out + ((long)iz * plane + idx[p]) * 3;
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        if constexpr (sizeof(OutT) == 8)
          o[q] = OutT{(float)E[p][q].x, (float)E[p][q].y};
        else
          o[q] = OutT{E[p][q].x, E[p][q].y};
      }
    }
  }
}

}  // namespace

void launch_field_direct(const FieldTable& t, const GridSpec& g, int tol_mode, int out_f64, void* out,
                         hipStream_t s) {
  CWX_CHECK(g.nx > 0 && g.ny > 0 && g.nz > 0, "field grid must be non-empty");
  CWX_CHECK(t.nwaves > 0 && t.layers >= 0, "no wave table uploaded");
  const long plane = (long)g.nx * g.ny;
  if (t.precision == 4) {
    CWX_CHECK(!out_f64, "an FP32 wave table produces complex64 output");
    constexpr int P = 1;
    dim3 grid(cdiv(plane, kFieldThreads * P), g.nz);
    hipLaunchKernelGGL((k_field_f32<P>), grid, dim3(kFieldThreads), 0, s, (const WaveF32*)t.waves,
                       t.z_layers, t.wb, t.we, t.layers, tol_mode, g, (float2*)out);
  } else {
    constexpr int P = 1;
    dim3 grid(cdiv(plane, kFieldThreads * P), g.nz);
    if (out_f64)
      hipLaunchKernelGGL((k_field_f64r<P, double2>), grid, dim3(kFieldThreads), 0, s,
                         (const WaveF64*)t.waves, t.z_layers, t.wb, t.we, t.layers, tol_mode, g,
                         (double2*)out);
    else
      hipLaunchKernelGGL((k_field_f64r<P, float2>), grid, dim3(kFieldThreads), 0, s,
                         (const WaveF64*)t.waves, t.z_layers, t.wb, t.we, t.layers, tol_mode, g,
                         (float2*)out);
  }
  CWX_HIP(hipGetLastError());
}

}  // namespace cwx
