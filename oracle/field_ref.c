/*
 * CPU restatement of the reference's plane-wave field evaluators.  TEST INFRASTRUCTURE ONLY: used as
 * the checker and as the CPU baseline of the field benchmark, never by the product path.
 *
 *   cpuEvaluator.cpp:294-407  cpu_cw_evaluate_xy / _yz / _xz: E(r) = sum_w E0_w exp(i k_w . r) over the
 *                             waves of the region of r, complex<float> arithmetic;
 *   gpuEvaluator.cu:32-37     the region test z >= z_l - 0.001 (tol_mode 0) or z >= z_l (tol_mode 1).
 *
 * Grid coordinates are origin + step * i rounded once (fmaf), as the GPU kernels compute them.
 * cw_eval_grid_f32 keeps the reference's FP32 arithmetic; cw_eval_grid_f64 evaluates the same sum in
 * FP64 on FP64 waves.  Both are OpenMP-parallel over points.
 *
 * Pins: tests/test_capi_cpu.py compares both against oracle/cw_oracle.field_points, which is
 * checked against the reference-produced x64/a.cw continuity and closed-form TE stack.
 */
#include <complex.h>
#include <math.h>
#include <stddef.h>

static int region(float z, const float* zl, int layers, int tol_mode) {
  int l = 0;
  for (int li = 0; li < layers; ++li) {
    if (tol_mode == 0) {
      if ((double)z >= (double)zl[li] - 0.001) l = li + 1;
    } else {
      if (z >= zl[li]) l = li + 1;
    }
  }
  return l;
}

/* W: nwaves x {E0x, E0y, E0z, kx, ky, kz} complex float (interleaved re, im); out: [nz][ny][nx][3] */
void cw_eval_grid_f32(const float complex* W, const float* zl, const int* wb, const int* we, int layers,
                      int tol_mode, float x0, float y0, float z0, float dx, float dy, float dz, int nx, int ny,
                      int nz, float complex* out) {
  const long plane = (long)nx * ny;
#pragma omp parallel for schedule(dynamic, 64)
  for (long p = 0; p < plane * nz; ++p) {
    const int iz = (int)(p / plane);
    const long q = p % plane;
    const float x = fmaf(dx, (float)(q % nx), x0), y = fmaf(dy, (float)(q / nx), y0), z = fmaf(dz, (float)iz, z0);
    const int l = region(z, zl, layers, tol_mode);
    float complex E0 = 0, E1 = 0, E2 = 0;
    for (int w = wb[l]; w < we[l]; ++w) {
      const float complex* r = W + 6L * w;
      const float complex kdr = x * r[3] + y * r[4] + z * r[5];
      const float complex ph = cexpf(I * kdr);
      E0 += r[0] * ph;
      E1 += r[1] * ph;
      E2 += r[2] * ph;
    }
    out[3 * p + 0] = E0;
    out[3 * p + 1] = E1;
    out[3 * p + 2] = E2;
  }
}

void cw_eval_grid_f64(const double complex* W, const float* zl, const int* wb, const int* we, int layers,
                      int tol_mode, float x0, float y0, float z0, float dx, float dy, float dz, int nx, int ny,
                      int nz, double complex* out) {
  const long plane = (long)nx * ny;
#pragma omp parallel for schedule(dynamic, 64)
  for (long p = 0; p < plane * nz; ++p) {
    const int iz = (int)(p / plane);
    const long q = p % plane;
    const float xf = fmaf(dx, (float)(q % nx), x0), yf = fmaf(dy, (float)(q / nx), y0), zf = fmaf(dz, (float)iz, z0);
    const double x = xf, y = yf, z = zf;
    const int l = region(zf, zl, layers, tol_mode);
    double complex E0 = 0, E1 = 0, E2 = 0;
    for (int w = wb[l]; w < we[l]; ++w) {
      const double complex* r = W + 6L * w;
      const double complex kdr = x * r[3] + y * r[4] + z * r[5];
      const double complex ph = cexp(I * kdr);
      E0 += r[0] * ph;
      E1 += r[1] * ph;
      E2 += r[2] * ph;
    }
    out[3 * p + 0] = E0;
    out[3 * p + 1] = E1;
    out[3 * p + 2] = E2;
  }
}
