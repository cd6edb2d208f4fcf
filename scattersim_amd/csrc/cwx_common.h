// Shared definitions for the cwx HIP library: complex arithmetic, error handling, launch helpers.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>

namespace cwx {

// Complex double, binary compatible with std::complex<double> and the C-ABI cwx_complex.
struct zc {
  double re, im;
};

__host__ __device__ inline zc mk(double r, double i = 0.0) { return zc{r, i}; }
__host__ __device__ inline zc operator+(zc a, zc b) { return zc{a.re + b.re, a.im + b.im}; }
__host__ __device__ inline zc operator-(zc a, zc b) { return zc{a.re - b.re, a.im - b.im}; }
__host__ __device__ inline zc operator-(zc a) { return zc{-a.re, -a.im}; }
__host__ __device__ inline zc operator*(zc a, zc b) {
  return zc{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ inline zc operator*(double s, zc a) { return zc{s * a.re, s * a.im}; }
__host__ __device__ inline zc operator*(zc a, double s) { return zc{s * a.re, s * a.im}; }
__host__ __device__ inline zc& operator+=(zc& a, zc b) { a.re += b.re; a.im += b.im; return a; }
__host__ __device__ inline zc& operator-=(zc& a, zc b) { a.re -= b.re; a.im -= b.im; return a; }
__host__ __device__ inline zc conj(zc a) { return zc{a.re, -a.im}; }
__host__ __device__ inline double abs1(zc a) { return fabs(a.re) + fabs(a.im); }
__host__ __device__ inline double absz(zc a) { return hypot(a.re, a.im); }
__host__ __device__ inline double norm2(zc a) { return a.re * a.re + a.im * a.im; }
// a += b * c
__host__ __device__ inline void fma_acc(zc& a, zc b, zc c) {
  a.re = fma(b.re, c.re, a.re);
  a.re = fma(-b.im, c.im, a.re);
  a.im = fma(b.re, c.im, a.im);
  a.im = fma(b.im, c.re, a.im);
}
// Smith's algorithm for a / b.
__host__ __device__ inline zc zdiv(zc a, zc b) {
  if (fabs(b.re) >= fabs(b.im)) {
    if (b.re == 0.0 && b.im == 0.0) return zc{a.re / b.re, a.im / b.re};
    double r = b.im / b.re, den = b.re + b.im * r;
    return zc{(a.re + a.im * r) / den, (a.im - a.re * r) / den};
  }
  double r = b.re / b.im, den = b.im + b.re * r;
  return zc{(a.re * r + a.im) / den, (a.im * r - a.re) / den};
}
__host__ __device__ inline zc operator/(zc a, zc b) { return zdiv(a, b); }
// Principal square root (branch cut on the negative real axis, sign of a zero imaginary part kept).
__host__ __device__ inline zc zsqrt(zc a) {
  if (a.re == 0.0 && a.im == 0.0) return zc{0.0, a.im};
  double m = hypot(a.re, a.im);
  double t = sqrt(0.5 * (m + fabs(a.re)));
  if (a.re >= 0.0) return zc{t, 0.5 * a.im / t};
  return zc{0.5 * fabs(a.im) / t, copysign(t, a.im)};
}
__host__ __device__ inline zc zexp(zc a) {
  double e = exp(a.re);
  double s, c;
  sincos(a.im, &s, &c);
  return zc{e * c, e * s};
}

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define CWX_HIP(call)                                                                          \
  do {                                                                                         \
    hipError_t _e = (call);                                                                    \
    if (_e != hipSuccess)                                                                      \
      throw ::cwx::Error(-2, std::string(#call " failed: ") + hipGetErrorString(_e) + " at " + \
                                 __FILE__ + ":" + std::to_string(__LINE__));                   \
  } while (0)

#define CWX_CHECK(cond, msg)                                  \
  do {                                                        \
    if (!(cond)) throw ::cwx::Error(-1, std::string(msg));    \
  } while (0)

inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// LDS bytes one workgroup may allocate on the current device (160 KiB on gfx950), queried once per device.
inline size_t lds_limit() {
  static size_t cached[64] = {};
  int dev = 0;
  CWX_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return 64 * 1024;
  if (!cached[dev]) {
    int v = 0;
    CWX_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    cached[dev] = v > 0 ? (size_t)v : 64 * 1024;
  }
  return cached[dev];
}

// Device memory of the library's buffers goes through a per-device cache: a released block is kept and
// handed to the next request of the same rounded size instead of being returned with hipFree (which
// synchronises the device and unmaps, a few hundred microseconds per large block).  Blocks released since
// the last synchronisation are reused only after one device-wide synchronisation (the guarantee hipFree
// gave for each), taken at the next allocation; during a stream capture by this thread (capture_scope) no
// synchronisation is made and those blocks wait.  The cache holds at most a bounded number of bytes per
// device and is emptied when an allocation fails; cwx_set_param("alloc_cache", 0) empties and disables it.
void* dev_alloc(size_t bytes, int* dev);
void dev_free(void* p, size_t bytes, int dev);
struct capture_scope {  // marks this thread as capturing a stream for the lifetime of the object
  capture_scope();
  ~capture_scope();
};

// Device buffer owned by the library (released on destruction).
template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  int dev = 0;
  DevBuf() = default;
  explicit DevBuf(size_t count) { alloc(count); }
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n), dev(o.dev) { o.p = nullptr; o.n = 0; }
  DevBuf& operator=(DevBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; dev = o.dev; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DevBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count) p = static_cast<T*>(dev_alloc(count * sizeof(T), &dev));
    n = count;
  }
  void ensure(size_t count) { if (count > n) alloc(count); }
  void release() {
    if (p) dev_free(p, n * sizeof(T), dev);
    p = nullptr;
    n = 0;
  }
};

}  // namespace cwx
