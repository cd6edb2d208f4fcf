// extern "C" entry points of libcwx (declared in include/cwx.h).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <atomic>
#include <mutex>
#include <random>
#include <string>
#include <vector>

#include "../../include/cwx.h"
#include "cwx_layer.h"
#include "cwx_mlp.h"
#include "cwx_solver.h"
#include "cwx_tune.h"

struct cwx_context {
  cwx::Context c;
};
struct cwx_solution {
  cwx::Solution s;
};

namespace cwx {

// ---- device memory cache of the library's buffers (declared in cwx_common.h) ----
namespace {
constexpr int kCacheDevices = 64;
constexpr size_t kCacheMaxBytes = size_t(32) << 30;  // per device
struct DevCache {
  std::mutex m;
  std::multimap<size_t, void*> ready;           // reusable now, by rounded size
  std::vector<std::pair<void*, size_t>> pending;  // released since the last device synchronisation
  size_t bytes = 0;                             // held in ready and pending
};
// never destroyed, so that buffers released during the process's exit still find it
DevCache* const g_cache = new DevCache[kCacheDevices];
std::atomic<bool> g_cache_on{true};
thread_local int t_capturing = 0;

size_t cache_round(size_t b) {
  const size_t g = b < (size_t(1) << 20) ? 512 : (size_t(2) << 20);
  return (b + g - 1) / g * g;
}

// Returns every cached block of c to the runtime (c.m held).  Pending blocks only outside a capture.
void cache_flush_locked(DevCache& c) {
  for (auto& kv : c.ready) (void)hipFree(kv.second);
  c.ready.clear();
  if (!t_capturing) {
    for (auto& b : c.pending) (void)hipFree(b.first);
    c.pending.clear();
  }
  c.bytes = 0;
  for (auto& b : c.pending) c.bytes += b.second;
}
}  // namespace

capture_scope::capture_scope() { ++t_capturing; }
capture_scope::~capture_scope() { --t_capturing; }

void* dev_alloc(size_t bytes, int* dev) {
  CWX_HIP(hipGetDevice(dev));
  CWX_CHECK(*dev >= 0 && *dev < kCacheDevices, "device index out of range");
  const size_t rb = cache_round(bytes);
  DevCache& c = g_cache[*dev];
  std::lock_guard<std::mutex> lk(c.m);
  if (g_cache_on.load()) {
    if (!c.pending.empty() && !t_capturing) {
      // work queued before those releases may still use them: one synchronisation covers them all
      CWX_HIP(hipDeviceSynchronize());
      for (auto& b : c.pending) c.ready.emplace(b.second, b.first);
      c.pending.clear();
    }
    auto it = c.ready.find(rb);
    if (it != c.ready.end()) {
      void* p = it->second;
      c.ready.erase(it);
      c.bytes -= rb;
      return p;
    }
  }
  void* p = nullptr;
  if (hipMalloc(&p, rb) != hipSuccess) {
    (void)hipGetLastError();
    if (!t_capturing) CWX_HIP(hipDeviceSynchronize());
    cache_flush_locked(c);
    CWX_HIP(hipMalloc(&p, rb));
  }
  return p;
}

void dev_free(void* p, size_t bytes, int dev) {
  if (!p) return;
  const size_t rb = cache_round(bytes);
  if (dev < 0 || dev >= kCacheDevices) {
    (void)hipFree(p);
    return;
  }
  DevCache& c = g_cache[dev];
  std::lock_guard<std::mutex> lk(c.m);
  if ((g_cache_on.load() && c.bytes + rb <= kCacheMaxBytes) || t_capturing) {
    c.pending.emplace_back(p, rb);
    c.bytes += rb;
    return;
  }
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(dev);
  (void)hipFree(p);
  (void)hipSetDevice(cur);
}

// Empties every device's cache and turns it on or off (cwx_set_param "alloc_cache").
void dev_cache_set(bool on) {
  g_cache_on.store(on);
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (int d = 0; d < kCacheDevices; ++d) {
    DevCache& c = g_cache[d];
    std::lock_guard<std::mutex> lk(c.m);
    if (c.ready.empty() && c.pending.empty()) continue;
    if (hipSetDevice(d) != hipSuccess) continue;
    if (!t_capturing) (void)hipDeviceSynchronize();
    cache_flush_locked(c);
  }
  (void)hipSetDevice(cur);
}

}  // namespace cwx

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

template <typename F>
int guard(F&& f) {
  try {
    g_err.clear();
    return f();
  } catch (const cwx::Error& e) {
    return fail(e.code, e.what());
  } catch (const std::bad_alloc&) {
    return fail(CWX_ERR_DEVICE, "host allocation failed");
  } catch (const std::exception& e) {
    return fail(CWX_ERR_ARG, e.what());
  }
}

using cwx::zc;

void set_device(cwx_context* ctx) { CWX_HIP(hipSetDevice(ctx->c.device)); }

// Scratch of cwx_zgesv_batched_dev_noalloc: per-matrix status, then the pivots.  Every GPU holds its own
// copy of the array for the life of the process; host code reaches it only through the device address
// hipGetSymbolAddress returns for the current device (scratch_ptr), never through the array's name.
static __device__ __align__(256) unsigned char scratch[CWX_NOALLOC_SCRATCH_BYTES];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Called by cwx_create with `device` current: looks up this device's copy of the scratch array once.
void scratch_init(int device) {
  CWX_CHECK(device >= 0 && device < kMaxDevices, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void* p = nullptr;
  CWX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  CWX_CHECK(p != nullptr, "no device address for the scratch array");
  g_scratch_ptr[device] = p;
}

void* scratch_ptr(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  void* p = (device >= 0 && device < kMaxDevices) ? g_scratch_ptr[device] : nullptr;
  CWX_CHECK(p != nullptr, "the scratch array of this device was not set up (cwx_create)");
  return p;
}

// host -> device copy of a column-major (rows x cols) block with leading dimension ld, packed
void put_matrix(const cwx_complex* h, int rows, int cols, int ld, zc* d, hipStream_t s) {
  if (ld == rows) {
    CWX_HIP(hipMemcpyAsync(d, h, sizeof(zc) * (size_t)rows * cols, hipMemcpyHostToDevice, s));
  } else {
    CWX_HIP(hipMemcpy2DAsync(d, sizeof(zc) * rows, h, sizeof(zc) * ld, sizeof(zc) * rows, cols,
                             hipMemcpyHostToDevice, s));
  }
}
void get_matrix(const zc* d, int rows, int cols, cwx_complex* h, int ld, hipStream_t s) {
  if (ld == rows) {
    CWX_HIP(hipMemcpyAsync(h, d, sizeof(zc) * (size_t)rows * cols, hipMemcpyDeviceToHost, s));
  } else {
    CWX_HIP(hipMemcpy2DAsync(h, sizeof(zc) * ld, d, sizeof(zc) * rows, sizeof(zc) * rows, cols,
                             hipMemcpyDeviceToHost, s));
  }
}

cwx::VolumeOptions to_options(const cwx_volume_params* p) {
  CWX_CHECK(p != nullptr, "params must not be NULL");
  cwx::VolumeOptions o;
  o.lambda = p->lambda;
  for (int i = 0; i < 3; ++i) o.direction[i] = p->direction[i];
  for (int i = 0; i < 2; ++i) {
    o.ex[i] = p->ex[i];
    o.ey[i] = p->ey[i];
    o.ez[i] = p->ez[i];
    o.n[i] = p->n[i];
    o.coef[i] = p->coef[i];
  }
  o.kappa = p->kappa;
  for (int i = 0; i < 3; ++i) o.size[i] = p->size[i];
  o.z = p->z;
  CWX_CHECK(p->ncoef >= 0 && p->ncoef <= 2, "--coef takes 0, 1 or 2 values");
  o.ncoef = p->ncoef;
  o.psf = p->psf != 0;
  o.external = p->external != 0;
  o.npz = p->npz != 0;
  o.no_cutoff = p->no_cutoff != 0;
  CWX_CHECK(p->eig_mode == 0 || p->eig_mode == 1, "eig_mode must be 0 or 1");
  o.eig_mode = p->eig_mode;
  return o;
}

}  // namespace

extern "C" {

const char* cwx_last_error(void) { return g_err.c_str(); }

int cwx_set_param(const char* name, int64_t value) {
  return guard([&]() -> int {
    CWX_CHECK(name != nullptr, "name must not be NULL");
    const std::string k(name);
    const bool unset = value == CWX_PARAM_TABLE;
    if (k == "qr_team") {
      CWX_CHECK(unset || (value >= -1 && value <= 64), "qr_team takes -1 (built-in choice), 0 (off) or a helper count");
    } else if (k == "lahr2_team") {
      CWX_CHECK(unset || (value >= -1 && value <= 64), "lahr2_team takes -1 (built-in choice), 0 (off) or a helper count");
    } else if (k == "qr_threads") {
      CWX_CHECK(unset || value == -1 || value == 256 || value == 512, "qr_threads takes 256 or 512");
    } else if (k == "batch_chunk_mib") {
      CWX_CHECK(unset || value == -1 || value >= 64, "batch_chunk_mib takes at least 64");
    } else if (k == "chain_block") {
      CWX_CHECK(unset || (value >= -1 && value <= 1),
                "chain_block takes -1 (built-in choice), 0 (the order-4MF chain) or 1 (the reduced mode's block form)");
    } else if (k == "lu_panel") {
      CWX_CHECK(unset || (value >= -1 && value <= 2),
                "lu_panel takes -1 (built-in choice), 0 (rows per lane pair), 1 (a column pair per wavefront) or 2 "
                "(whole rows per lane)");
    } else if (k == "getrs_inv") {
      CWX_CHECK(unset || (value >= -1 && value <= 1),
                "getrs_inv takes -1 (built-in choice), 0 (substitution by diagonal blocks) or 1 (their inverses)");
    } else if (k == "alloc_cache") {
      CWX_CHECK(unset || (value >= -1 && value <= 1),
                "alloc_cache takes -1 or 1 (keep released device buffers for reuse, the default) or 0 (return "
                "them at once; the cache is emptied)");
      cwx::dev_cache_set(unset || value != 0);
      return CWX_OK;
    } else if (k == "zgemm_kernel") {
      CWX_CHECK(unset || (value >= -1 && value <= 3),
                "zgemm_kernel takes -1 (built-in choice), 0 (vector) or 1-3 (matrix cores, 64x64, 128x64, 64x128)");
    } else {
      return fail(CWX_ERR_ARG, "unknown parameter: " + k);
    }
    cwx::tune_override(k, (long)value, !unset && value != -1);
    return CWX_OK;
  });
}

int64_t cwx_get_param(const char* name, int64_t a, int64_t b) {
  if (!name) return -1;
  return cwx::tune_lookup(name, (long)a, (long)b, -1);
}

const char* cwx_tuning_table(void) {
  static thread_local std::string p;
  p = cwx::tune_table_path();
  return p.c_str();
}
const char* cwx_version(void) { return "cwx 0.1 (gfx950)"; }

int cwx_create(int device, cwx_context** ctx) {
  return guard([&]() -> int {
    CWX_CHECK(ctx != nullptr, "ctx must not be NULL");
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0) return fail(CWX_ERR_DEVICE, "no HIP device available");
    CWX_CHECK(device >= 0 && device < count, "device index out of range");
    auto c = std::make_unique<cwx_context>();
    c->c.device = device;
    CWX_HIP(hipSetDevice(device));
    scratch_init(device);
    CWX_HIP(hipStreamCreateWithFlags(&c->c.stream, hipStreamNonBlocking));
    c->c.own_stream = true;
    *ctx = c.release();
    return CWX_OK;
  });
}

int cwx_destroy(cwx_context* ctx) {
  return guard([&]() -> int {
    if (!ctx) return CWX_OK;
    (void)hipSetDevice(ctx->c.device);
    if (ctx->c.stream) (void)hipStreamSynchronize(ctx->c.stream);
    hipStream_t own = ctx->c.own_stream ? ctx->c.stream : nullptr;
    delete ctx;
    if (own) (void)hipStreamDestroy(own);
    return CWX_OK;
  });
}

int cwx_set_stream(cwx_context* ctx, void* hip_stream) {
  return guard([&]() -> int {
    CWX_CHECK(ctx != nullptr, "ctx must not be NULL");
    set_device(ctx);
    if (ctx->c.own_stream) CWX_HIP(hipStreamSynchronize(ctx->c.stream));
    if (hip_stream) {
      if (ctx->c.own_stream) CWX_HIP(hipStreamDestroy(ctx->c.stream));
      ctx->c.stream = (hipStream_t)hip_stream;
      ctx->c.own_stream = false;
    } else if (!ctx->c.own_stream) {
      CWX_HIP(hipStreamCreateWithFlags(&ctx->c.stream, hipStreamNonBlocking));
      ctx->c.own_stream = true;
    }
    return CWX_OK;
  });
}

int cwx_synchronize(cwx_context* ctx) {
  return guard([&]() -> int {
    CWX_CHECK(ctx != nullptr, "ctx must not be NULL");
    set_device(ctx);
    CWX_HIP(hipStreamSynchronize(ctx->c.stream));
    return CWX_OK;
  });
}

// ---------------------------------------------------------------------------------------------
int cwx_dev_alloc(cwx_context* ctx, int64_t bytes, void** p) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && p && bytes >= 0, "bad argument");
    set_device(ctx);
    *p = nullptr;
    if (bytes) CWX_HIP(hipMalloc(p, (size_t)bytes));
    return CWX_OK;
  });
}

int cwx_dev_free(cwx_context* ctx, void* p) {
  return guard([&]() -> int {
    CWX_CHECK(ctx != nullptr, "ctx must not be NULL");
    set_device(ctx);
    if (p) {
      CWX_HIP(hipStreamSynchronize(ctx->c.stream));
      CWX_HIP(hipFree(p));
    }
    return CWX_OK;
  });
}

int cwx_memcpy(cwx_context* ctx, void* dst, const void* src, int64_t bytes, int kind) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && dst && src && bytes >= 0, "bad argument");
    CWX_CHECK(kind >= 0 && kind <= 2, "kind must be 0 (host to device), 1 (device to host) or 2 (device to device)");
    set_device(ctx);
    const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : kind == 1 ? hipMemcpyDeviceToHost
                                                                          : hipMemcpyDeviceToDevice;
    CWX_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, k, ctx->c.stream));
    CWX_HIP(hipStreamSynchronize(ctx->c.stream));
    return CWX_OK;
  });
}

int cwx_zgeev(cwx_context* ctx, int n, cwx_complex* A, int lda, cwx_complex* w, cwx_complex* VR, int ldvr) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && A && w && VR, "NULL argument");
    CWX_CHECK(n >= 1 && lda >= n && ldvr >= n, "bad matrix dimensions");
    set_device(ctx);
    hipStream_t s = ctx->c.stream;
    cwx::DevBuf<zc> dA((size_t)n * n), dW(n), dV((size_t)n * n);
    cwx::DevBuf<int> dinfo(1);
    put_matrix(A, n, n, lda, dA.p, s);
    cwx::zgeev_batched(n, dA.p, n, (long)n * n, dW.p, n, dV.p, n, (long)n * n, 1, ctx->c.eig, dinfo.p, s);
    int info = 0;
    get_matrix(dA.p, n, n, A, lda, s);
    get_matrix(dV.p, n, n, VR, ldvr, s);
    CWX_HIP(hipMemcpyAsync(w, dW.p, sizeof(zc) * n, hipMemcpyDeviceToHost, s));
    CWX_HIP(hipMemcpyAsync(&info, dinfo.p, sizeof(int), hipMemcpyDeviceToHost, s));
    CWX_HIP(hipStreamSynchronize(s));
    if (info) return fail(info, "zgeev: QR iteration did not converge");
    return CWX_OK;
  });
}

int cwx_zgesv(cwx_context* ctx, int n, int nrhs, cwx_complex* A, int lda, int* ipiv, cwx_complex* B, int ldb) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && A && B, "NULL argument");
    CWX_CHECK(n >= 1 && nrhs >= 1 && lda >= n && ldb >= n, "bad matrix dimensions");
    set_device(ctx);
    hipStream_t s = ctx->c.stream;
    cwx::DevBuf<zc> dA((size_t)n * n), dB((size_t)n * nrhs);
    cwx::DevBuf<int> dP(n), dinfo(1);
    put_matrix(A, n, n, lda, dA.p, s);
    put_matrix(B, n, nrhs, ldb, dB.p, s);
    cwx::zgetrf_batched(n, dA.p, n, (long)n * n, dP.p, n, dinfo.p, 1, s);
    cwx::zgetrs_batched(n, nrhs, dA.p, n, (long)n * n, dP.p, n, dB.p, n, (long)n * nrhs, 1, s);
    get_matrix(dA.p, n, n, A, lda, s);
    get_matrix(dB.p, n, nrhs, B, ldb, s);
    std::vector<int> hp(n);
    int info = 0;
    CWX_HIP(hipMemcpyAsync(hp.data(), dP.p, sizeof(int) * n, hipMemcpyDeviceToHost, s));
    CWX_HIP(hipMemcpyAsync(&info, dinfo.p, sizeof(int), hipMemcpyDeviceToHost, s));
    CWX_HIP(hipStreamSynchronize(s));
    if (ipiv)
      for (int i = 0; i < n; ++i) ipiv[i] = hp[i] + 1;
    if (info) return fail(info, "zgesv: exactly singular matrix");
    return CWX_OK;
  });
}

int cwx_zinverse(cwx_context* ctx, int n, const cwx_complex* A, int lda, cwx_complex* Ainv, int ldinv) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && A && Ainv, "NULL argument");
    CWX_CHECK(n >= 1 && lda >= n && ldinv >= n, "bad matrix dimensions");
    std::vector<cwx_complex> a((size_t)n * n), b((size_t)n * n);
    for (int c = 0; c < n; ++c)
      for (int r = 0; r < n; ++r) {
        a[(size_t)c * n + r] = A[(size_t)c * lda + r];
        b[(size_t)c * n + r] = cwx_complex{r == c ? 1.0 : 0.0, 0.0};
      }
    int st = cwx_zgesv(ctx, n, n, a.data(), n, nullptr, b.data(), n);
    if (st != CWX_OK) return st;
    for (int c = 0; c < n; ++c)
      for (int r = 0; r < n; ++r) Ainv[(size_t)c * ldinv + r] = b[(size_t)c * n + r];
    return CWX_OK;
  });
}

int cwx_zgemm(cwx_context* ctx, int m, int n, int k, cwx_complex alpha, const cwx_complex* A, int lda,
              const cwx_complex* B, int ldb, cwx_complex beta, cwx_complex* C, int ldc) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && A && B && C, "NULL argument");
    CWX_CHECK(m >= 1 && n >= 1 && k >= 1 && lda >= m && ldb >= k && ldc >= m, "bad matrix dimensions");
    set_device(ctx);
    hipStream_t s = ctx->c.stream;
    cwx::DevBuf<zc> dA((size_t)m * k), dB((size_t)k * n), dC((size_t)m * n);
    put_matrix(A, m, k, lda, dA.p, s);
    put_matrix(B, k, n, ldb, dB.p, s);
    put_matrix(C, m, n, ldc, dC.p, s);
    cwx::zgemm_batched(cwx::OpN, cwx::OpN, m, n, k, zc{alpha.re, alpha.im}, dA.p, m, 0, dB.p, k, 0,
                       zc{beta.re, beta.im}, dC.p, m, 0, 1, s);
    get_matrix(dC.p, m, n, C, ldc, s);
    CWX_HIP(hipStreamSynchronize(s));
    return CWX_OK;
  });
}

int cwx_zgeev_batched_dev(cwx_context* ctx, int n, int batch, void* dA, int64_t strideA, void* dW,
                          int64_t strideW, void* dV, int64_t strideV, int* h_info) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && dA && dW && dV, "NULL argument");
    CWX_CHECK(n >= 1 && batch >= 1 && strideA >= (int64_t)n * n && strideV >= (int64_t)n * n && strideW >= n,
              "bad dimensions");
    set_device(ctx);
    hipStream_t s = ctx->c.stream;
    ctx->c.info.ensure(batch);
    cwx::zgeev_batched(n, (zc*)dA, n, strideA, (zc*)dW, strideW, (zc*)dV, n, strideV, batch, ctx->c.eig,
                       ctx->c.info.p, s);
    if (h_info) {
      CWX_HIP(hipMemcpyAsync(h_info, ctx->c.info.p, sizeof(int) * batch, hipMemcpyDeviceToHost, s));
      CWX_HIP(hipStreamSynchronize(s));
      for (int b = 0; b < batch; ++b)
        if (h_info[b]) return fail(h_info[b], "zgeev: QR iteration did not converge");
    }
    return CWX_OK;
  });
}

int cwx_zgetrf_batched_dev(cwx_context* ctx, int n, int batch, void* dA, int64_t strideA, void* dIpiv,
                           int64_t strideP, int* h_info) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && dA && dIpiv, "NULL argument");
    CWX_CHECK(n >= 1 && batch >= 1 && strideA >= (int64_t)n * n && strideP >= n, "bad dimensions");
    set_device(ctx);
    hipStream_t s = ctx->c.stream;
    ctx->c.info.ensure(batch);
    cwx::zgetrf_batched(n, (zc*)dA, n, strideA, (int*)dIpiv, strideP, ctx->c.info.p, batch, s);
    if (h_info) {
      CWX_HIP(hipMemcpyAsync(h_info, ctx->c.info.p, sizeof(int) * batch, hipMemcpyDeviceToHost, s));
      CWX_HIP(hipStreamSynchronize(s));
      for (int b = 0; b < batch; ++b)
        if (h_info[b]) return fail(h_info[b], "zgetrf: exactly singular matrix");
    }
    return CWX_OK;
  });
}

int cwx_zgetrs_batched_dev(cwx_context* ctx, int n, int nrhs, int batch, const void* dLU, int64_t strideA,
                           const void* dIpiv, int64_t strideP, void* dB, int64_t strideB) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && dLU && dIpiv && dB, "NULL argument");
    CWX_CHECK(n >= 1 && nrhs >= 1 && batch >= 1, "bad dimensions");
    set_device(ctx);
    cwx::zgetrs_batched(n, nrhs, (const zc*)dLU, n, strideA, (const int*)dIpiv, strideP, (zc*)dB, n, strideB,
                        batch, ctx->c.stream);
    return CWX_OK;
  });
}

int cwx_zgesv_batched_dev_noalloc(cwx_context* ctx, int n, int nrhs, int batch, void* dA, int64_t strideA, void* dB,
                                  int64_t strideB, int* h_info) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && dA && dB, "NULL argument");
    CWX_CHECK(n >= 1 && nrhs >= 1 && batch >= 1 && batch <= 65535, "bad dimensions");
    CWX_CHECK(strideA >= (int64_t)n * n && strideB >= (int64_t)n * nrhs, "strides smaller than the matrices");
    if ((long long)n * batch > (long long)CWX_NOALLOC_MAX_PIVOTS)
      return fail(CWX_ERR_ARG, "zgesv_noalloc: n * batch = " + std::to_string((long long)n * batch) +
                                   " pivots exceed the scratch limit of " + std::to_string(CWX_NOALLOC_MAX_PIVOTS));
    // status words first (room for the largest batch), then the pivots
    const size_t info_bytes = (size_t)65536 * sizeof(int);
    static_assert((size_t)65536 * sizeof(int) + (size_t)CWX_NOALLOC_MAX_PIVOTS * sizeof(int) <=
                      (size_t)CWX_NOALLOC_SCRATCH_BYTES, "scratch layout");
    set_device(ctx);
    unsigned char* base = static_cast<unsigned char*>(scratch_ptr(ctx->c.device));
    int* info = reinterpret_cast<int*>(base);
    int* piv = reinterpret_cast<int*>(base + info_bytes);
    hipStream_t s = ctx->c.stream;
    // zgetrf_batched clears info itself (a fill kernel), so the last call's contents do not leak in
    cwx::zgetrf_batched(n, (zc*)dA, n, strideA, piv, n, info, batch, s);
    cwx::zgetrs_batched(n, nrhs, (const zc*)dA, n, strideA, piv, n, (zc*)dB, n, strideB, batch, s);
    CWX_HIP(hipGetLastError());
    if (h_info) {
      CWX_HIP(hipMemcpyAsync(h_info, info, sizeof(int) * batch, hipMemcpyDeviceToHost, s));
      CWX_HIP(hipStreamSynchronize(s));
      for (int b = 0; b < batch; ++b)
        if (h_info[b]) return fail(h_info[b], "zgesv_noalloc: exactly singular matrix");
    }
    return CWX_OK;
  });
}

int cwx_zgemm_batched_dev(cwx_context* ctx, int transa, int transb, int m, int n, int k, cwx_complex alpha,
                          const void* dA, int lda, int64_t strideA, const void* dB, int ldb, int64_t strideB,
                          cwx_complex beta, void* dC, int ldc, int64_t strideC, int batch) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && dA && dB && dC, "NULL argument");
    CWX_CHECK(m >= 1 && n >= 1 && k >= 1 && batch >= 1, "bad dimensions");
    CWX_CHECK(transa >= 0 && transa <= 2 && transb >= 0 && transb <= 2, "trans must be 0 (N), 1 (T) or 2 (C)");
    set_device(ctx);
    cwx::zgemm_batched((cwx::Op)transa, (cwx::Op)transb, m, n, k, zc{alpha.re, alpha.im}, (const zc*)dA, lda,
                       strideA, (const zc*)dB, ldb, strideB, zc{beta.re, beta.im}, (zc*)dC, ldc, strideC, batch,
                       ctx->c.stream);
    return CWX_OK;
  });
}

int cwx_zgemm_split32_batched_dev(cwx_context* ctx, int transa, int transb, int m, int n, int k, cwx_complex alpha,
                                  const void* dA, int lda, int64_t strideA, const void* dB, int ldb, int64_t strideB,
                                  cwx_complex beta, void* dC, int ldc, int64_t strideC, int batch, int terms) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && dA && dB && dC, "NULL argument");
    CWX_CHECK(m >= 1 && n >= 1 && k >= 1 && batch >= 1, "bad dimensions");
    CWX_CHECK(transa >= 0 && transa <= 2 && transb >= 0 && transb <= 2, "trans must be 0 (N), 1 (T) or 2 (C)");
    CWX_CHECK(terms == 1 || terms == 3, "terms must be 1 or 3");
    set_device(ctx);
    cwx::zgemm_split32_set((cwx::Op)transa, (cwx::Op)transb, m, n, k, zc{alpha.re, alpha.im},
                           cwx::strided((const zc*)dA, strideA), lda, cwx::strided((const zc*)dB, strideB), ldb,
                           zc{beta.re, beta.im}, cwx::strided((zc*)dC, strideC), ldc, batch, terms == 3,
                           ctx->c.stream);
    return CWX_OK;
  });
}

// ---------------------------------------------------------------------------------------------
void cwx_volume_params_default(cwx_volume_params* p) {
  if (!p) return;
  std::memset(p, 0, sizeof(*p));
  p->lambda = 1.0;
  p->direction[2] = 1.0;
  p->ey[0] = 1.0;
  p->n[0] = p->n[1] = 1.0;
  p->size[0] = p->size[1] = 10.0;
  p->size[2] = 2.0;
}

static int solve_common(cwx_context* ctx, const cwx_volume_params* p, const zc* d_sample, int L, int s1, int s2,
                        cwx_solution** out) {
  auto sol = std::make_unique<cwx_solution>();
  cwx::solve_volume(ctx->c, to_options(p), d_sample, L, s1, s2, sol->s);
  const int info = sol->s.info;
  const std::string msg = sol->s.info_msg;
  *out = sol.release();
  if (info) return fail(info, msg);
  return CWX_OK;
}

int cwx_solve_volume(cwx_context* ctx, const cwx_volume_params* p, const cwx_complex* sample, int L, int s1,
                     int s2, cwx_solution** out) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && p && sample && out, "NULL argument");
    CWX_CHECK(L >= 1 && s1 >= 1 && s2 >= 1, "sample dimensions must be positive");
    set_device(ctx);
    const size_t n = (size_t)L * s1 * s2;
    cwx::DevBuf<zc> d(n);
    CWX_HIP(hipMemcpyAsync(d.p, sample, sizeof(zc) * n, hipMemcpyHostToDevice, ctx->c.stream));
    return solve_common(ctx, p, d.p, L, s1, s2, out);
  });
}

int cwx_solve_volume_dev(cwx_context* ctx, const cwx_volume_params* p, const void* d_sample, int L, int s1,
                         int s2, cwx_solution** out) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && p && d_sample && out, "NULL argument");
    CWX_CHECK(L >= 1 && s1 >= 1 && s2 >= 1, "sample dimensions must be positive");
    set_device(ctx);
    return solve_common(ctx, p, (const zc*)d_sample, L, s1, s2, out);
  });
}

int cwx_solve_batch_dev(cwx_context* ctx, int count, const cwx_volume_params* params, const int* L,
                        const void* d_samples, int s1, int s2, int64_t max_bytes, void* d_x, int* h_info, int* M0,
                        int* M1) {
  return guard([&]() -> int {
    CWX_CHECK(params && L, "NULL argument");
    CWX_CHECK(count >= 1, "count must be positive");
    CWX_CHECK(s1 >= 1 && s2 >= 1, "sample dimensions must be positive");
    CWX_CHECK(max_bytes >= 0, "max_bytes must be non-negative");
    std::vector<cwx::VolumeOptions> opts(count);
    std::vector<int> Ls(L, L + count);
    for (int i = 0; i < count; ++i) {
      CWX_CHECK(Ls[i] >= 1, "instance " + std::to_string(i) + " has no layers");
      opts[i] = to_options(params + i);
    }
    const cwx::Setup st = cwx::make_setup(opts[0], Ls[0], s1, s2);
    if (M0) *M0 = st.M0;
    if (M1) *M1 = st.M1;
    if (!d_x) return CWX_OK;
    CWX_CHECK(ctx && d_samples, "NULL argument");
    set_device(ctx);
    std::vector<int> info;
    std::string msg;
    const int st_code = cwx::solve_batch(ctx->c, opts, (const zc*)d_samples, Ls, s1, s2, (size_t)max_bytes,
                                         (zc*)d_x, info, msg);
    if (h_info) std::copy(info.begin(), info.end(), h_info);
    if (st_code) return fail(st_code, msg);
    return CWX_OK;
  });
}

int cwx_solve_batch_graph_dev(cwx_context* ctx, int count, const cwx_volume_params* params, const int* L,
                              const void* d_samples, int s1, int s2, int chunk, void* d_x, int* h_info,
                              int* captured) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && params && L && d_samples && d_x, "NULL argument");
    CWX_CHECK(count >= 1 && s1 >= 1 && s2 >= 1 && chunk >= 1, "bad batch arguments");
    std::vector<cwx::VolumeOptions> opts(count);
    std::vector<int> Ls(L, L + count);
    for (int i = 0; i < count; ++i) {
      CWX_CHECK(Ls[i] >= 1, "instance " + std::to_string(i) + " has no layers");
      opts[i] = to_options(params + i);
    }
    set_device(ctx);
    std::vector<int> info;
    std::string msg;
    const int st_code = cwx::solve_batch_graph(ctx->c, opts, (const zc*)d_samples, Ls, s1, s2, chunk, (zc*)d_x,
                                               info, msg, captured);
    if (h_info) std::copy(info.begin(), info.end(), h_info);
    if (st_code) return fail(st_code, msg);
    return CWX_OK;
  });
}

int cwx_solve_batch_mixed_dev(cwx_context* ctx, int count, const cwx_volume_params* params, const int* L,
                              const void* d_samples, int s1, int s2, int64_t max_bytes, void* d_x, int64_t* x_offsets,
                              int* h_info, int* M0s, int* M1s) {
  return guard([&]() -> int {
    CWX_CHECK(params && L && x_offsets, "NULL argument");
    CWX_CHECK(count >= 1 && s1 >= 1 && s2 >= 1 && max_bytes >= 0, "bad batch arguments");
    std::vector<cwx::VolumeOptions> opts(count);
    std::vector<int> Ls(L, L + count);
    std::vector<long> off(count + 1, 0);
    for (int i = 0; i < count; ++i) {
      CWX_CHECK(Ls[i] >= 1, "instance " + std::to_string(i) + " has no layers");
      opts[i] = to_options(params + i);
      const cwx::Setup st = cwx::make_setup(opts[i], Ls[i], s1, s2);
      if (M0s) M0s[i] = st.M0;
      if (M1s) M1s[i] = st.M1;
      off[i + 1] = off[i] + 6L * st.MF;
    }
    for (int i = 0; i <= count; ++i) x_offsets[i] = off[i];
    if (!d_x) return CWX_OK;
    CWX_CHECK(ctx && d_samples, "NULL argument");
    set_device(ctx);
    std::vector<int> info;
    std::string msg;
    const int code = cwx::solve_batch_mixed(ctx->c, opts, (const zc*)d_samples, Ls, s1, s2, (size_t)max_bytes,
                                            (zc*)d_x, off, info, msg);
    if (h_info) std::copy(info.begin(), info.end(), h_info);
    if (code) return fail(code, msg);
    return CWX_OK;
  });
}

int cwx_solve_fields_dev(cwx_context* ctx, const cwx_volume_params* p, const void* d_sample, int L, int s1, int s2,
                         int K, const void* d_EF, void* d_x, int* M0, int* M1) {
  return guard([&]() -> int {
    CWX_CHECK(p, "NULL argument");
    CWX_CHECK(L >= 1 && s1 >= 1 && s2 >= 1, "sample dimensions must be positive");
    const cwx::VolumeOptions o = to_options(p);
    const cwx::Setup st = cwx::make_setup(o, L, s1, s2);
    if (M0) *M0 = st.M0;
    if (M1) *M1 = st.M1;
    if (!d_x) return CWX_OK;
    CWX_CHECK(ctx && d_sample && d_EF, "NULL argument");
    CWX_CHECK(K >= 1, "K must be positive");
    set_device(ctx);
    std::string msg;
    const int code = cwx::solve_fields(ctx->c, o, (const zc*)d_sample, L, s1, s2, K, (const zc*)d_EF, (zc*)d_x, msg);
    if (code) return fail(code, msg);
    return CWX_OK;
  });
}

int cwx_phi_build_dev(cwx_context* ctx, const cwx_volume_params* p, const void* d_sample, int L, int s1, int s2,
                      int blocks, void* d_out, int* n_out) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && p && n_out, "NULL argument");
    CWX_CHECK(L >= 1 && s1 >= 1 && s2 >= 1, "sample dimensions must be positive");
    CWX_CHECK(blocks == 0 || blocks == 1, "blocks must be 0 (Phi) or 1 (P and Q)");
    const cwx::VolumeOptions o = to_options(p);
    const cwx::Setup st = cwx::make_setup(o, L, s1, s2);
    *n_out = blocks ? 2 * st.MF : 4 * st.MF;
    if (!d_out) return CWX_OK;
    CWX_CHECK(d_sample != nullptr, "NULL sample");
    set_device(ctx);
    cwx::build_phi_dev(o, (const zc*)d_sample, L, s1, s2, blocks, (zc*)d_out, ctx->c.stream);
    return CWX_OK;
  });
}

int cwx_solution_get_info(const cwx_solution* sol, cwx_solution_info* info) {
  return guard([&]() -> int {
    CWX_CHECK(sol && info, "NULL argument");
    const auto& s = sol->s;
    info->M0 = s.st.M0;
    info->M1 = s.st.M1;
    info->MF = s.st.MF;
    info->L = s.st.L;
    info->cutoff = s.st.cutoff ? 1 : 0;
    info->info = s.info;
    info->k = s.st.k;
    info->nrec = s.nrec;
    info->nlists = (int)s.seg_cnt.size();
    return CWX_OK;
  });
}

int64_t cwx_solution_size(const cwx_solution* sol, int what) {
  if (!sol) return -1;
  const auto& s = sol->s;
  const int64_t MF = s.st.MF, L = s.st.L, n4 = 4 * MF;
  switch (what) {
    case CWX_SOL_X: return 6 * MF;
    case CWX_SOL_GAMMA: return L * n4;
    case CWX_SOL_BETA: return s.beta.p ? L * n4 : 0;
    case CWX_SOL_QCH: return s.Qch.p ? L * 3 * MF * n4 : 0;
    case CWX_SOL_Z: return L + 1;
    case CWX_SOL_RECORDS: return s.nrec * 6;
    case CWX_SOL_SX: case CWX_SOL_SY: case CWX_SOL_SZ0: case CWX_SOL_SZ1: return MF;
    case CWX_SOL_EF: return 3 * MF;
    case CWX_SOL_GD: case CWX_SOL_GC: return L * n4 * n4;
    case CWX_SOL_NIF: return L * MF;
    default: return -1;
  }
}

const void* cwx_solution_device_ptr(const cwx_solution* sol, int what) {
  if (!sol) return nullptr;
  const auto& s = sol->s;
  switch (what) {
    case CWX_SOL_X: return s.x.p;
    case CWX_SOL_GAMMA: return s.gam.p;
    case CWX_SOL_BETA: return s.beta.p;
    case CWX_SOL_QCH: return s.Qch.p;
    case CWX_SOL_RECORDS: return s.rec.p;
    case CWX_SOL_GD: return s.Gd.p;
    case CWX_SOL_GC: return s.Gc.p;
    case CWX_SOL_NIF: return s.NIF.p;
    default: return nullptr;
  }
}

int cwx_solution_copy(cwx_context* ctx, const cwx_solution* sol, int what, void* host_dst) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && sol && host_dst, "NULL argument");
    const int64_t n = cwx_solution_size(sol, what);
    CWX_CHECK(n >= 0, "unknown solution array");
    const auto& s = sol->s;
    auto hostcpy = [&](const std::vector<zc>& v) {
      std::memcpy(host_dst, v.data(), sizeof(zc) * v.size());
      return CWX_OK;
    };
    switch (what) {
      case CWX_SOL_Z: std::memcpy(host_dst, s.st.z.data(), sizeof(double) * s.st.z.size()); return CWX_OK;
      case CWX_SOL_SX: return hostcpy(s.st.Sx);
      case CWX_SOL_SY: return hostcpy(s.st.Sy);
      case CWX_SOL_SZ0: return hostcpy(s.st.Sz0);
      case CWX_SOL_SZ1: return hostcpy(s.st.Sz1);
      case CWX_SOL_EF: return hostcpy(s.st.EF);
      default: break;
    }
    const void* d = cwx_solution_device_ptr(sol, what);
    if (n == 0) return CWX_OK;
    CWX_CHECK(d != nullptr, "solution array not available (--external skips beta and Q)");
    set_device(ctx);
    CWX_HIP(hipMemcpyAsync(host_dst, d, sizeof(zc) * n, hipMemcpyDeviceToHost, ctx->c.stream));
    CWX_HIP(hipStreamSynchronize(ctx->c.stream));
    return CWX_OK;
  });
}

int cwx_solution_lists(const cwx_solution* sol, int64_t* offsets, int64_t* counts) {
  return guard([&]() -> int {
    CWX_CHECK(sol && offsets && counts, "NULL argument");
    for (size_t i = 0; i < sol->s.seg_cnt.size(); ++i) {
      offsets[i] = sol->s.seg_off[i];
      counts[i] = sol->s.seg_cnt[i];
    }
    return CWX_OK;
  });
}

int cwx_solution_write_cw(cwx_context* ctx, const cwx_solution* sol, const char* path) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && sol && path, "NULL argument");
    set_device(ctx);
    try {
      cwx::cw_write_file(sol->s, path, ctx->c.stream);
    } catch (const cwx::Error& e) {
      if (e.code == -1) return fail(CWX_ERR_IO, e.what());
      throw;
    }
    return CWX_OK;
  });
}

int cwx_solution_write_cw2(cwx_context* ctx, const cwx_solution* sol, const char* path) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && sol && path, "NULL argument");
    if (sol->s.opt.external)
      return fail(CWX_ERR_ARG, "an --external solution has no interior mode amplitudes (beta); "
                               "cwx_solution_write_cw already writes only its external waves");
    set_device(ctx);
    try {
      cwx::cw2_write_file(sol->s, path, ctx->c.stream);
    } catch (const cwx::Error& e) {
      if (e.code == -1) return fail(CWX_ERR_IO, e.what());
      throw;
    }
    return CWX_OK;
  });
}

int cwx_solution_read_cw2(cwx_context* ctx, const char* path, cwx_solution** out) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && path && out, "NULL argument");
    *out = nullptr;
    set_device(ctx);
    std::unique_ptr<cwx_solution> sol(new cwx_solution);
    try {
      cwx::cw2_read_file(ctx->c, path, sol->s);
    } catch (const cwx::Error& e) {
      if (e.code == -1) return fail(CWX_ERR_IO, e.what());
      throw;
    }
    *out = sol.release();
    return CWX_OK;
  });
}

int cwx_solution_free(cwx_solution* sol) {
  return guard([&]() -> int {
    delete sol;
    return CWX_OK;
  });
}

// ---------------------------------------------------------------------------------------------
namespace {

// the stack as the device kernels take it: L complex indices (the first real) and L - 1 boundaries
void layer_tables(int L, const double* n, const double* kappa, const double* z, cwx::DevBuf<zc>& ri,
                  cwx::DevBuf<double>& dz, hipStream_t s) {
  CWX_CHECK(L >= 2 && n && z && (L == 1 || kappa), "a stack needs n[L], kappa[L-1] and z[L-1] with L >= 2");
  std::vector<zc> h(L);
  h[0] = cwx::mk(n[0], 0.0);
  for (int l = 1; l < L; ++l) h[l] = cwx::mk(n[l], kappa[l - 1]);
  for (int l = 1; l + 1 < L; ++l)
    CWX_CHECK(z[l] >= z[l - 1], "layer boundaries must be in increasing z");
  ri.alloc(L);
  dz.alloc(L - 1);
  CWX_HIP(hipMemcpyAsync(ri.p, h.data(), sizeof(zc) * L, hipMemcpyHostToDevice, s));
  CWX_HIP(hipMemcpyAsync(dz.p, z, sizeof(double) * (L - 1), hipMemcpyHostToDevice, s));
}

}  // namespace

int cwx_layer_solve_dev(cwx_context* ctx, int L, const double* n, const double* kappa, const double* z,
                        double lambda, int64_t P, const cwx_wave_f64* d_incident, cwx_wave_f64* d_reflected,
                        cwx_wave_f64* d_transmitted) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && (P == 0 || (d_incident && d_reflected && d_transmitted)), "NULL argument");
    set_device(ctx);
    hipStream_t s = ctx->c.stream;
    cwx::DevBuf<zc> ri;
    cwx::DevBuf<double> dz;
    layer_tables(L, n, kappa, z, ri, dz, s);
    const int ns = cwx::layer_solve_dev(L, ri.p, dz.p, lambda, P, reinterpret_cast<const zc*>(d_incident),
                                        reinterpret_cast<zc*>(d_reflected), reinterpret_cast<zc*>(d_transmitted), s);
    if (ns) return fail(ns, std::to_string(ns) + " incident wave(s) gave an exactly singular system");
    return CWX_OK;
  });
}

int cwx_layer_solve(cwx_context* ctx, int L, const double* n, const double* kappa, const double* z, double lambda,
                    int64_t P, const cwx_wave_f64* incident, cwx_wave_f64* reflected, cwx_wave_f64* transmitted) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && (P == 0 || (incident && reflected && transmitted)), "NULL argument");
    CWX_CHECK(L >= 2, "a stack needs at least two layers");
    set_device(ctx);
    hipStream_t s = ctx->c.stream;
    const size_t nout = (size_t)P * (L - 1);
    cwx::DevBuf<zc> di((size_t)P * 6), dr(nout * 6), dt(nout * 6);
    if (P) CWX_HIP(hipMemcpyAsync(di.p, incident, sizeof(cwx_wave_f64) * P, hipMemcpyHostToDevice, s));
    const int st = cwx_layer_solve_dev(ctx, L, n, kappa, z, lambda, P, reinterpret_cast<cwx_wave_f64*>(di.p),
                                       reinterpret_cast<cwx_wave_f64*>(dr.p), reinterpret_cast<cwx_wave_f64*>(dt.p));
    if (st < 0) return st;
    const std::string msg = g_err;
    if (P) {
      CWX_HIP(hipMemcpyAsync(reflected, dr.p, sizeof(cwx_wave_f64) * nout, hipMemcpyDeviceToHost, s));
      CWX_HIP(hipMemcpyAsync(transmitted, dt.p, sizeof(cwx_wave_f64) * nout, hipMemcpyDeviceToHost, s));
    }
    CWX_HIP(hipStreamSynchronize(s));
    if (st > 0) return fail(st, msg);
    return CWX_OK;
  });
}

int cwx_layer_write_cw(const char* path, int L, const double* z, int64_t P, const cwx_wave_f64* incident,
                       const cwx_wave_f64* reflected, const cwx_wave_f64* transmitted) {
  return guard([&]() -> int {
    CWX_CHECK(path && z && (P == 0 || (incident && reflected && transmitted)) && L >= 2, "NULL argument");
    std::ofstream f(path, std::ios::binary);
    if (!f) return fail(CWX_ERR_IO, std::string("cannot open ") + path + " for writing");
    auto put = [&](const void* p, size_t nb) { f.write(static_cast<const char*>(p), (std::streamsize)nb); };
    auto put_u64 = [&](uint64_t v) { put(&v, 8); };
    const unsigned char is_volume = 0;
    put(&is_volume, 1);
    put_u64(8);
    put_u64((uint64_t)P);
    put(incident, sizeof(cwx_wave_f64) * (size_t)P);
    put_u64((uint64_t)(L - 1));
    for (int l = 0; l < L - 1; ++l) {
      put(&z[l], 8);
      put_u64((uint64_t)P);
      for (int64_t p = 0; p < P; ++p) put(&reflected[(size_t)p * (L - 1) + l], sizeof(cwx_wave_f64));
      put_u64((uint64_t)P);
      for (int64_t p = 0; p < P; ++p) put(&transmitted[(size_t)p * (L - 1) + l], sizeof(cwx_wave_f64));
    }
    if (!f) return fail(CWX_ERR_IO, std::string("write to ") + path + " failed");
    return CWX_OK;
  });
}

// ---------------------------------------------------------------------------------------------
namespace {

// unit vector d and the rotation that takes +z to it (Rodrigues), applied to v
void rotate_from_z(const double d[3], const double v[3], double out[3]) {
  const double c = d[2];
  if (c > 1.0 - 1e-15) {
    for (int i = 0; i < 3; ++i) out[i] = v[i];
    return;
  }
  if (c < -1.0 + 1e-15) {  // half turn about x
    out[0] = v[0];
    out[1] = -v[1];
    out[2] = -v[2];
    return;
  }
  // axis a = z x d / |z x d|, angle acos(c)
  const double ax = -d[1], ay = d[0], sn = std::sqrt(ax * ax + ay * ay);
  const double ux = ax / sn, uy = ay / sn;  // uz = 0
  const double dot = ux * v[0] + uy * v[1];
  const double cr[3] = {uy * v[2], -ux * v[2], ux * v[1] - uy * v[0]};
  for (int i = 0; i < 3; ++i) {
    const double ui = i == 0 ? ux : (i == 1 ? uy : 0.0);
    out[i] = v[i] * c + cr[i] * sn + ui * dot * (1.0 - c);
  }
}

// one sample of a focused beam: direction (theta, phi) about the beam axis, amplitude weight w
void beam_wave(const double k[3], double kmag, const cwx_complex E[3], double theta, double phi, double w,
               cwx_wave_f64& out) {
  const double local[3] = {std::sin(theta) * std::cos(phi), std::sin(theta) * std::sin(phi), std::cos(theta)};
  const double axis[3] = {k[0] / kmag, k[1] / kmag, k[2] / kmag};
  double d[3];
  rotate_from_z(axis, local, d);
  const double kk[3] = {d[0] * kmag, d[1] * kmag, d[2] * kmag};
  const cwx_complex ex{w * E[0].re, w * E[0].im}, ey{w * E[1].re, w * E[1].im};
  cwx_complex ez{w * E[2].re, w * E[2].im};
  if (kk[2] != 0.0)  // transverse: Ez from k . E0 = 0
    ez = cwx_complex{-(kk[0] * ex.re + kk[1] * ey.re) / kk[2], -(kk[0] * ex.im + kk[1] * ey.im) / kk[2]};
  out.E0[0] = ex;
  out.E0[1] = ey;
  out.E0[2] = ez;
  for (int c = 0; c < 3; ++c) out.k[c] = cwx_complex{kk[c], 0.0};
}

}  // namespace

int cwx_beam_polar(double alpha, double beta, const double k[3], const cwx_complex E[3], int N0, int N1,
                   cwx_wave_f64* out) {
  return guard([&]() -> int {
    CWX_CHECK(k && E && out, "NULL argument");
    CWX_CHECK(N0 >= 1 && N1 >= 1, "the beam needs at least one sample per angle");
    CWX_CHECK(alpha > beta && beta >= 0.0, "need 0 <= beta < alpha");
    const double kmag = std::sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
    CWX_CHECK(kmag > 0.0, "zero wave vector");
    const double dth = (alpha - beta) / N0, dph = 2.0 * M_PI / N1;
    for (int j = 0; j < N1; ++j)
      for (int i = 0; i < N0; ++i) {
        const double th = beta + i * dth, ph = j * dph;
        beam_wave(k, kmag, E, th, ph, std::sin(th) * dth * dph, out[(size_t)j * N0 + i]);
      }
    return CWX_OK;
  });
}

int cwx_beam_montecarlo(double alpha, double beta, const double k[3], const cwx_complex E[3], int64_t N,
                        uint64_t seed, cwx_wave_f64* out) {
  return guard([&]() -> int {
    CWX_CHECK(k && E && (N == 0 || out), "NULL argument");
    CWX_CHECK(N >= 0 && alpha > beta && beta >= 0.0, "need N >= 0 and 0 <= beta < alpha");
    const double kmag = std::sqrt(k[0] * k[0] + k[1] * k[1] + k[2] * k[2]);
    CWX_CHECK(kmag > 0.0, "zero wave vector");
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(0.0, 1.0);
    const double ca = std::cos(alpha), cb = std::cos(beta);
    const double w = 2.0 * M_PI * (cb - ca) / (double)std::max<int64_t>(N, 1);  // cap solid angle / N
    for (int64_t i = 0; i < N; ++i) {
      const double ct = cb - u(rng) * (cb - ca), ph = 2.0 * M_PI * u(rng);
      beam_wave(k, kmag, E, std::acos(ct), ph, w, out[i]);
    }
    return CWX_OK;
  });
}

// ---------------------------------------------------------------------------------------------
int cwx_bf16_gemm(cwx_context* ctx, int M, int N, int K, const void* A, int64_t sam, int64_t sak, const void* B,
                  int64_t sbk, int64_t sbn, float alpha, float beta, float* C, int64_t ldc) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && (M == 0 || N == 0 || (A && B && C)), "NULL argument");
    CWX_CHECK(M >= 0 && N >= 0 && K >= 0 && ldc >= N, "bad GEMM dimensions");
    set_device(ctx);
    cwx::gemm_bf16(M, N, K, (const uint16_t*)A, sam, sak, (const uint16_t*)B, sbk, sbn, alpha, beta, C, ldc,
                   ctx->c.stream);
    return CWX_OK;
  });
}

int cwx_mlp_dense_fwd(cwx_context* ctx, int M, int N, int K, const void* X, int64_t ldx, const void* W, int64_t ldw,
                      const float* bias, int act, float* Z, void* H_bf16, float* H_f32) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && (M == 0 || (X && W)), "NULL argument");
    CWX_CHECK(ldx >= K && ldw >= K, "bad dense-layer strides");
    set_device(ctx);
    cwx::mlp_dense_fwd(M, N, K, (const uint16_t*)X, ldx, (const uint16_t*)W, ldw, bias, act, Z, (uint16_t*)H_bf16,
                       H_f32, ctx->c.stream);
    return CWX_OK;
  });
}

int cwx_mlp_act_fwd(cwx_context* ctx, int M, int N, int act, const float* Zacc, const float* bias, float* Z,
                    void* H_bf16, float* H_f32) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && Zacc, "NULL argument");
    CWX_CHECK(M >= 0 && N >= 1 && (act == 0 || act == 1), "bad activation arguments");
    set_device(ctx);
    cwx::mlp_act_fwd(M, N, act, Zacc, bias, Z, (uint16_t*)H_bf16, H_f32, ctx->c.stream);
    return CWX_OK;
  });
}

int cwx_mlp_act_bwd(cwx_context* ctx, int M, int N, int act, const float* Z, const float* dH, float* dZ,
                    void* dZ_bf16, float* dbias) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && dH && (act == 0 || Z), "NULL argument");
    CWX_CHECK(M >= 0 && N >= 1 && (act == 0 || act == 1), "bad activation arguments");
    set_device(ctx);
    cwx::mlp_act_bwd(M, N, act, Z, dH, dZ, (uint16_t*)dZ_bf16, dbias, ctx->c.stream);
    return CWX_OK;
  });
}

int cwx_mlp_mse(cwx_context* ctx, int M, int N, const float* Y, const float* T, float* dY, void* dY_bf16,
                float* loss) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && Y && T, "NULL argument");
    CWX_CHECK(M >= 1 && N >= 1, "bad loss dimensions");
    set_device(ctx);
    if (loss) CWX_HIP(hipMemsetAsync(loss, 0, sizeof(float), ctx->c.stream));
    cwx::mlp_mse(M, N, Y, T, dY, (uint16_t*)dY_bf16, loss, ctx->c.stream);
    return CWX_OK;
  });
}

int cwx_mlp_adam(cwx_context* ctx, int64_t n, float lr, float beta1, float beta2, float eps, int step, float* w,
                 const float* g, float* m, float* v, void* w_bf16) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && w && g && m && v, "NULL argument");
    CWX_CHECK(n >= 0, "negative size");
    set_device(ctx);
    if (n) cwx::mlp_adam(n, lr, beta1, beta2, eps, step, w, g, m, v, (uint16_t*)w_bf16, ctx->c.stream);
    return CWX_OK;
  });
}

int cwx_f32_to_bf16(cwx_context* ctx, int64_t n, const float* x, void* y) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && (n == 0 || (x && y)), "NULL argument");
    set_device(ctx);
    if (n) cwx::to_bf16(n, x, (uint16_t*)y, ctx->c.stream);
    return CWX_OK;
  });
}

// ---------------------------------------------------------------------------------------------
int cwx_cw_read_unpacked(const char* path, int64_t* nwaves, int* layers, cwx_wave_f64* W, double* z_layers,
                         int* waves_begin, int* waves_end) {
  return guard([&]() -> int {
    CWX_CHECK(path && nwaves && layers, "NULL argument");
    std::ifstream f(path, std::ios::binary);
    if (!f) return fail(CWX_ERR_IO, std::string("cannot open ") + path);
    std::vector<unsigned char> buf((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    size_t off = 0;
    auto need = [&](size_t n) {
      if (off + n > buf.size()) throw cwx::Error(CWX_ERR_IO, std::string("truncated .cw file ") + path);
    };
    auto u64 = [&]() { need(8); uint64_t v; std::memcpy(&v, &buf[off], 8); off += 8; return v; };
    auto f64 = [&]() { need(8); double v; std::memcpy(&v, &buf[off], 8); off += 8; return v; };
    int rec_c = 5;
    if (buf.size() >= 9 && (buf[0] == 0 || buf[0] == 1)) {
      uint64_t p;
      std::memcpy(&p, &buf[1], 8);
      if (p == 4 || p == 8) { off = 1; rec_c = 6; }
    }
    const uint64_t prec = u64();
    if (prec != 8) return fail(CWX_ERR_IO, "only double-precision .cw files are supported");
    struct List { size_t off; uint64_t n; };
    std::vector<List> lists;
    std::vector<double> zs;
    auto read_list = [&]() {
      const uint64_t n = u64();
      need(n * rec_c * 16);
      lists.push_back(List{off, n});
      off += n * rec_c * 16;
    };
    read_list();
    const uint64_t nb = u64();
    for (uint64_t l = 0; l < nb; ++l) {
      zs.push_back(f64());
      read_list();
      read_list();
    }
    int64_t total = 0;
    for (auto& L : lists) total += (int64_t)L.n;
    *nwaves = total;
    *layers = (int)nb;
    if (!W) return CWX_OK;
    CWX_CHECK(z_layers && waves_begin && waves_end, "NULL argument");
    int64_t idx = 0;
    for (auto& L : lists)
      for (uint64_t i = 0; i < L.n; ++i) {
        const double* r = (const double*)&buf[L.off + i * rec_c * 16];
        cwx_wave_f64& w = W[idx++];
        if (rec_c == 6) {
          for (int c = 0; c < 3; ++c) {
            w.E0[c] = cwx_complex{r[2 * c], r[2 * c + 1]};
            w.k[c] = cwx_complex{r[6 + 2 * c], r[6 + 2 * c + 1]};
          }
        } else {
          w.E0[0] = cwx_complex{r[0], r[1]};
          w.E0[1] = cwx_complex{r[2], r[3]};
          w.E0[2] = cwx_complex{0.0, 0.0};
          for (int c = 0; c < 3; ++c) w.k[c] = cwx_complex{r[4 + 2 * c], r[4 + 2 * c + 1]};
        }
      }
    // region map of cw_allocate (cpuEvaluator.cpp:34-53)
    int64_t t = (int64_t)lists[0].n;
    waves_begin[0] = 0;
    for (uint64_t l = 0; l < nb; ++l) {
      z_layers[l] = zs[l];
      t += (int64_t)lists[1 + 2 * l].n;
      waves_end[l] = (int)t;
      waves_begin[l + 1] = (int)t;
      t += (int64_t)lists[2 + 2 * l].n;
    }
    waves_end[nb] = (int)t;
    return CWX_OK;
  });
}

// ---------------------------------------------------------------------------------------------
static int field_init(cwx_context* ctx, const void* W, size_t recsize, int prec, int64_t nwaves,
                      const float* z_layers, int layers, const int* wb, const int* we) {
  CWX_CHECK(ctx && W && wb && we && (layers == 0 || z_layers), "NULL argument");
  CWX_CHECK(nwaves >= 1 && layers >= 0, "empty wave table");
  for (int l = 0; l <= layers; ++l)
    CWX_CHECK(wb[l] >= 0 && we[l] >= wb[l] && we[l] <= nwaves, "region bounds outside the wave table");
  set_device(ctx);
  auto& c = ctx->c;
  hipStream_t s = c.stream;
  c.ftab.alloc(recsize * nwaves);
  CWX_HIP(hipMemcpyAsync(c.ftab.p, W, recsize * nwaves, hipMemcpyHostToDevice, s));
  c.fz.alloc(std::max(1, layers));
  if (layers) CWX_HIP(hipMemcpyAsync(c.fz.p, z_layers, sizeof(float) * layers, hipMemcpyHostToDevice, s));
  c.fwb.alloc(layers + 1);
  c.fwe.alloc(layers + 1);
  CWX_HIP(hipMemcpyAsync(c.fwb.p, wb, sizeof(int) * (layers + 1), hipMemcpyHostToDevice, s));
  CWX_HIP(hipMemcpyAsync(c.fwe.p, we, sizeof(int) * (layers + 1), hipMemcpyHostToDevice, s));
  CWX_HIP(hipStreamSynchronize(s));
  c.table = cwx::FieldTable{c.ftab.p, prec, (long)nwaves, c.fz.p, c.fwb.p, c.fwe.p, layers};
  return CWX_OK;
}

int cwx_field_init_f32(cwx_context* ctx, const cwx_wave_f32* W, int64_t nwaves, const float* z_layers, int layers,
                       const int* waves_begin, const int* waves_end) {
  return guard([&]() -> int {
    return field_init(ctx, W, sizeof(cwx_wave_f32), 4, nwaves, z_layers, layers, waves_begin, waves_end);
  });
}

int cwx_field_init_f64(cwx_context* ctx, const cwx_wave_f64* W, int64_t nwaves, const float* z_layers, int layers,
                       const int* waves_begin, const int* waves_end) {
  return guard([&]() -> int {
    return field_init(ctx, W, sizeof(cwx_wave_f64), 8, nwaves, z_layers, layers, waves_begin, waves_end);
  });
}

int cwx_field_init_solution(cwx_context* ctx, const cwx_solution* sol) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && sol, "NULL argument");
    const auto& s = sol->s;
    CWX_CHECK(s.nrec > 0, "solution has no records");
    set_device(ctx);
    auto& c = ctx->c;
    const int nb = s.st.L + 1;
    std::vector<float> zl(nb);
    std::vector<int> wb(nb + 1), we(nb + 1);
    long t = s.seg_cnt[0];
    wb[0] = 0;
    for (int l = 0; l < nb; ++l) {
      zl[l] = (float)((s.opt.external && l > 0 && l < nb - 1) ? 0.0 : s.st.z[l]);
      t += s.seg_cnt[1 + 2 * l];
      we[l] = (int)t;
      wb[l + 1] = (int)t;
      t += s.seg_cnt[2 + 2 * l];
    }
    we[nb] = (int)t;
    hipStream_t st = c.stream;
    c.ftab.release();
    c.fz.alloc(nb);
    c.fwb.alloc(nb + 1);
    c.fwe.alloc(nb + 1);
    CWX_HIP(hipMemcpyAsync(c.fz.p, zl.data(), sizeof(float) * nb, hipMemcpyHostToDevice, st));
    CWX_HIP(hipMemcpyAsync(c.fwb.p, wb.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice, st));
    CWX_HIP(hipMemcpyAsync(c.fwe.p, we.data(), sizeof(int) * (nb + 1), hipMemcpyHostToDevice, st));
    CWX_HIP(hipStreamSynchronize(st));
    c.table = cwx::FieldTable{s.rec.p, 8, s.nrec, c.fz.p, c.fwb.p, c.fwe.p, nb};
    return CWX_OK;
  });
}

int cwx_field_grid_dev(cwx_context* ctx, void* d_out, int out_f64, int tol_mode, float x0, float y0, float z0,
                       float dx, float dy, float dz, int nx, int ny, int nz) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && d_out, "NULL argument");
    if (ctx->c.table.nwaves <= 0) return fail(CWX_ERR_STATE, "no wave table: call cwx_field_init_* first");
    CWX_CHECK(nx >= 1 && ny >= 1 && nz >= 1 && nz <= 65535, "bad grid dimensions");
    CWX_CHECK(tol_mode == 0 || tol_mode == 1, "tol_mode must be 0 or 1");
    set_device(ctx);
    cwx::GridSpec g{x0, y0, z0, dx, dy, dz, nx, ny, nz};
    cwx::launch_field_direct(ctx->c.table, g, tol_mode, out_f64, d_out, ctx->c.stream);
    return CWX_OK;
  });
}

int cwx_field_solution_grid_dev(cwx_context* ctx, const cwx_solution* sol, void* d_out, int out_f64, int tol_mode,
                                float x0, float y0, float z0, float dx, float dy, float dz, int nx, int ny,
                                int nz) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && sol && d_out, "NULL argument");
    CWX_CHECK(nx >= 1 && ny >= 1 && nz >= 1 && nz <= 65535, "bad grid dimensions");
    CWX_CHECK(tol_mode == 0 || tol_mode == 1, "tol_mode must be 0 or 1");
    set_device(ctx);
    cwx::field_separable(sol->s, cwx::GridSpec{x0, y0, z0, dx, dy, dz, nx, ny, nz}, tol_mode, out_f64, d_out,
                         ctx->c.sep, ctx->c.stream);
    return CWX_OK;
  });
}

int cwx_field_evaluate(cwx_context* ctx, cwx_complexf* E_xy, cwx_complexf* E_xz, cwx_complexf* E_yz, float x_start,
                       float y_start, float z_start, float x, float y, float z, float d, int N) {
  return cwx_field_evaluate_rule(ctx, E_xy, E_xz, E_yz, x_start, y_start, z_start, x, y, z, d, N, 0);
}

int cwx_field_evaluate_rule(cwx_context* ctx, cwx_complexf* E_xy, cwx_complexf* E_xz, cwx_complexf* E_yz,
                            float x_start, float y_start, float z_start, float x, float y, float z, float d, int N,
                            int tol_mode) {
  return guard([&]() -> int {
    CWX_CHECK(ctx && E_xy && E_xz && E_yz, "NULL argument");
    CWX_CHECK(N >= 1 && N <= 65535, "bad slice resolution");
    CWX_CHECK(tol_mode == 0 || tol_mode == 1, "tol_mode must be 0 (z >= z_l - 0.001) or 1 (z >= z_l)");
    if (ctx->c.table.nwaves <= 0) return fail(CWX_ERR_STATE, "no wave table: call cwx_field_init_* first");
    set_device(ctx);
    hipStream_t s = ctx->c.stream;
    const size_t bytes = sizeof(cwx_complexf) * 3 * (size_t)N * N;
    cwx::DevBuf<unsigned char> dE(bytes);
    // kernel_xy: plane z, rows along y (gpuEvaluator.cu:70-89)
    cwx::launch_field_direct(ctx->c.table, cwx::GridSpec{x_start, y_start, z, d, d, 0.f, N, N, 1}, tol_mode, 0, dE.p,
                             s);
    CWX_HIP(hipMemcpyAsync(E_xy, dE.p, bytes, hipMemcpyDeviceToHost, s));
    CWX_HIP(hipStreamSynchronize(s));
    // kernel_xz: plane y, rows along z
    cwx::launch_field_direct(ctx->c.table, cwx::GridSpec{x_start, y, z_start, d, 0.f, d, N, 1, N}, tol_mode, 0, dE.p,
                             s);
    CWX_HIP(hipMemcpyAsync(E_xz, dE.p, bytes, hipMemcpyDeviceToHost, s));
    CWX_HIP(hipStreamSynchronize(s));
    // kernel_yz: plane x, columns along y, rows along z
    cwx::launch_field_direct(ctx->c.table, cwx::GridSpec{x, y_start, z_start, 0.f, d, d, 1, N, N}, tol_mode, 0, dE.p,
                             s);
    CWX_HIP(hipMemcpyAsync(E_yz, dE.p, bytes, hipMemcpyDeviceToHost, s));
    CWX_HIP(hipStreamSynchronize(s));
    return CWX_OK;
  });
}

}  // extern "C"
