/*
 * cwx — MI355X-native coupled-wave solver and plane-wave field evaluator (C ABI).
 *
 * Every entry point returns an int status: CWX_OK (0), a negative error class, or a positive
 * LAPACK-style numerical failure index (non-convergence, zero pivot).  The message of the last
 * failure on the calling thread is available from cwx_last_error().  Complex numbers are
 * interleaved {re, im} pairs, binary compatible with std::complex<double>/<float>; matrices are
 * column-major with explicit leading dimensions, as LAPACK.  A context owns one HIP device and one
 * stream; contexts are independent, a single context must not be used by two threads at once.
 *
 * Reference interfaces replaced (paths relative to the scattersim checkout):
 *   third_Lapack.h:38-85       MKL_eigensolve / MKL_linearsolve / MKL_inverse / MKL_multiply
 *   FourierWave.h:303-397      volume::CalculateD / phi (the per-layer property matrix)
 *   ScatterVolume.cpp:506-1019 the scattervolume solve (argv + .npy -> CoupledWaveStructure)
 *   CoupledWaveStructure.h:63-125  CoupledWaveStructure::save / load
 *   cpuEvaluator.h:6-24        UnpackedWave, cw_allocate, cw_unpack
 *   gpuEvaluator.h:14-17       gpu_initialize, gpu_cw_evaluate
 */
#ifndef CWX_H
#define CWX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct cwx_context cwx_context;
typedef struct cwx_solution cwx_solution;

typedef struct { double re, im; } cwx_complex;
typedef struct { float re, im; } cwx_complexf;

/* UnpackedWave<float> (cpuEvaluator.h:6-10): 48 bytes */
typedef struct { cwx_complexf E0[3]; cwx_complexf k[3]; } cwx_wave_f32;
/* .cw record of this library: {E0.x, E0.y, E0.z, k.x, k.y, k.z}, 96 bytes */
typedef struct { cwx_complex E0[3]; cwx_complex k[3]; } cwx_wave_f64;

enum {
  CWX_OK = 0,
  CWX_ERR_ARG = -1,    /* invalid argument */
  CWX_ERR_DEVICE = -2, /* HIP runtime failure or no device */
  CWX_ERR_IO = -3,     /* file could not be read or written */
  CWX_ERR_STATE = -4   /* call out of order (e.g. evaluate before init) */
};

/* Message of the last failing call on this thread ("" if none). */
const char* cwx_last_error(void);
/* Launch shapes (no counterpart in the reference).  The library reads the table that
 * `python -m scattersim_amd.tune` writes (data/tuning.txt beside the library's lib/ directory, or the
 * file named by CWX_TUNING_FILE); cwx_set_param overrides one key for the process, -1 selecting the
 * built-in choice and CWX_PARAM_TABLE returning to the table.  Keys:
 *   "qr_team"          helper workgroups per matrix of the QR iteration when the matrices do not fill
 *                      the device (0: none); results do not depend on it
 *   "lahr2_team"       helper workgroups per matrix of each Hessenberg panel in that case (0: none)
 *   "qr_threads"       256 or 512 threads per QR workgroup when they do
 *   "batch_chunk_mib"  device memory of one chunk of cwx_solve_batch_dev when max_bytes is 0
 *   "alloc_cache"      1 (built-in): device buffers the library releases are kept (per device, at most
 *                      32 GiB) and reused by later calls instead of being returned with hipFree; 0 returns
 *                      them at once (and empties the cache).  Memory from cwx_dev_alloc is never cached
 *   "zgemm_kernel"     the FP64 complex GEMM: 1 (built-in) the matrix-core kernel with 64x64 tiles, 2 and 3
 *                      its 128x64 and 64x128 tilings, 0 the vector kernel
 *   "chain_block"      the reduced eigen mode's transfer chain: 1 (built-in) in block form from the order-2MF
 *                      eigenvectors and Q, 0 through LUs of the order-4MF Gd (the reference's formulation)
 *   "lu_panel"         zgetrf's panel kernel: 2 (built-in) whole rows per lane, 1 one column pair per
 *                      wavefront (both: 32-column panels up to 512 rows, 16 columns up to 1024), 0 rows per
 *                      lane pair (32 columns); the same pivots
 *   "getrs_inv"        the solves with many right-hand sides in the transfer chain and for S^-1 f3: 1
 *                      (built-in) with the inverses of the factors' 64 x 64 diagonal blocks applied as
 *                      matrix-core products, 0 by substitution in 32 x 32 diagonal blocks
 * zgemm_kernel, chain_block, lu_panel and getrs_inv are the keys that change results, in the last bits (each
 * form sums in its own order); a process that compares bits across calls keeps one setting.
 * cwx_tuning_table returns the path of the table in use ("" if none).  Host-only calls: no device work. */
#define CWX_PARAM_TABLE INT64_MIN
int cwx_set_param(const char* name, int64_t value);
/* The value in force for key `name` at shape (a, b): the override, else the table's rule, else -1. */
int64_t cwx_get_param(const char* name, int64_t a, int64_t b);
const char* cwx_tuning_table(void);
/* Library version string. */
const char* cwx_version(void);

/* ---- context ------------------------------------------------------------------------------ */
/* Creates a context on HIP device `device` with its own stream. */
int cwx_create(int device, cwx_context** ctx);
int cwx_destroy(cwx_context* ctx);
/* Use a caller-owned hipStream_t (passed as void*) for all later work; NULL restores the own one. */
int cwx_set_stream(cwx_context* ctx, void* hip_stream);
int cwx_synchronize(cwx_context* ctx);

/* Device memory of the context's device for callers without their own allocator (no reference
 * counterpart).  cwx_memcpy kind: 0 host to device, 1 device to host, 2 device to device; it
 * synchronises the context stream. */
int cwx_dev_alloc(cwx_context* ctx, int64_t bytes, void** p);
int cwx_dev_free(cwx_context* ctx, void* p);
int cwx_memcpy(cwx_context* ctx, void* dst, const void* src, int64_t bytes, int kind);

/* ---- dense complex FP64 linear algebra: the LAPACK seam (third_Lapack.h:38-85) --------------- */
/* Replaces MKL_eigensolve (third_Lapack.h:38-45, LAPACKE_zgeev 'N','V').  A (n x n, host) is
 * overwritten by its Schur form; w (n) eigenvalues; VR (n x n) right eigenvectors of unit 2-norm
 * with their largest component real.  Returns >0 when the QR iteration does not converge. */
int cwx_zgeev(cwx_context* ctx, int n, cwx_complex* A, int lda, cwx_complex* w, cwx_complex* VR,
              int ldvr);
/* Replaces MKL_linearsolve (third_Lapack.h:47-54, LAPACKE_zgesv): A is overwritten by its LU
 * factors, B (n x nrhs) by the solution, ipiv (n, may be NULL) receives 1-based pivots. */
int cwx_zgesv(cwx_context* ctx, int n, int nrhs, cwx_complex* A, int lda, int* ipiv, cwx_complex* B,
              int ldb);
/* Replaces MKL_inverse (third_Lapack.h:56-64, zgesv with an identity right-hand side). */
int cwx_zinverse(cwx_context* ctx, int n, const cwx_complex* A, int lda, cwx_complex* Ainv, int ldinv);
/* Replaces MKL_multiply (third_Lapack.h:75-85, cblas_zgemm): C = alpha A B + beta C. */
int cwx_zgemm(cwx_context* ctx, int m, int n, int k, cwx_complex alpha, const cwx_complex* A, int lda,
              const cwx_complex* B, int ldb, cwx_complex beta, cwx_complex* C, int ldc);

/* Device-pointer, batched forms of the same operations (no reference counterpart: the reference
 * calls them one matrix at a time from ScatterVolume.cpp:262, 331-333, 739, 768-781).  Matrices
 * are packed with leading dimension n at element strides; work is queued on the context stream.
 * h_info (host, batch entries, may be NULL) receives per-matrix status after a synchronisation. */
int cwx_zgeev_batched_dev(cwx_context* ctx, int n, int batch, void* dA, int64_t strideA, void* dW,
                          int64_t strideW, void* dV, int64_t strideV, int* h_info);
int cwx_zgetrf_batched_dev(cwx_context* ctx, int n, int batch, void* dA, int64_t strideA, void* dIpiv,
                           int64_t strideP, int* h_info);
int cwx_zgetrs_batched_dev(cwx_context* ctx, int n, int nrhs, int batch, const void* dLU,
                           int64_t strideA, const void* dIpiv, int64_t strideP, void* dB, int64_t strideB);
/* No counterpart in the reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  zgesv (MKL_linearsolve,
 * third_Lapack.h:47-54) of `batch` device matrices with the kernels of cwx_zgetrf_batched_dev +
 * cwx_zgetrs_batched_dev: dA is overwritten by its LU factors, dB (nrhs columns) by the solution.  The
 * pivots and per-matrix status live in a static device array of CWX_NOALLOC_SCRATCH_BYTES per GPU,
 * shared by every call on that GPU: n * batch may be at most CWX_NOALLOC_MAX_PIVOTS, else CWX_ERR_ARG.
 * Calls that overlap on one GPU (two streams, two threads) race on the scratch.  No hipMalloc, hipFree
 * or allocating library call is made.  h_info (host, may be NULL) as cwx_zgetrf_batched_dev. */
#define CWX_NOALLOC_SCRATCH_BYTES (4 << 20)
#define CWX_NOALLOC_MAX_PIVOTS (((4 << 20) - 65536 * 4) / 4)
int cwx_zgesv_batched_dev_noalloc(cwx_context* ctx, int n, int nrhs, int batch, void* dA, int64_t strideA,
                                  void* dB, int64_t strideB, int* h_info);
int cwx_zgemm_batched_dev(cwx_context* ctx, int transa, int transb, int m, int n, int k,
                          cwx_complex alpha, const void* dA, int lda, int64_t strideA, const void* dB,
                          int ldb, int64_t strideB, cwx_complex beta, void* dC, int ldc, int64_t strideC,
                          int batch);
/* The product of cwx_zgemm_batched_dev (MKL_multiply, third_Lapack.h:75-85, as called by the transfer
 * chain at ScatterVolume.cpp:330-334 and 768-781) on the FP32 matrix cores: every FP64 operand is split
 * into hi = fl32(x) and lo = fl32(x - hi), a real product taken as hi hi + hi lo + lo hi, 16-term partial
 * sums accumulated in FP32 and added in FP64.  Relative error about 1e-7 of |A| |B|, not FP64: for
 * callers that accept that accuracy.  terms = 3 takes that split form, terms = 1 only hi hi (the operands
 * rounded to FP32, a third of the matrix-core work).  Otherwise the arguments and errors of
 * cwx_zgemm_batched_dev. */
int cwx_zgemm_split32_batched_dev(cwx_context* ctx, int transa, int transb, int m, int n, int k,
                                  cwx_complex alpha, const void* dA, int lda, int64_t strideA, const void* dB,
                                  int ldb, int64_t strideB, cwx_complex beta, void* dC, int ldc,
                                  int64_t strideC, int batch, int terms);

/* ---- the scattervolume solve (ScatterVolume.cpp:506-1019) ------------------------------------ */
/* Options of ScatterVolume.cpp:510-530; cwx_volume_params_default fills the reference defaults. */
typedef struct {
  double lambda;       /* --lambda (1.0) */
  double direction[3]; /* --direction (0 0 1) */
  double ex[2];        /* --ex (0 0) as (re, im) */
  double ey[2];        /* --ey (1 0) */
  double ez[2];        /* --ez (0 0) */
  double n[2];         /* --n (1 1): upper and lower half-space indices */
  double kappa;        /* --kappa (0): absorbance of the lower half-space */
  double size[3];      /* --size (10 10 2) */
  double z;            /* --z (0) */
  int ncoef;           /* number of --coef values: 0, 1 or 2 */
  int coef[2];
  int psf;             /* --psf */
  int external;        /* --external */
  int npz;             /* --npz */
  int no_cutoff;       /* library only: bypass the silent order cut-off of ScatterVolume.cpp:613-642 */
  int eig_mode;        /* 0: zgeev of the 4MF x 4MF Phi (reference); 1: 2MF eigenproblem of P Q */
} cwx_volume_params;
void cwx_volume_params_default(cwx_volume_params* p);

/* Solve one sample: (L, s1, s2) complex128 refractive indices in C order, on the host
 * (cwx_solve_volume) or already on the device (cwx_solve_volume_dev).  A 2-D sample is L = 1.
 * Returns >0 if an eigensolve did not converge or a factorisation met a zero pivot (the solution
 * object is still returned for inspection). */
int cwx_solve_volume(cwx_context* ctx, const cwx_volume_params* p, const cwx_complex* sample, int L,
                     int s1, int s2, cwx_solution** out);
int cwx_solve_volume_dev(cwx_context* ctx, const cwx_volume_params* p, const void* d_sample, int L,
                         int s1, int s2, cwx_solution** out);

/* Solve `count` independent samples together (no counterpart in the reference, which runs one
 * scattervolume process per sample, e.g. spectroscopy.py:66-90).  Instance i has params[i] and L[i]
 * layers; the layers of all instances lie back to back in d_samples ((sum L) x s1 x s2 complex128,
 * C order, device).  All instances must resolve to the same Fourier orders and eigen mode.  Each chunk
 * of instances whose device workspace stays below max_bytes (0: 64 GiB) is one launch set over all of
 * its layers.  d_x (device, count x 6MF complex) receives [Rx Ry Rz Tx Ty Tz] of every instance,
 * h_info (host, count, may be NULL) its status; *M0 and *M1 (may be NULL) the orders.  With
 * d_x == NULL only the orders are computed (ctx may then be NULL).  Returns the first non-zero
 * instance status. */
int cwx_solve_batch_dev(cwx_context* ctx, int count, const cwx_volume_params* params, const int* L,
                        const void* d_samples, int s1, int s2, int64_t max_bytes, void* d_x, int* h_info, int* M0,
                        int* M1);
/* cwx_solve_batch_dev with every chunk of `chunk` instances one launch set captured in a HIP graph (no
 * counterpart in the reference).  The first chunk of a given shape (layer counts in order, orders, eigen
 * mode) is captured, and later chunks of the same shape replay that graph with their own samples and
 * tables; the graphs stay with the context until cwx_destroy.  Results equal cwx_solve_batch_dev's.
 * Teams of the QR iteration are not used inside a graph; an instance whose reduced eigenproblem has a
 * nearly defective pair is solved again outside the graph.  *captured (may be NULL) receives the number
 * of graphs this call captured. */
int cwx_solve_batch_graph_dev(cwx_context* ctx, int count, const cwx_volume_params* params, const int* L,
                              const void* d_samples, int s1, int s2, int chunk, void* d_x, int* h_info,
                              int* captured);

/* cwx_solve_batch_dev for instances whose Fourier orders (or eigen modes) differ, as wavelength and angle
 * sweeps with the order cut-off produce: the instances are grouped by orders and each group is one
 * batched solve.  x_offsets (host, count + 1) receives the start of every instance's [R; T] in d_x
 * (6 M0_i M1_i complex each, instances in input order; x_offsets[count] is the total); M0s, M1s (host,
 * count, may be NULL) the orders.  With d_x == NULL only these are computed (ctx and d_samples may be
 * NULL), so the caller can size d_x first.  No counterpart in the reference. */
int cwx_solve_batch_mixed_dev(cwx_context* ctx, int count, const cwx_volume_params* params, const int* L,
                              const void* d_samples, int s1, int s2, int64_t max_bytes, void* d_x, int64_t* x_offsets,
                              int* h_info, int* M0s, int* M1s);

/* Many incident fields on one sample, sharing every layer's eigen-decomposition, the transfer chain
 * and the boundary factorisation.  The reference builds one incident field, the Fourier
 * coefficients EF of the plane wave (ScatterVolume.cpp:660-674; with --psf every order carries E),
 * and solves A x = [0; -f1 EF] for it (ScatterVolume.cpp:739-740); here K fields are solved with one
 * factorisation of A.  Field k is K x 3MF complex in d_EF (device): its Ex, Ey and Ez
 * coefficients, MF each, in the order of CWX_SOL_EF (order (r, c) of the M1 x M0 grid at r + c M1;
 * a single non-zero order is a plane wave of transverse wave vector (Sx, Sy) k of that order).
 * The wavelength, direction and sample come from p (its ex/ey/ez are not used).  d_x (device,
 * K x 6MF) receives [R; T] per field.  With d_x == NULL only *M0, *M1 are set (ctx may be NULL).
 * No counterpart in the reference, which runs one process per incident field. */
int cwx_solve_fields_dev(cwx_context* ctx, const cwx_volume_params* p, const void* d_sample, int L, int s1, int s2,
                         int K, const void* d_EF, void* d_x, int* M0, int* M1);

/* Replaces volume::CalculateD / phi (FourierWave.h:303-397): the property matrix of every layer,
 * Phi = [[0, P], [Q, 0]] (4MF x 4MF, column-major, blocks = 0, L matrices) or its blocks P and Q
 * (2MF x 2MF each, blocks = 1: the L P blocks, then the L Q blocks) into device memory d_out.
 * *n_out receives the matrix order; with d_out == NULL only the order is computed. */
int cwx_phi_build_dev(cwx_context* ctx, const cwx_volume_params* p, const void* d_sample, int L, int s1,
                      int s2, int blocks, void* d_out, int* n_out);

typedef struct {
  int M0, M1, MF, L;   /* Fourier orders along x, y; MF = M0 M1; layers */
  int cutoff;          /* the order cut-off heuristic fired */
  int info;            /* numerical status of the solve */
  double k;            /* 2 PI / lambda with PI = 3.141592653 (FourierWave.h:18) */
  int64_t nrec;        /* plane-wave records of the .cw */
  int nlists;          /* 1 + 2 (L + 1): Pi, then Pr and Pt of every boundary */
} cwx_solution_info;
int cwx_solution_get_info(const cwx_solution* sol, cwx_solution_info* info);

enum {
  CWX_SOL_X = 0,   /* 6MF complex: [Rx Ry Rz Tx Ty Tz] (ScatterVolume.cpp:126-131) */
  CWX_SOL_GAMMA,   /* L x 4MF complex: sorted eigenvalues, even = down, odd = up */
  CWX_SOL_BETA,    /* L x 4MF complex: mode amplitudes (ScatterVolume.cpp:766-783) */
  CWX_SOL_QCH,     /* L x (3MF x 4MF) complex, column-major: [Q_check | Q_hat] */
  CWX_SOL_Z,       /* L + 1 double: boundary positions */
  CWX_SOL_RECORDS, /* nrec x 6 complex: .cw records in file order */
  CWX_SOL_SX, CWX_SOL_SY, CWX_SOL_SZ0, CWX_SOL_SZ1, /* MF complex each */
  CWX_SOL_EF,      /* 3MF complex */
  CWX_SOL_GD,      /* L x 4MF x 4MF complex */
  CWX_SOL_GC,
  CWX_SOL_NIF      /* L x MF complex: Fourier coefficients of n^-2, (M1 x M0) column-major */
};
/* Number of elements (complex, or double for CWX_SOL_Z) of a solution array. */
int64_t cwx_solution_size(const cwx_solution* sol, int what);
int cwx_solution_copy(cwx_context* ctx, const cwx_solution* sol, int what, void* host_dst);
/* Device address of a solution array (NULL for host-only arrays). */
const void* cwx_solution_device_ptr(const cwx_solution* sol, int what);
/* Offsets and counts (in records) of the Pi / Pr_l / Pt_l lists, nlists entries each. */
int cwx_solution_lists(const cwx_solution* sol, int64_t* offsets, int64_t* counts);
/* CoupledWaveStructure::save (CoupledWaveStructure.h:63-92) with 96-byte records. */
int cwx_solution_write_cw(cwx_context* ctx, const cwx_solution* sol, const char* path);
/* Compact heterogeneous .cw (docs/cw_file_format.pptx slide 1, read by ScatterView.py:68-178):
 * is_volume = 1, the incident, reflected and transmitted waves only, then per layer the mode data
 * beta, gamma, gg (4MF x 4MF, row-major) and the n^-2 coefficients, instead of every interior plane
 * wave.  A trailer after the layers (past the reference reader's end) keeps the wavelength, the
 * incident field and the boundary solution, so cwx_solution_read_cw2 can rebuild the solution.
 * An --external solution (no beta) is refused with CWX_ERR_ARG. */
int cwx_solution_write_cw2(cwx_context* ctx, const cwx_solution* sol, const char* path);
/* Rebuilds a device solution from a file written by cwx_solution_write_cw2 (the interior wave
 * records, Q rows and tables are recomputed from the mode data, as cpuEvaluator.cpp:109-210 does
 * from the reference's compact files).  Files without the trailer are refused with CWX_ERR_IO. */
int cwx_solution_read_cw2(cwx_context* ctx, const char* path, cwx_solution** out);
int cwx_solution_free(cwx_solution* sol);

/* ---- .cw files (CoupledWaveStructure.h:94-125 + cpuEvaluator.cpp:34-102) --------------------- */
/* Reads a .cw file (this library's 96-byte records, or the reference's legacy 80-byte records
 * {E0.x, E0.y, k.x, k.y, k.z}, E0.z taken as 0) and unpacks it: call once with W == NULL to get the
 * sizes, then with buffers of nwaves records, `layers` z values and layers + 1 region bounds. */
int cwx_cw_read_unpacked(const char* path, int64_t* nwaves, int* layers, cwx_wave_f64* W,
                         double* z_layers, int* waves_begin, int* waves_end);

/* ---- homogeneous layer stacks (ScatterLayer.cpp:110-420) -------------------------------------- */
/* The per-incident-wave loop of scatterlayer (ScatterLayer.cpp:390-420) for P incident waves at once:
 * for each wave the 6L x 6L system of ScatterLayer.cpp:152-232 (incident amplitude in layer 0, no
 * reflected wave in layer L-1, Gauss's law per wave, Ex Ey Hx Hy continuity per interface) is
 * solved and turned into the reflected wave of layer l and the transmitted wave of layer l+1 at
 * boundary l, wound to z[l] (mat2waves, ScatterLayer.cpp:240-266).  n: L real indices, kappa: L-1
 * absorbances of layers 1..L-1, z: L-1 increasing boundaries (the expansion of the command-line
 * lists, ScatterLayer.cpp:327-344, is the caller's).  incident: P records with real k;
 * reflected / transmitted: P x (L-1) records, wave-major.  A positive return counts incident waves
 * whose system had an exactly zero pivot. */
int cwx_layer_solve(cwx_context* ctx, int L, const double* n, const double* kappa, const double* z, double lambda,
                    int64_t P, const cwx_wave_f64* incident, cwx_wave_f64* reflected, cwx_wave_f64* transmitted);
/* The same with incident / reflected / transmitted in device memory (n, kappa, z stay on the host). */
int cwx_layer_solve_dev(cwx_context* ctx, int L, const double* n, const double* kappa, const double* z,
                        double lambda, int64_t P, const cwx_wave_f64* d_incident, cwx_wave_f64* d_reflected,
                        cwx_wave_f64* d_transmitted);
/* CoupledWaveStructure::save (CoupledWaveStructure.h:63-92) of a layer solve as scatterlayer writes it
 * (ScatterLayer.cpp:390-424): all P incident waves, then per boundary z[l] its P reflected and P
 * transmitted waves; 96-byte records behind an is_volume byte of 0, as cwx_solution_write_cw. */
int cwx_layer_write_cw(const char* path, int L, const double* z, int64_t P, const cwx_wave_f64* incident,
                       const cwx_wave_f64* reflected, const cwx_wave_f64* transmitted);

/* ---- focused-beam incidence (host only) ------------------------------------------------------ */
/* tira::planewave<double>::SolidAnglePolar, an external dependency of the reference called at
 * ScatterLayer.cpp:383-384 and ScatterPlane.cpp:240-241: the cone of half-angle alpha (obscured below
 * beta) about k, sampled at theta_i = beta + i (alpha - beta) / N0 (i < N0) and phi_j = 2 pi j / N1
 * (j < N1, the outer index: out[j N0 + i]).  Each wave has |k| of the axis, amplitude weight
 * sin(theta) dtheta dphi on E's x and y components, and Ez from k . E0 = 0.  A tilted axis rotates
 * the cone about z x k (Rodrigues).  This reproduces the incident lists of x64/a.cw (axis +z) and
 * data/psf.cw (axis (1, 0, 1)); an obscuration beta > 0 is covered by no reference output. */
int cwx_beam_polar(double alpha, double beta, const double k[3], const cwx_complex E[3], int N0, int N1,
                   cwx_wave_f64* out);
/* tira::planewave<double>::SolidAngleMC (ScatterLayer.cpp:381-382, ScatterPlane.cpp:238-239): N
 * directions uniform in solid angle over the same cone (std::mt19937_64 seeded with seed), each with
 * weight (cap solid angle) / N.  The reference's random sequence is tira's and is not reproduced. */
int cwx_beam_montecarlo(double alpha, double beta, const double k[3], const cwx_complex E[3], int64_t N,
                        uint64_t seed, cwx_wave_f64* out);

/* ---- learned closure (no counterpart in the reference) ------------------------------------------ */
/* Kernels of a small multilayer perceptron that maps a low-order solution to high-order reflection and
 * transmission coefficients, trained on the solver's own solutions.  Device pointers throughout;
 * BF16 arrays are uint16 bit patterns, FP32 accumulation, row-major activations (rows = samples).
 * GEMM: C (M x N, row stride ldc) = alpha op(A) op(B) + beta C with A(m, k) = A[m sam + k sak] and
 * B(k, n) = B[k sbk + n sbn] in BF16 on the matrix cores (v_mfma_f32_16x16x16_bf16). */
int cwx_bf16_gemm(cwx_context* ctx, int M, int N, int K, const void* A, int64_t sam, int64_t sak, const void* B,
                  int64_t sbk, int64_t sbn, float alpha, float beta, float* C, int64_t ldc);
/* One dense layer in one launch: Z = X W^T + bias and H = act(Z), with X (M x K, row stride ldx) and W
 * (N x K, row stride ldw) in BF16; K, ldx and ldw multiples of 4, X and W 8-byte aligned.  Outputs as
 * cwx_mlp_act_fwd (row stride N). */
int cwx_mlp_dense_fwd(cwx_context* ctx, int M, int N, int K, const void* X, int64_t ldx, const void* W, int64_t ldw,
                      const float* bias, int act, float* Z, void* H_bf16, float* H_f32);
/* Z = Zacc + bias (bias per column, may be NULL); H = act(Z) as BF16 (H_bf16) and/or FP32 (H_f32);
 * act 0 is the identity, 1 the tanh form of GELU.  Z, H_bf16 and H_f32 may be NULL. */
int cwx_mlp_act_fwd(cwx_context* ctx, int M, int N, int act, const float* Zacc, const float* bias, float* Z,
                    void* H_bf16, float* H_f32);
/* dZ = dH act'(Z) as FP32 (dZ) and/or BF16 (dZ_bf16); dbias (may be NULL) = column sums of dZ. */
int cwx_mlp_act_bwd(cwx_context* ctx, int M, int N, int act, const float* Z, const float* dH, float* dZ,
                    void* dZ_bf16, float* dbias);
/* *loss (device, may be NULL) = mean((Y - T)^2) over M x N; dY = 2 (Y - T) / (M N) (FP32 and/or BF16). */
int cwx_mlp_mse(cwx_context* ctx, int M, int N, const float* Y, const float* T, float* dY, void* dY_bf16,
                float* loss);
/* One bias-corrected Adam step (step >= 1) on n FP32 weights w with gradient g and moments m, v;
 * w_bf16 (may be NULL) receives the rounded copy used by the GEMMs. */
int cwx_mlp_adam(cwx_context* ctx, int64_t n, float lr, float beta1, float beta2, float eps, int step, float* w,
                 const float* g, float* m, float* v, void* w_bf16);
/* y = x rounded to BF16 (to nearest, ties to even). */
int cwx_f32_to_bf16(cwx_context* ctx, int64_t n, const float* x, void* y);

/* ---- field evaluation (gpuEvaluator.h:14-17) ------------------------------------------------- */
/* gpu_initialize (gpuEvaluator.cu:15-26): upload an unpacked FP32 wave table and region map. */
int cwx_field_init_f32(cwx_context* ctx, const cwx_wave_f32* W, int64_t nwaves, const float* z_layers,
                       int layers, const int* waves_begin, const int* waves_end);
/* The same with an FP64 table; evaluation then reduces phases in FP64 (~1e-7 relative error). */
int cwx_field_init_f64(cwx_context* ctx, const cwx_wave_f64* W, int64_t nwaves, const float* z_layers,
                       int layers, const int* waves_begin, const int* waves_end);
/* Use the device-resident records of a solve as the FP64 table (no copy; the solution must
 * outlive the evaluations). */
int cwx_field_init_solution(cwx_context* ctx, const cwx_solution* sol);
/* gpu_cw_evaluate (gpuEvaluator.cu:131-165): three N x N slices at z (xy), y (xz) and x (yz) into
 * caller-owned host buffers of N*N*3 complex<float>, rows along the second axis. */
int cwx_field_evaluate(cwx_context* ctx, cwx_complexf* E_xy, cwx_complexf* E_xz, cwx_complexf* E_yz,
                       float x_start, float y_start, float z_start, float x, float y, float z, float d,
                       int N);
/* Replaces cpu_cw_evaluate_xy / _yz / _xz (cpuEvaluator.cpp:294-407), which scatterview runs for
 * --cuda -1 (ScatterView.cpp:363-370): the same three slices with the region test selected by tol_mode,
 * 1 being the CPU evaluator's z >= z_l without the 0.001 tolerance (cpuEvaluator.cpp:303,349,386).  The
 * sum itself still runs on the GPU; tol_mode 0 is cwx_field_evaluate. */
int cwx_field_evaluate_rule(cwx_context* ctx, cwx_complexf* E_xy, cwx_complexf* E_xz, cwx_complexf* E_yz,
                            float x_start, float y_start, float z_start, float x, float y, float z, float d, int N,
                            int tol_mode);
/* Evaluate a regular grid into device memory: out[iz][iy][ix][3] complex<float> (out_f64 = 0) or
 * complex<double> (out_f64 = 1, FP64 tables only) at x0 + dx*ix etc.  tol_mode 0 uses the GPU
 * region test z >= z_l - 0.001 (gpuEvaluator.cu:34), 1 the CPU one z >= z_l (cpuEvaluator.cpp:303). */
int cwx_field_grid_dev(cwx_context* ctx, void* d_out, int out_f64, int tol_mode, float x0, float y0,
                       float z0, float dx, float dy, float dz, int nx, int ny, int nz);

/* Field of a solution on a regular grid without a wave table: the records of every region share
 * the per-order transverse wavevectors, so the sum of gpuEvaluator.cu:53-66 is regrouped as
 * exp(i kx x) . C(z) . exp(i ky y) with C(z) from the mode amplitudes (FP64 throughout).  Same
 * output layout, region test and tol_mode as cwx_field_grid_dev; no counterpart in the reference. */
int cwx_field_solution_grid_dev(cwx_context* ctx, const cwx_solution* sol, void* d_out, int out_f64,
                                int tol_mode, float x0, float y0, float z0, float dx, float dy, float dz,
                                int nx, int ny, int nz);

#ifdef __cplusplus
}
#endif

#endif /* CWX_H */
