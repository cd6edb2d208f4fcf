// Host-side coupled-wave solver: option handling, device pipeline and solution storage.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "cwx_common.h"
#include "cwx_kernels.h"

namespace cwx {

// Options of scattervolume (ScatterVolume.cpp:510-530).
struct VolumeOptions {
  double lambda = 1.0;
  double direction[3] = {0.0, 0.0, 1.0};
  double ex[2] = {0.0, 0.0}, ey[2] = {1.0, 0.0}, ez[2] = {0.0, 0.0};
  double n[2] = {1.0, 1.0};
  double kappa = 0.0;
  double size[3] = {10.0, 10.0, 2.0};
  double z = 0.0;
  int ncoef = 0;
  int coef[2] = {0, 0};
  bool psf = false;
  bool external = false;
  bool npz = false;
  bool no_cutoff = false;  // library callers may bypass the silent Fourier-order cut-off
  int eig_mode = 0;        // 0: zgeev of the full 4MF Phi; 1: 2MF reduction through P Q
};

// Everything computed on the host before the per-layer work (ScatterVolume.cpp:574-685).
struct Setup {
  int M0 = 0, M1 = 0, MF = 0, L = 0, s1 = 0, s2 = 0;
  double k = 0.0;
  double dir_n[3] = {0, 0, 0};
  zc e[3];
  zc ni[2];
  std::vector<double> z;            // L + 1 boundary positions
  std::vector<double> up, wq;       // M0, M1
  std::vector<zc> Sx, Sy, Sz0, Sz1; // MF each, column-major flattening of the (M1 x M0) meshes
  std::vector<zc> EF;               // 3 MF
  std::vector<int> rows, cols;      // DFT frequency index kept at each cropped row / column
  bool cutoff = false;              // the order cut-off heuristic fired
};

Setup make_setup(const VolumeOptions& o, int L, int s1, int s2);

struct Context;

// Device-resident solution of one sample.
struct Solution {
  VolumeOptions opt;
  Setup st;
  int n4 = 0;                 // 4 MF
  DevBuf<zc> gam;             // L x 4MF sorted eigenvalues (even: down, odd: up)
  DevBuf<zc> Gd, Gc;          // L x (4MF)^2
  DevBuf<zc> x;               // 6 MF boundary solution [R; T]
  DevBuf<zc> beta;            // L x 4MF
  DevBuf<zc> Qch;             // L x (3MF x 4MF) = [Q_check | Q_hat] per layer
  DevBuf<zc> NIF;             // L x MF (M1 x M0 column-major)
  DevBuf<zc> tabs;            // Sx | Sy | Sz0 | Sz1 (MF each) | EF (3MF)
  DevBuf<zc> rec;             // .cw records, nrec x 6 (E0xyz, kxyz)
  std::vector<long> seg_off;  // per list: Pi, then Pr_0, Pt_0, Pr_1, ...
  std::vector<long> seg_cnt;
  long nrec = 0;
  int info = 0;               // >0: numerical failure (eig non-convergence or singular pivot)
  std::string info_msg;
};

// Solve one sample given as (L, s1, s2) complex128 refractive indices in C order on the device.
void solve_volume(Context& ctx, const VolumeOptions& o, const zc* d_sample, int L, int s1, int s2,
                  Solution& sol);

// solve_batch with the launches of each chunk of `chunk` instances captured in a HIP graph the first time a
// chunk of that shape (layer counts, orders, eigen mode) comes, and replayed for the others; the graphs
// stay in the context.  Results equal solve_batch's.  captured (may be null) receives the number of
// graphs this call captured.
int solve_batch_graph(Context& ctx, const std::vector<VolumeOptions>& opts, const zc* d_samples,
                      const std::vector<int>& Ls, int s1, int s2, int chunk, zc* d_x, std::vector<int>& info,
                      std::string& first_msg, int* captured);

// Solve many independent samples that share the grid and the Fourier orders: instance i has opts[i]
// and Ls[i] layers, stored back to back in d_samples.  Instances are processed in chunks whose
// device workspace stays below max_bytes (0: 64 GiB); each chunk is one launch set over all of
// its layers.  d_x (device, count x 6MF) receives every instance's [R; T]; info its status.
// Returns the first non-zero status (first_msg describes it).
int solve_batch(Context& ctx, const std::vector<VolumeOptions>& opts, const zc* d_samples, const std::vector<int>& Ls,
                int s1, int s2, size_t max_bytes, zc* d_x, std::vector<int>& info, std::string& first_msg);

// Like solve_batch, for instances whose orders (and eigen modes) differ: instances are grouped by orders,
// each group is one batched solve, and instance i's [R; T] (6 M0_i M1_i) lands at d_x + x_off[i].
int solve_batch_mixed(Context& ctx, const std::vector<VolumeOptions>& opts, const zc* d_samples,
                      const std::vector<int>& Ls, int s1, int s2, size_t max_bytes, zc* d_x,
                      const std::vector<long>& x_off, std::vector<int>& info, std::string& first_msg);

// K incident fields on one sample with one eigen-decomposition, one transfer chain and one boundary
// factorisation: d_EF holds K x 3MF incident Fourier coefficients (component-major, order index
// as Setup::EF), d_x receives K x 6MF boundary solutions [R; T].  Returns the solve's status.
int solve_fields(Context& ctx, const VolumeOptions& o, const zc* d_sample, int L, int s1, int s2, int K,
                 const zc* d_EF, zc* d_x, std::string& msg);

// Compact heterogeneous .cw (mode data instead of interior plane waves) and its reverse: a device
// solution rebuilt from such a file (see cwx_solver.hip for the layout).
void cw2_write_file(const Solution& sol, const std::string& path, hipStream_t s);
void cw2_read_file(Context& ctx, const std::string& path, Solution& sol);

// Work buffers of the separable field evaluation.
struct SepWork {
  DevBuf<zc> C, X, Y;
  DevBuf<float2> Tf, Yf;          // complex64 path: T = X C per plane, y phase table
  DevBuf<unsigned char> pack;     // boundary z (float, double) and the plane groups of the coefficient pass
  std::vector<unsigned char> pack_host;  // what pack holds
};

// Field of a solution on a regular grid by order-separable synthesis (see cwx_fieldsep.hip).
void field_separable(const Solution& sol, const GridSpec& g, int tol_mode, int out_f64, void* out,
                     SepWork& w, hipStream_t s);

// Work areas of the set pipeline (eigen workspace, stacked matrices, pivots), grown on demand.
struct SetWork {
  EigWork eig;
  DevBuf<zc> w0, w1, w2, w3;
  DevBuf<int> ipiv;
};

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  EigWork eig;      // the LAPACK-seam entry points
  SetWork work;     // the solver
  std::vector<std::shared_ptr<void>> graphs;  // captured batched solves (solve_batch_graph)
  DevBuf<int> info;
  // uploaded field table
  DevBuf<unsigned char> ftab;
  DevBuf<float> fz;
  DevBuf<int> fwb, fwe;
  FieldTable table;
  std::vector<float> h_z;
  SepWork sep;
};

// Phi of every layer (blocks = 0, L x 4MF x 4MF) or its P, Q blocks (blocks = 1) into device memory.
Setup build_phi_dev(const VolumeOptions& o, const zc* d_sample, int L, int s1, int s2, int blocks, zc* out,
                    hipStream_t s);

void cw_write_file(const Solution& sol, const std::string& path, hipStream_t s);

}  // namespace cwx
