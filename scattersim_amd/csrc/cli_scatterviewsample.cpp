// scatterviewsample --nogui: in-sample field slices of a coupled-wave solution (ScatterViewSample.cpp:674-785)
// on top of libcwx.  The reference sums the external waves on the GPU and adds the field inside the sample
// from per-layer mode data: per-order coefficients at each z, then an inverse DFT onto the slice
// (cpuEvaluator.cpp:109-292, FourierWave.h:202-229).  Here a compact heterogeneous .cw is rebuilt into a
// device solution and the whole slice, inside and outside the sample, comes from the separable evaluator,
// which is that algorithm: FP64 per-order coefficients per plane, then the x and y phase sums on the matrix
// cores.  A full .cw (interior plane waves, no mode data) is evaluated from its wave table instead.
// Options, defaults, slice geometry (EvaluateVectorSlices, ScatterViewSample.cpp:354-376) and the (N, N, 3)
// complex64 .npy output follow the reference; the program ends with status 1 after saving, as it does.
#include <chrono>
#include <cmath>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/cwx.h"
#include "cli_args.h"
#include "cwx_npy.h"

int main(int argc, char** argv) {
  cwx_cli::Args a;
  try {
    a = cwx_cli::parse(argc, argv, {{'c', "cuda"}, {'v', "verbose"}}, {"help", "nogui", "verbose", "sample"});
  } catch (const std::exception& e) {
    std::cerr << e.what() << std::endl;
    return 1;
  }
  if (a.has("help")) {
    std::cout << "Allowed options:\n"
                 "  --input arg (=c.cw)       coupled wave structure to evaluate (also positional)\n"
                 "  --help                    produce help message\n"
                 "  -c [ --cuda ] arg (=0)    device number (-1 is CPU-only, not provided by this implementation)\n"
                 "  --nogui                   save an output file without loading the GUI\n"
                 "  -v [ --verbose ]          produce verbose output\n"
                 "  --sample                  load a 3D sample stored as a grid (*.npy)\n"
                 "  --resolution arg (=8)     resolution of the sample field (N = 2^resolution)\n"
                 "  --extent arg (=30)        size of the sample being visualized\n"
                 "  --output arg (=xz.npy)    output file written when the --nogui option is used\n"
                 "  --axis arg (=1)           axis to cut (0 = X, 1 = Y, 2 = Z)\n"
                 "  --center arg (=0 0 0)     center position of the sampled volume\n"
                 "  --slice arg (=0)          coordinate along the specified axis\n"
              << std::endl;
    return 1;
  }
  auto first = [&](const std::string& k, const std::string& def) {
    return (a.has(k) && !a.vals(k).empty()) ? a.vals(k)[0] : def;
  };
  std::string input = first("input", a.positional.empty() ? "c.cw" : a.positional.back());
  int device = std::stoi(first("cuda", "0"));
  float extent = std::stof(first("extent", "30"));
  int resolution = std::stoi(first("resolution", "8"));
  std::string out = first("output", "xz.npy");
  int axis = std::stoi(first("axis", "1"));
  float slice = std::stof(first("slice", "0"));
  float center[3] = {0, 0, 0};
  if (a.has("center")) {
    auto& v = a.vals("center");
    for (size_t i = 0; i < 3 && i < v.size(); ++i) center[i] = std::stof(v[i]);
  }
  const bool verbose = a.has("verbose");
  std::cout << "Visualization working." << std::endl;
  if (!a.has("nogui")) {
    std::cerr << "error: the interactive viewer is not part of this implementation; use --nogui" << std::endl;
    return 1;
  }
  if (device < 0) {
    std::cerr << "error: --cuda -1 (CPU evaluation) is not provided by this implementation" << std::endl;
    return 2;
  }
  if (axis < 0 || axis > 2) {
    std::cout << "Wrong click at the wrong region. " << std::endl;
    return 1;
  }
  if (resolution < 1 || resolution > 14) {
    std::cerr << "error: --resolution must be in 1..14" << std::endl;
    return 1;
  }
  // the plane position of the cut axis is the --slice value (ScatterViewSample.cpp:712-718)
  float plane[3] = {0, 0, 0};
  plane[axis] = slice;
  const auto t0 = std::chrono::system_clock::now();

  bool compact = false;
  {
    std::ifstream f(input, std::ios::binary);
    if (!f) {
      std::cout << "ERROR: file " << input << " not found" << std::endl;
      return 1;
    }
    char b0 = 0;
    compact = f.read(&b0, 1) && b0 == 1;
  }
  const int N = 1 << resolution;
  const float d = extent / (N - 1);
  const float xs = center[0] - extent / 2, ys = center[1] - extent / 2, zs = center[2] - extent / 2;
  std::vector<cwx_complexf> E((size_t)N * N * 3);
  cwx_context* ctx = nullptr;
  cwx_solution* sol = nullptr;
  void* dE = nullptr;
  auto finish = [&](int rc) {
    if (dE) cwx_dev_free(ctx, dE);
    if (sol) cwx_solution_free(sol);
    if (ctx) cwx_destroy(ctx);
    return rc;
  };
  auto fail = [&]() {
    std::cerr << "error: " << cwx_last_error() << std::endl;
    return finish(2);
  };
  if (cwx_create(device, &ctx) != CWX_OK) return fail();
  if (verbose) std::cout << "Using device " << device << " (" << cwx_version() << ")" << std::endl;
  if (compact) {
    if (cwx_solution_read_cw2(ctx, input.c_str(), &sol) != CWX_OK) return fail();
    // one slice as a grid with a single plane along the cut axis: [z][y][x] order gives the reference's
    // E_xy[y][x], E_xz[z][x] and E_yz[z][y] layouts
    const int nx = axis == 0 ? 1 : N, ny = axis == 1 ? 1 : N, nz = axis == 2 ? 1 : N;
    const float x0 = axis == 0 ? plane[0] : xs, y0 = axis == 1 ? plane[1] : ys, z0 = axis == 2 ? plane[2] : zs;
    if (cwx_dev_alloc(ctx, (int64_t)sizeof(cwx_complexf) * E.size(), &dE) != CWX_OK) return fail();
    if (cwx_field_solution_grid_dev(ctx, sol, dE, 0, 0, x0, y0, z0, d, d, d, nx, ny, nz) != CWX_OK ||
        cwx_memcpy(ctx, E.data(), dE, (int64_t)sizeof(cwx_complexf) * E.size(), 1) != CWX_OK)
      return fail();
  } else {
    int64_t nw = 0;
    int layers = 0;
    if (cwx_cw_read_unpacked(input.c_str(), &nw, &layers, nullptr, nullptr, nullptr, nullptr) != CWX_OK) {
      std::cout << "ERROR: file " << input << " not found" << std::endl;
      return finish(1);
    }
    std::vector<cwx_wave_f64> W(nw);
    std::vector<double> zl(layers);
    std::vector<int> wb(layers + 1), we(layers + 1);
    if (cwx_cw_read_unpacked(input.c_str(), &nw, &layers, W.data(), zl.data(), wb.data(), we.data()) != CWX_OK)
      return fail();
    std::vector<cwx_wave_f32> Wf(nw);
    for (int64_t i = 0; i < nw; ++i)
      for (int c = 0; c < 3; ++c) {
        Wf[i].E0[c] = cwx_complexf{(float)W[i].E0[c].re, (float)W[i].E0[c].im};
        Wf[i].k[c] = cwx_complexf{(float)W[i].k[c].re, (float)W[i].k[c].im};
      }
    std::vector<float> zf(zl.begin(), zl.end());
    std::vector<cwx_complexf> Exy(E.size()), Exz(E.size()), Eyz(E.size());
    if (cwx_field_init_f32(ctx, Wf.data(), nw, zf.data(), layers, wb.data(), we.data()) != CWX_OK ||
        cwx_field_evaluate(ctx, Exy.data(), Exz.data(), Eyz.data(), xs, ys, zs, plane[0], plane[1], plane[2], d,
                           N) != CWX_OK)
      return fail();
    E = axis == 0 ? Eyz : (axis == 1 ? Exz : Exy);
  }
  const auto t1 = std::chrono::system_clock::now();
  if (verbose)
    std::cout << "Evaluation takes:" << std::chrono::duration<double>(t1 - t0).count() << "s" << std::endl;
  try {
    cwx_npy::save_c64(out, {(size_t)N, (size_t)N, 3}, E.data());
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return finish(2);
  }
  return finish(1);
}
