// scatterview --nogui: load a .cw, evaluate the xy / xz / yz slices on the GPU and save one of them
// as an (N, N, 3) complex64 .npy (ScatterView.cpp:678-795).  Options and defaults follow
// ScatterView.cpp:681-697; like the reference, --nogui exits with status 1 after saving
// (ScatterView.cpp:794).  --cuda -1 selects the reference's CPU evaluator (ScatterView.cpp:363-370,
// cpuEvaluator.cpp:294-407); here that means its region rule (z >= z_l, no tolerance) on device 0, the
// sum still running on the GPU.  The GUI is not part of this implementation.
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
  a = cwx_cli::parse(argc, argv, {{'c', "cuda"}, {'v', "verbose"}}, {"help", "nogui", "verbose", "sample"});
  if (a.has("help")) {
    std::cout << "Allowed options:\n"
                 "  --input arg (=c.cw)     coupled wave structure to evaluate (also positional)\n"
                 "  --help                  produce help message\n"
                 "  -c [ --cuda ] arg (=0)  device number (-1: the CPU evaluator's region rule z >= z_l, on device 0)\n"
                 "  --nogui                 save an output file without loading the GUI\n"
                 "  -v [ --verbose ]        produce verbose output\n"
                 "  --size arg (=100)       size of the sample being visualized\n"
                 "  --resolution arg (=8)   resolution of the field (N = 2^resolution)\n"
                 "  --output arg (=xz.npy)  output file written when the --nogui option is used\n"
                 "  --axis arg (=1)         axis to cut (0 = X, 1 = Y, 2 = Z)\n"
                 "  --center arg (=0 0 0)   center position of the sampled volume\n"
                 "  --slice arg (=0)        coordinate along the specified axis relative to 'center'\n"
              << std::endl;
    return 1;
  }
  auto first = [&](const std::string& k, const std::string& def) {
    return (a.has(k) && !a.vals(k).empty()) ? a.vals(k)[0] : def;
  };
  std::string input = first("input", a.positional.empty() ? "c.cw" : a.positional.back());
  int device = std::stoi(first("cuda", "0"));
  float extent = std::stof(first("size", "100"));
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
  if (!a.has("nogui")) {
    std::cerr << "error: the interactive viewer is not part of this implementation; use --nogui" << std::endl;
    return 1;
  }
  // the reference's CPU path differs from its GPU path only in the region test (cpuEvaluator.cpp:303)
  const int tol_mode = device < 0 ? 1 : 0;
  if (device < 0) device = 0;
  float plane[3] = {0, 0, 0};
  if (axis >= 0 && axis <= 2) plane[axis] = slice;
  const auto t0 = std::chrono::system_clock::now();

  // A compact heterogeneous file (is_volume byte 1, docs/cw_file_format.pptx slide 1) holds mode data
  // instead of the interior waves: rebuild the solution from it and evaluate its FP64 records.
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
  cwx_context* ctx = nullptr;
  cwx_solution* sol = nullptr;
  if (compact) {
    if (cwx_create(device, &ctx) != CWX_OK || cwx_solution_read_cw2(ctx, input.c_str(), &sol) != CWX_OK ||
        cwx_field_init_solution(ctx, sol) != CWX_OK) {
      std::cerr << "error: " << cwx_last_error() << std::endl;
      if (sol) cwx_solution_free(sol);
      if (ctx) cwx_destroy(ctx);
      return 2;
    }
  } else {
    int64_t nw = 0;
    int layers = 0;
    if (cwx_cw_read_unpacked(input.c_str(), &nw, &layers, nullptr, nullptr, nullptr, nullptr) != CWX_OK) {
      std::cout << "ERROR: file " << input << " not found" << std::endl;
      return 1;
    }
    std::vector<cwx_wave_f64> W(nw);
    std::vector<double> zl(layers);
    std::vector<int> wb(layers + 1), we(layers + 1);
    if (cwx_cw_read_unpacked(input.c_str(), &nw, &layers, W.data(), zl.data(), wb.data(), we.data()) != CWX_OK) {
      std::cerr << "error: " << cwx_last_error() << std::endl;
      return 1;
    }
    // cw_unpack (cpuEvaluator.cpp:61-102): the evaluator works on FP32 copies of the waves
    std::vector<cwx_wave_f32> Wf(nw);
    for (int64_t i = 0; i < nw; ++i)
      for (int c = 0; c < 3; ++c) {
        Wf[i].E0[c] = cwx_complexf{(float)W[i].E0[c].re, (float)W[i].E0[c].im};
        Wf[i].k[c] = cwx_complexf{(float)W[i].k[c].re, (float)W[i].k[c].im};
      }
    std::vector<float> zf(zl.begin(), zl.end());
    if (cwx_create(device, &ctx) != CWX_OK ||
        cwx_field_init_f32(ctx, Wf.data(), nw, zf.data(), layers, wb.data(), we.data()) != CWX_OK) {
      std::cerr << "error: " << cwx_last_error() << std::endl;
      if (ctx) cwx_destroy(ctx);
      return 2;
    }
  }
  auto finish = [&](int rc) {
    if (sol) cwx_solution_free(sol);
    cwx_destroy(ctx);
    return rc;
  };
  if (verbose) std::cout << "Using device " << device << " (" << cwx_version() << ")" << std::endl;
  // EvaluateVectorSlices (ScatterView.cpp:354-376)
  const unsigned N = (unsigned)std::pow(2, resolution);
  const float d = extent / (N - 1);
  const float xs = center[0] - extent / 2, ys = center[1] - extent / 2, zs = center[2] - extent / 2;
  std::vector<cwx_complexf> Exy((size_t)N * N * 3), Exz((size_t)N * N * 3), Eyz((size_t)N * N * 3);
  if (cwx_field_evaluate_rule(ctx, Exy.data(), Exz.data(), Eyz.data(), xs, ys, zs, plane[0], plane[1], plane[2], d,
                              (int)N, tol_mode) != CWX_OK) {
    std::cerr << "error: " << cwx_last_error() << std::endl;
    return finish(2);
  }
  const auto t1 = std::chrono::system_clock::now();
  std::cout << "Evaluation (no gui) takes:" << std::chrono::duration<double>(t1 - t0).count() << "s" << std::endl;
  const std::vector<size_t> shape{N, N, 3};
  try {
    if (axis == 1) cwx_npy::save_c64(out, shape, Exz.data());
    else if (axis == 2) cwx_npy::save_c64(out, shape, Exy.data());
    else if (axis == 0) cwx_npy::save_c64(out, shape, Eyz.data());
    else {
      std::cout << "Wrong click at the wrong region. " << std::endl;
      return finish(1);
    }
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return finish(2);
  }
  return finish(1);
}
