// scattervolume / scatterthinvolume: the reference's command-line surface (ScatterVolume.cpp:506-1019)
// on top of libcwx.  Options and defaults follow ScatterVolume.cpp:510-530; --help prints them and
// returns 1 (ScatterVolume.cpp:538-541).  Solver failures exit with status 2 and a message (the
// reference ignores LAPACK's info, third_Lapack.h:43,53,62).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <fstream>
#include <iostream>
#include <string>
#include <vector>

#include "../../include/cwx.h"
#include "cli_args.h"
#include "cwx_npy.h"

namespace {

const char* kHelp =
    "Allowed options:\n"
    "  --help                 produce help message\n"
    "  --sample arg           input sample as an .npy file\n"
    "  --lambda arg (=1)      incident field vacuum wavelength\n"
    "  --direction arg (=0, 0, 1)  incoming field direction\n"
    "  --ex arg (=0, 0)       x component of the electrical field\n"
    "  --ey arg (=1, 0)       y component of the electrical field\n"
    "  --ez arg (=0 0)        z component of the electrical field\n"
    "  --n arg (=1, 1)        real refractive index (optical path length) of the upper and lower layers\n"
    "  --kappa arg (=0.00)    absorbance of the lower layer (upper layer is always 0.0)\n"
    "  --size arg (=10, 10, 4) The real size of the single-layer sample\n"
    "  --z arg (=0.0)         the center of the sample along z-axis\n"
    "  --output arg (=c.cw)   output filename for the coupled wave structure\n"
    "  --coef arg             number of Fourier coefficients (can be specified in 2 dimensions)\n"
    "  --psf                  generate the point spread function(PSF) for the optic system\n"
    "  --external             save waves for visualization of external field only\n"
    "  --log                  produce a log file\n"
    "  --npz                  calculate Px and Py only\n"
    "  --reduced              (cwx) solve the 2MF eigenproblem of P Q instead of the 4MF Phi\n"
    "  --no-cutoff            (cwx) keep the requested orders (skip the order cut-off heuristic)\n"
    "  --compact              (cwx) write the compact heterogeneous .cw (mode data per layer)\n"
    "  --device arg (=0)      (cwx) HIP device\n";

std::vector<double> nums(const cwx_cli::Args& a, const std::string& k, std::vector<double> def) {
  if (!a.has(k)) return def;
  std::vector<double> v;
  for (auto& s : a.vals(k)) v.push_back(std::stod(s));
  return v;
}

}  // namespace

int main(int argc, char** argv) {
  const auto t0 = std::chrono::steady_clock::now();
  cwx_cli::Args a;
  try {
    a = cwx_cli::parse(argc, argv, {}, {"help", "psf", "external", "log", "npz", "reduced", "no-cutoff", "compact"});
  } catch (const std::exception& e) {
    std::cerr << e.what() << std::endl;
    return 1;
  }
  if (a.has("help")) {
    std::cout << kHelp << std::endl;
    return 1;
  }
  cwx_volume_params p;
  cwx_volume_params_default(&p);
  std::string sample, output = "c.cw";
  int device = 0;
  std::ofstream logfile;
  try {
    if (a.has("sample") && !a.vals("sample").empty()) sample = a.vals("sample")[0];
    if (a.has("output")) output = a.vals("output").empty() ? "" : a.vals("output")[0];
    p.lambda = nums(a, "lambda", {1.0})[0];
    auto dir = nums(a, "direction", {0, 0, 1});
    auto ex = nums(a, "ex", {0, 0}), ey = nums(a, "ey", {1, 0}), ez = nums(a, "ez", {0, 0});
    auto n = nums(a, "n", {1, 1});
    auto kap = nums(a, "kappa", {0});
    auto size = nums(a, "size", {10, 10, 2});
    if (dir.size() != 3 || size.size() != 3 || ex.size() != 2 || ey.size() != 2 || ez.size() != 2 || kap.empty())
      throw std::runtime_error("--direction/--size take 3 values, --ex/--ey/--ez take (re, im) pairs");
    if (n.size() != 2) throw std::runtime_error("--n takes the indices of the two half-spaces");
    for (int i = 0; i < 3; ++i) { p.direction[i] = dir[i]; p.size[i] = size[i]; }
    for (int i = 0; i < 2; ++i) { p.ex[i] = ex[i]; p.ey[i] = ey[i]; p.ez[i] = ez[i]; p.n[i] = n[i]; }
    p.kappa = kap[0];
    p.z = nums(a, "z", {0.0})[0];
    if (a.has("coef")) {
      auto c = nums(a, "coef", {});
      if (c.size() > 2) throw std::runtime_error("--coef takes 0, 1 or 2 values");
      p.ncoef = (int)c.size();
      for (size_t i = 0; i < c.size(); ++i) p.coef[i] = (int)c[i];
    }
    p.psf = a.has("psf");
    p.external = a.has("external");
    p.npz = a.has("npz");
    p.eig_mode = a.has("reduced") ? 1 : 0;
    p.no_cutoff = a.has("no-cutoff");
    device = (int)nums(a, "device", {0})[0];
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
  if (a.has("log")) {
    logfile.open(std::to_string((long)std::time(nullptr)) + "_scattervolume.log");
    for (int c = 0; c < argc; ++c) logfile << argv[c] << " ";
    logfile << std::endl;
  }
  if (sample.empty()) {
    std::cerr << "error: --sample is required" << std::endl;
    return 1;
  }
  cwx_npy::Array arr;
  try {
    arr = cwx_npy::load(sample);
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
  int L = 1, s1, s2;
  if (arr.shape.size() == 3) { L = (int)arr.shape[0]; s1 = (int)arr.shape[1]; s2 = (int)arr.shape[2]; }
  else if (arr.shape.size() == 2) { s1 = (int)arr.shape[0]; s2 = (int)arr.shape[1]; }
  else { std::cerr << "error: the sample must be a 2-D or 3-D array" << std::endl; return 1; }

  cwx_context* ctx = nullptr;
  if (cwx_create(device, &ctx) != CWX_OK) {
    std::cerr << "error: " << cwx_last_error() << std::endl;
    return 2;
  }
  cwx_solution* sol = nullptr;
  const auto t1 = std::chrono::steady_clock::now();
  int st = cwx_solve_volume(ctx, &p, (const cwx_complex*)arr.data.data(), L, s1, s2, &sol);
  const auto t2 = std::chrono::steady_clock::now();
  if (st != CWX_OK) {
    std::cerr << "error: solve failed (" << st << "): " << cwx_last_error() << std::endl;
    if (sol) cwx_solution_free(sol);
    cwx_destroy(ctx);
    return 2;
  }
  cwx_solution_info info;
  cwx_solution_get_info(sol, &info);
  if (info.cutoff) {
    std::cout << "Cutting off invalid waves with imagnery sz..." << std::endl;
    std::cout << "[WARNING] Propagation directions for decomposed waves are not all downward. We suggest to "
                 "increase in_size or decrease the wavelength to tolerate higher Fourier coefficients."
              << std::endl;
  }
  std::cout << "Field saved in " << output << "." << std::endl;
  std::cout << "Number of pixels (x, y): [" << s2 << "," << s1 << "]" << std::endl;
  std::cout << "Number of sublayers: " << L << std::endl;
  std::cout << "Number of Fourier coefficients (Mx, My): [" << info.M0 << "," << info.M1 << "]" << std::endl;
  std::cout << "Effective number of Fourier coefficients: [" << info.MF << "]" << std::endl;
  if (!output.empty()) {
    const int wst = a.has("compact") ? cwx_solution_write_cw2(ctx, sol, output.c_str())
                                     : cwx_solution_write_cw(ctx, sol, output.c_str());
    if (wst != CWX_OK) {
      std::cerr << "error: " << cwx_last_error() << std::endl;
      cwx_solution_free(sol);
      cwx_destroy(ctx);
      return 2;
    }
  }
  const auto t3 = std::chrono::steady_clock::now();
  if (logfile) {
    auto sec = [](auto d) { return std::chrono::duration<double>(d).count(); };
    logfile << "Time to load the sample: " << sec(t1 - t0) << "s" << std::endl;
    logfile << "Time to solve (device): " << sec(t2 - t1) << "s" << std::endl;
    logfile << "Time to save the field " << sec(t3 - t2) << "s" << std::endl;
    logfile << "Total time:" << sec(t3 - t0) << "s" << std::endl;
  }
  cwx_solution_free(sol);
  cwx_destroy(ctx);
  return 0;
}
