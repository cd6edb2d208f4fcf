// scatterlayer: a focused beam (or one plane wave) through a stack of homogeneous layers
// (ScatterLayer.cpp:281-535) on top of libcwx.  Options and defaults follow ScatterLayer.cpp:284-301:
// --help prints them and returns 1, and there are no single-letter options, so negative numbers parse
// as values.  Every incident wave is solved in one call (cwx_layer_solve); the .cw output is the
// reference's list layout (all incident waves, then per boundary its reflected and transmitted waves).
// Differences: the per-incidence dump of the solution vector (ScatterLayer.cpp:401) is not printed,
// --device selects the HIP device and --incident takes the incident waves from an (P, 6) complex .npy.
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdio>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/cwx.h"
#include "cli_args.h"
#include "cwx_npy.h"

namespace {

const char* kHelp =
    "Allowed options:\n"
    "  --help                       produce help message\n"
    "  --lambda arg (=1)            incident field vacuum wavelength\n"
    "  --direction arg (=0, 0, 1)   incoming field direction\n"
    "  --ex arg (=0, 0)             x component of the electrical field\n"
    "  --ey arg (=1, 0)             y component of the electrical field\n"
    "  --ez arg (=0 0)              z component of the electrical field\n"
    "  --n arg (=1, 1.4, 1.4, 1.0)  real refractive index (optical path length) of all L layers\n"
    "  --kappa arg (=0.05, 0.00, 0.00) absorbance of layers 2+ (layer 1 is always 0.0)\n"
    "  --z arg (=-3.0, 0.0, 3.0)    position of each layer boundary\n"
    "  --output_sample arg (=sample.npy) output the sample as the input for scattervolume\n"
    "  --output arg (=a.cw)         output filename for the coupled wave structure\n"
    "  --alpha arg (=1)             angle used to focus the incident field\n"
    "  --beta arg (=0)              internal obscuration angle (for simulating reflective optics)\n"
    "  --na arg                     focus angle expressed as a numerical aperture (overrides --alpha)\n"
    "  --samples arg (=64 64)       number of samples (can be specified in 2 dimensions)\n"
    "  --mode arg (=polar)          sampling mode (polar, montecarlo)\n"
    "  --log                        produce a log file\n"
    "  --incident arg               (cwx) incident waves as a (P, 6) complex .npy: E0.xyz, k.xyz\n"
    "  --device arg (=0)            (cwx) HIP device\n";

std::vector<double> nums(const cwx_cli::Args& a, const std::string& k, std::vector<double> def) {
  if (!a.has(k)) return def;
  std::vector<double> v;
  for (auto& s : a.vals(k)) v.push_back(std::stod(s));
  return v;
}

std::string str(const cwx_cli::Args& a, const std::string& k, const std::string& def) {
  if (!a.has(k)) return def;
  return a.vals(k).empty() ? std::string() : a.vals(k)[0];
}

using cplx = std::complex<double>;

std::string vec3(const cplx* v, int w) {
  std::stringstream ss;
  const bool real = v[0].imag() == 0.0 && v[1].imag() == 0.0 && v[2].imag() == 0.0;
  for (int i = 0; i < 3; ++i) {
    if (real) ss << std::setw(w) << std::left << v[i].real();
    else ss << std::setw(w) << std::left << v[i];
  }
  return ss.str();
}

cplx at(const cwx_complex& c) { return cplx(c.re, c.im); }

}  // namespace

int main(int argc, char** argv) {
  cwx_cli::Args a;
  try {
    a = cwx_cli::parse(argc, argv, {}, {"help", "log"});
  } catch (const std::exception& e) {
    std::cerr << e.what() << std::endl;
    return 1;
  }
  if (a.has("help")) {
    std::cout << kHelp << std::endl;
    return 1;
  }
  double lambda, alpha, beta;
  std::vector<double> dir, ex, ey, ez, n, kappa, z, samples;
  std::string output, output_sample, mode, incident_file;
  int device = 0;
  try {
    lambda = nums(a, "lambda", {1.0})[0];
    dir = nums(a, "direction", {0, 0, 1});
    ex = nums(a, "ex", {0, 0});
    ey = nums(a, "ey", {1, 0});
    ez = nums(a, "ez", {0, 0});
    n = nums(a, "n", {1, 1.4, 1.4, 1.0});
    kappa = nums(a, "kappa", {0.05});
    z = nums(a, "z", {-5.0, 0.0, 5.0});
    output_sample = str(a, "output_sample", "sample.npy");
    output = str(a, "output", "a.cw");
    alpha = nums(a, "alpha", {1.0})[0];
    beta = nums(a, "beta", {0.0})[0];
    samples = nums(a, "samples", {64, 64});
    mode = str(a, "mode", "polar");
    incident_file = str(a, "incident", "");
    device = (int)nums(a, "device", {0})[0];
    if (dir.size() != 3 || ex.size() != 2 || ey.size() != 2 || ez.size() != 2)
      throw std::runtime_error("--direction takes 3 values, --ex/--ey/--ez take (re, im) pairs");
    if (n.empty() || kappa.empty() || z.empty() || samples.empty())
      throw std::runtime_error("--n, --kappa, --z and --samples need at least one value");
    if (mode != "polar" && mode != "montecarlo") throw std::runtime_error("--mode is polar or montecarlo");
  } catch (const std::exception& e) {
    std::cerr << "error: " << e.what() << std::endl;
    return 1;
  }
  std::ofstream logfile;
  if (a.has("log")) logfile.open(std::to_string((long)std::time(nullptr)) + "_scatterlayer.log");
  if (a.has("na")) alpha = std::asin(nums(a, "na", {0.0})[0] / n[0]);  // ScatterLayer.cpp:324-326

  // layer count and padding (ScatterLayer.cpp:327-344)
  size_t L = n.size();
  if (kappa.size() + 1 > L) L = kappa.size() + 1;
  if (z.size() + 1 > L) L = z.size() + 1;
  n.resize(L, n.back());
  kappa.resize(L - 1, 0.0);
  while (z.size() + 1 < L) z.push_back(z.back() + 10.0);
  z.resize(L - 1);

  // reference wave (ScatterLayer.cpp:346-353): orthogonalised E, direction scaled by n0
  double d[3] = {dir[0], dir[1], dir[2]};
  const double dn = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
  if (dn == 0.0) {
    std::cerr << "error: --direction must be non-zero" << std::endl;
    return 1;
  }
  for (double& v : d) v /= dn;
  cplx e[3] = {cplx(ex[0], ex[1]), cplx(ey[0], ey[1]), cplx(ez[0], ez[1])};
  {
    const cplx mag = std::sqrt(e[0] * e[0] + e[1] * e[1] + e[2] * e[2]);
    for (cplx& v : e) v /= mag;
    const cplx dc[3] = {d[0], d[1], d[2]};
    const cplx s[3] = {e[1] * dc[2] - e[2] * dc[1], e[2] * dc[0] - e[0] * dc[2], e[0] * dc[1] - e[1] * dc[0]};
    e[0] = dc[1] * s[2] - dc[2] * s[1];
    e[1] = dc[2] * s[0] - dc[0] * s[2];
    e[2] = dc[0] * s[1] - dc[1] * s[0];
  }
  const double kvac = 2.0 * M_PI / lambda;
  const double k3[3] = {d[0] * n[0] * kvac, d[1] * n[0] * kvac, d[2] * n[0] * kvac};
  const cwx_complex E3[3] = {{e[0].real(), e[0].imag()}, {e[1].real(), e[1].imag()}, {e[2].real(), e[2].imag()}};

  // incident waves (ScatterLayer.cpp:355-385)
  unsigned N[2];
  if (samples.size() == 1) {
    if (mode == "montecarlo") { N[0] = (unsigned)samples[0]; N[1] = 1; }
    else N[0] = N[1] = (unsigned)std::sqrt(samples[0]);
  } else {
    N[0] = (unsigned)samples[0];
    N[1] = (unsigned)samples[1];
  }
  if (alpha == 0) N[0] = N[1] = 1;
  std::vector<cwx_wave_f64> I;
  if (!incident_file.empty()) {
    try {
      cwx_npy::Array arr = cwx_npy::load(incident_file);
      if (arr.shape.size() != 2 || arr.shape[1] != 6) throw std::runtime_error("--incident must be (P, 6)");
      I.resize(arr.shape[0]);
      for (size_t p = 0; p < I.size(); ++p)
        for (int c = 0; c < 3; ++c) {
          I[p].E0[c] = cwx_complex{arr.data[p * 6 + c].real(), arr.data[p * 6 + c].imag()};
          I[p].k[c] = cwx_complex{arr.data[p * 6 + 3 + c].real(), arr.data[p * 6 + 3 + c].imag()};
        }
    } catch (const std::exception& ex2) {
      std::cerr << "error: " << ex2.what() << std::endl;
      return 1;
    }
  } else if (alpha == 0 || N[0] * N[1] == 1) {
    I.resize(1);
    for (int c = 0; c < 3; ++c) {
      I[0].E0[c] = E3[c];
      I[0].k[c] = cwx_complex{k3[c], 0.0};
    }
  } else {
    int st;
    if (mode == "montecarlo") {
      I.resize((size_t)N[0] * N[1]);
      st = cwx_beam_montecarlo(alpha, beta, k3, E3, (int64_t)I.size(), 0, I.data());
    } else {
      I.resize((size_t)N[0] * N[1]);
      st = cwx_beam_polar(alpha, beta, k3, E3, (int)N[0], (int)N[1], I.data());
    }
    if (st != CWX_OK) {
      std::cerr << "error: " << cwx_last_error() << std::endl;
      return 1;
    }
  }
  const int64_t P = (int64_t)I.size();

  cwx_context* ctx = nullptr;
  if (cwx_create(device, &ctx) != CWX_OK) {
    std::cerr << "error: " << cwx_last_error() << std::endl;
    return 2;
  }
  std::vector<cwx_wave_f64> R((size_t)P * (L - 1)), T((size_t)P * (L - 1));
  const auto t0 = std::chrono::steady_clock::now();
  int st = cwx_layer_solve(ctx, (int)L, n.data(), kappa.data(), z.data(), lambda, P, I.data(), R.data(), T.data());
  const auto t1 = std::chrono::steady_clock::now();
  if (st < 0) {
    std::cerr << "error: " << cwx_last_error() << std::endl;
    cwx_destroy(ctx);
    return 2;
  }
  if (st > 0) std::cerr << "warning: " << cwx_last_error() << std::endl;
  if (logfile) logfile << "solved " << P << " incident waves in " << std::chrono::duration<double>(t1 - t0).count()
                       << " s" << std::endl;
  if (!output.empty() &&
      cwx_layer_write_cw(output.c_str(), (int)L, z.data(), P, I.data(), R.data(), T.data()) != CWX_OK) {
    std::cerr << "error: " << cwx_last_error() << std::endl;
    cwx_destroy(ctx);
    return 2;
  }
  // the layer indices as a scattervolume sample: ri grows by L entries per incident wave
  // (ScatterLayer.cpp:118-125 runs once per wave) and entries 1 .. size-2 are written (426-434)
  if (!output_sample.empty()) {
    std::vector<cplx> ri;
    for (int64_t p = 0; p < P; ++p) {
      ri.push_back(cplx(n[0], 0.0));
      for (size_t l = 1; l < L; ++l) ri.push_back(cplx(n[l], kappa[l - 1]));
    }
    const size_t cnt = ri.size() >= 2 ? ri.size() - 2 : 0;
    try {
      cwx_npy::save_c128(output_sample, {cnt, 1, 1}, ri.data() + 1);
    } catch (const std::exception& ex2) {
      std::cerr << "error: " << ex2.what() << std::endl;
      cwx_destroy(ctx);
      return 2;
    }
  }

  // summary of the reference wave (ScatterLayer.cpp:436-535)
  std::vector<cwx_wave_f64> Ir(1), Rr(L - 1), Tr(L - 1);
  for (int c = 0; c < 3; ++c) {
    Ir[0].E0[c] = E3[c];
    Ir[0].k[c] = cwx_complex{k3[c], 0.0};
  }
  cwx_layer_solve(ctx, (int)L, n.data(), kappa.data(), z.data(), lambda, 1, Ir.data(), Rr.data(), Tr.data());
  const int nw = 20, num = 10;
  std::cout << std::setw(nw) << std::left << "vacuum lambda:" << std::setw(num) << std::left << lambda << std::endl;
  if (alpha != 0.0) {
    std::cout << std::setw(nw) << std::left << "focusing angle:" << std::setw(num) << alpha << " ("
              << std::sin(alpha) * n[0] << " NA)" << std::endl;
    if (beta > 0.0)
      std::cout << std::setw(nw) << std::left << "obscuration angle:" << std::setw(num) << std::left << beta << " ("
                << std::sin(beta) * n[0] << " NA)" << std::endl;
  }
  std::cout << std::endl << std::setw(nw) << std::left << "layers: " << L << std::endl;
  for (size_t l = 0; l < L; ++l) {
    const float zs = l == 0 ? -INFINITY : (float)z[l - 1], ze = l == L - 1 ? INFINITY : (float)z[l];
    std::stringstream rng;
    rng << "[" << zs << ", " << ze << "]:";
    std::cout << std::setw(4) << std::right << l << std::setw(nw - 4) << std::left << rng.str() << n[l] << " + "
              << (l == 0 ? 0.0 : kappa[l - 1]) << "i" << std::endl;
  }
  std::cout << std::endl;
  std::cout << std::setw(nw) << std::left << "samples: " << N[0] << " x " << N[1] << " = " << N[0] * N[1] << std::endl;
  std::cout << std::setw(nw) << std::left << "sampling mode: " << mode << std::endl << std::endl;
  cplx kc[3] = {k3[0], k3[1], k3[2]}, ec[3] = {e[0], e[1], e[2]};
  std::cout << std::setw(nw) << std::left << "vvvvvvvv    k:" << vec3(kc, num) << std::endl;
  std::cout << std::setw(nw) << std::left << "vvvvvvvv  |k|:" << n[0] * kvac << std::endl;
  std::cout << std::setw(nw) << std::left << "vvvvvvvv E(0):" << vec3(ec, num) << std::endl << std::endl;
  for (size_t l = 0; l + 1 < L; ++l) {
    cplx rk[3], rE[3], tk[3], tE[3];
    for (int c = 0; c < 3; ++c) {
      rk[c] = at(Rr[l].k[c]);
      rE[c] = at(Rr[l].E0[c]);
      tk[c] = at(Tr[l].k[c]);
      tE[c] = at(Tr[l].E0[c]);
    }
    std::cout << std::setw(nw) << std::left << "^^^^^^^^    k:" << vec3(rk, num) << std::endl;
    std::cout << std::setw(nw) << std::left << "^^^^^^^^ E(0):" << vec3(rE, num) << std::endl << std::endl;
    std::cout << "n = " << n[l] << " + " << (l == 0 ? 0.0 : kappa[l - 1]) << "i=========================" << std::endl;
    std::cout << "n = " << n[l + 1] << " + " << kappa[l] << "i=========================" << std::endl << std::endl;
    std::cout << std::setw(nw) << std::left << "vvvvvvvv    k:" << vec3(tk, num) << std::endl;
    std::cout << std::setw(nw) << std::left << "vvvvvvvv E(0):" << vec3(tE, num) << std::endl << std::endl << std::endl;
  }
  cwx_destroy(ctx);
  return 0;
}
