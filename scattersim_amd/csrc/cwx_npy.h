// Minimal .npy reader / writer for the command-line tools (NumPy format versions 1.0-3.0).
#pragma once

#include <complex>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace cwx_npy {

struct Array {
  std::vector<size_t> shape;
  std::vector<std::complex<double>> data;  // C order
};

inline std::string header_field(const std::string& h, const std::string& key) {
  size_t p = h.find("'" + key + "'");
  if (p == std::string::npos) throw std::runtime_error("npy header lacks " + key);
  p = h.find(':', p);
  size_t e = p + 1;
  while (e < h.size() && h[e] == ' ') ++e;
  if (h[e] == '(') return h.substr(e, h.find(')', e) - e + 1);
  if (h[e] == '\'') return h.substr(e + 1, h.find('\'', e + 1) - e - 1);
  size_t q = h.find_first_of(",}", e);
  return h.substr(e, q - e);
}

// Loads a 2-D or 3-D array of complex128 / float64 / complex64 / float32 as complex128 in C order.
// Fortran-ordered files are reordered into the logical C-order array.
inline Array load(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open " + path);
  char magic[6];
  f.read(magic, 6);
  if (!f || std::memcmp(magic, "\x93NUMPY", 6) != 0) throw std::runtime_error(path + " is not an .npy file");
  unsigned char ver[2];
  f.read((char*)ver, 2);
  uint32_t hlen = 0;
  if (ver[0] == 1) {
    uint16_t h16;
    f.read((char*)&h16, 2);
    hlen = h16;
  } else {
    f.read((char*)&hlen, 4);
  }
  std::string h(hlen, ' ');
  f.read(&h[0], hlen);
  const std::string descr = header_field(h, "descr");
  const bool fortran = header_field(h, "fortran_order").find("True") != std::string::npos;
  const std::string sh = header_field(h, "shape");
  Array a;
  size_t i = 1;
  while (i < sh.size()) {
    while (i < sh.size() && (sh[i] == ' ' || sh[i] == ',')) ++i;
    if (i >= sh.size() || sh[i] == ')') break;
    size_t j = i;
    while (j < sh.size() && isdigit((unsigned char)sh[j])) ++j;
    a.shape.push_back(std::stoull(sh.substr(i, j - i)));
    i = j;
  }
  size_t n = 1;
  for (size_t d : a.shape) n *= d;
  std::vector<std::complex<double>> raw(n);
  if (descr == "<c16") {
    f.read((char*)raw.data(), 16 * n);
  } else if (descr == "<f8") {
    std::vector<double> t(n);
    f.read((char*)t.data(), 8 * n);
    for (size_t k = 0; k < n; ++k) raw[k] = t[k];
  } else if (descr == "<c8") {
    std::vector<std::complex<float>> t(n);
    f.read((char*)t.data(), 8 * n);
    for (size_t k = 0; k < n; ++k) raw[k] = std::complex<double>(t[k].real(), t[k].imag());
  } else if (descr == "<f4") {
    std::vector<float> t(n);
    f.read((char*)t.data(), 4 * n);
    for (size_t k = 0; k < n; ++k) raw[k] = t[k];
  } else {
    throw std::runtime_error("unsupported npy dtype " + descr);
  }
  if (!f) throw std::runtime_error("truncated .npy file " + path);
  if (!fortran || a.shape.size() < 2) {
    a.data = std::move(raw);
    return a;
  }
  // Fortran order: element (i0, i1, ..., ik) at i0 + s0 (i1 + s1 (...)); rewrite in C order
  a.data.resize(n);
  const size_t nd = a.shape.size();
  std::vector<size_t> idx(nd, 0);
  for (size_t c = 0; c < n; ++c) {
    size_t fo = 0, mul = 1;
    for (size_t d = 0; d < nd; ++d) { fo += idx[d] * mul; mul *= a.shape[d]; }
    a.data[c] = raw[fo];
    for (size_t d = nd; d-- > 0;) {
      if (++idx[d] < a.shape[d]) break;
      idx[d] = 0;
    }
  }
  return a;
}

// Saves data of the given shape in C order; descr names the element type ('<c8' complex<float>,
// '<c16' complex<double>), esize its size in bytes.
inline void save(const std::string& path, const std::vector<size_t>& shape, const void* data, const char* descr,
                 size_t esize) {
  std::string dict = std::string("{'descr': '") + descr + "', 'fortran_order': False, 'shape': (";
  size_t n = 1;
  for (size_t i = 0; i < shape.size(); ++i) {
    dict += std::to_string(shape[i]);
    dict += (shape.size() == 1 || i + 1 < shape.size()) ? ", " : "";
    n *= shape[i];
  }
  dict += "), }";
  size_t total = 10 + dict.size() + 1;
  size_t pad = (64 - total % 64) % 64;
  dict += std::string(pad, ' ') + "\n";
  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot write " + path);
  f.write("\x93NUMPY\x01\x00", 8);
  uint16_t hl = (uint16_t)dict.size();
  f.write((const char*)&hl, 2);
  f.write(dict.data(), dict.size());
  f.write((const char*)data, esize * n);
}
inline void save_c64(const std::string& path, const std::vector<size_t>& shape, const void* data) {
  save(path, shape, data, "<c8", 8);
}
inline void save_c128(const std::string& path, const std::vector<size_t>& shape, const void* data) {
  save(path, shape, data, "<c16", 16);
}

}  // namespace cwx_npy
