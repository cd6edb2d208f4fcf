// Tuning table of libcwx (see cwx_tune.h).
#include "cwx_tune.h"

#include <dlfcn.h>

#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <vector>

namespace cwx {
namespace {

struct Rule {
  long a_lo, a_hi, b_lo, b_hi, value;
};

struct Table {
  std::once_flag once;
  std::mutex mu;
  std::map<std::string, std::vector<Rule>> rules;
  std::map<std::string, long> overrides;
  std::string path;
};

Table& table() {
  static Table t;
  return t;
}

std::string library_dir() {
  Dl_info info{};
  if (dladdr(reinterpret_cast<void*>(&tune_lookup), &info) == 0 || !info.dli_fname) return "";
  std::string p(info.dli_fname);
  const size_t s = p.find_last_of('/');
  return s == std::string::npos ? std::string(".") : p.substr(0, s);
}

void load(Table& t) {
  std::string path;
  if (const char* e = std::getenv("CWX_TUNING_FILE")) path = e;
  else path = library_dir() + "/../data/tuning.txt";
  std::ifstream f(path);
  if (!f) return;
  t.path = path;
  std::string line;
  while (std::getline(f, line)) {
    const size_t h = line.find('#');
    if (h != std::string::npos) line.resize(h);
    std::istringstream is(line);
    std::string key;
    Rule r{};
    if (is >> key >> r.a_lo >> r.a_hi >> r.b_lo >> r.b_hi >> r.value) t.rules[key].push_back(r);
  }
}

}  // namespace

long tune_lookup(const std::string& key, long a, long b, long fallback) {
  Table& t = table();
  std::call_once(t.once, [&]() { load(t); });
  std::lock_guard<std::mutex> lk(t.mu);
  const auto o = t.overrides.find(key);
  if (o != t.overrides.end()) return o->second;
  const auto it = t.rules.find(key);
  if (it == t.rules.end()) return fallback;
  for (const Rule& r : it->second)
    if (a >= r.a_lo && a <= r.a_hi && b >= r.b_lo && b <= r.b_hi) return r.value;
  return fallback;
}

void tune_override(const std::string& key, long value, bool set) {
  Table& t = table();
  std::call_once(t.once, [&]() { load(t); });
  std::lock_guard<std::mutex> lk(t.mu);
  if (set) t.overrides[key] = value;
  else t.overrides.erase(key);
}

std::string tune_table_path() {
  Table& t = table();
  std::call_once(t.once, [&]() { load(t); });
  return t.path;
}

}  // namespace cwx
