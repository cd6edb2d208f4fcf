// Small "--name v1 v2 ..." option parser for the command-line tools.  Values may be negative
// numbers (the reference disables short options for scattervolume so that "-1" parses as a value,
// ScatterVolume.cpp:531-535).
#pragma once

#include <cstdio>
#include <map>
#include <set>
#include <string>
#include <vector>

namespace cwx_cli {

struct Args {
  std::map<std::string, std::vector<std::string>> opts;
  std::vector<std::string> positional;
  bool has(const std::string& k) const { return opts.count(k) != 0; }
  const std::vector<std::string>& vals(const std::string& k) const { return opts.at(k); }
};

inline bool looks_numeric(const std::string& s) {
  if (s.size() < 2 || s[0] != '-') return false;
  return (s[1] >= '0' && s[1] <= '9') || s[1] == '.';
}

// shorts: single-letter aliases ('c' -> "cuda"); flags: options that take no value.
inline Args parse(int argc, char** argv, const std::map<char, std::string>& shorts = {},
                  const std::set<std::string>& flags = {}) {
  Args a;
  std::string cur;
  bool in_opt = false;
  for (int i = 1; i < argc; ++i) {
    std::string t = argv[i];
    if (t.size() > 2 && t[0] == '-' && t[1] == '-') {
      cur = t.substr(2);
      size_t eq = cur.find('=');
      a.opts[cur.substr(0, eq)];
      if (eq != std::string::npos) {
        a.opts[cur.substr(0, eq)].push_back(cur.substr(eq + 1));
        cur = cur.substr(0, eq);
      }
      in_opt = !flags.count(cur);
    } else if (t.size() == 2 && t[0] == '-' && !looks_numeric(t) && shorts.count(t[1])) {
      cur = shorts.at(t[1]);
      a.opts[cur];
      in_opt = !flags.count(cur);
    } else if (in_opt) {
      a.opts[cur].push_back(t);
    } else {
      a.positional.push_back(t);
    }
  }
  return a;
}

}  // namespace cwx_cli
