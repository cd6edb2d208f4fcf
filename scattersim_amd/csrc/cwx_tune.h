// Launch-shape choices of the kernels: a table written by the project's tuner
// (python -m scattersim_amd.tune), overridable per process with cwx_set_param.
//
// The table is plain text, one rule per line: `key a_lo a_hi b_lo b_hi value`, `#` starting a comment.
// A lookup of (key, a, b) returns the value of the first rule of that key with a_lo <= a <= a_hi and
// b_lo <= b <= b_hi.  The file is data/tuning.txt next to the library's directory (lib/../data), or
// the path in CWX_TUNING_FILE; without one, every lookup falls back to the built-in choice.
#pragma once

#include <string>

namespace cwx {

// value of the first matching rule, else `fallback`
long tune_lookup(const std::string& key, long a, long b, long fallback);
// process-wide overrides (cwx_set_param); `set` false clears the key
void tune_override(const std::string& key, long value, bool set);
// path of the table the library loaded ("" when none)
std::string tune_table_path();

}  // namespace cwx
