// The reference's text outputs and command-line convention, host only (no
// device code, so this pair of files also builds with a plain C++ compiler):
// print_energy / print_singular (src/useful.cpp:54-102) and set_para_val
// (src/functions.cpp:46-80).
#pragma once

#include <ostream>
#include <string>
#include <vector>

namespace dmrgx {

// "l \t r \t val[0..NEV)" in fixed notation with 10 decimals; when n < 6
// the missing values are written as the integer 0 (print_energy)
void write_energy_line(std::ostream& o, int l, int r, const double* val, int n);
// "l \t r \t entropy \t s[0..20) \t error" (print_singular): entropy
// -sum s^2 log(s^2 + 1e-10) and error 1 - sum_{i < min(cutoff, |s|)} s_i^2
// over all singular values s (descending); missing s written as 0
void write_singular_line(std::ostream& o, int l, int r, const std::vector<double>& s, int cutoff);

// set_para_val: reads "key=value" arguments (l, m, n, s and the model's
// parameter names; the first parameter defaults to 1, the others to 0, l to
// 4, m to 10, n to -1, s to 0 -- the mains' defaults) and returns the file
// suffix "_L_<l>_m_<m>_n_<n>_s_<s>_<name>_<value>..._.txt" with the values in
// the default stream format (6 significant digits)
struct Params {
  int sites = 4, cutoff = 10, sweep = -1, sector = 0;
  std::vector<double> values;
  std::string suffix;
};
Params set_para_val(int argc, const char* const* argv, const std::vector<std::string>& names);

}  // namespace dmrgx
