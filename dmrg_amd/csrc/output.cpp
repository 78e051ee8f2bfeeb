#include "output.hpp"

#include <cmath>
#include <cstdlib>
#include <ios>
#include <sstream>

namespace dmrgx {

namespace {
constexpr int ACC = 10;      // global.h:20
constexpr int NEV_OUT = 6;   // global.h:21
constexpr int NSI_OUT = 20;  // global.h:22
constexpr double TOL_OUT = 1e-10;

struct Fixed {
  std::ostream& o;
  std::ios::fmtflags f;
  std::streamsize p;
  explicit Fixed(std::ostream& s) : o(s), f(s.flags()), p(s.precision()) {
    o.precision(ACC);
    o.setf(std::ios::fixed);
    o.setf(std::ios::showpoint);
  }
  ~Fixed() {
    o.flags(f);
    o.precision(p);
  }
};
}  // namespace

void write_energy_line(std::ostream& o, int l, int r, const double* val, int n) {
  Fixed fx(o);
  o << l << "\t" << r;
  if (NEV_OUT <= n) {
    for (int i = 0; i < NEV_OUT; ++i) o << "\t" << val[i];
  } else {
    for (int i = 0; i < n; ++i) o << "\t" << val[i];
    for (int i = std::max(n, 0); i < NEV_OUT; ++i) o << "\t" << 0;
  }
  o << "\n";
}

void write_singular_line(std::ostream& o, int l, int r, const std::vector<double>& s, int cutoff) {
  Fixed fx(o);
  double entropy = 0;
  for (double x : s) entropy += -x * x * std::log(x * x + TOL_OUT);
  o << l << "\t" << r << "\t" << entropy;
  const int ns = (int)s.size();
  if (ns >= NSI_OUT) {
    for (int i = 0; i < NSI_OUT; ++i) o << "\t" << s[i];
  } else {
    for (int i = 0; i < ns; ++i) o << "\t" << s[i];
    for (int i = ns; i < NSI_OUT; ++i) o << "\t" << 0;
  }
  double error = 1;
  const int c = ns < cutoff ? ns : cutoff;
  for (int i = 0; i < c; ++i) error -= s[i] * s[i];
  o << "\t" << error << "\n";
}

Params set_para_val(int argc, const char* const* argv, const std::vector<std::string>& names) {
  Params p;
  p.values.assign(names.size(), 0.0);
  if (!p.values.empty()) p.values[0] = 1;
  for (int i = 0; i < argc; ++i) {
    const std::string a(argv[i]);
    const size_t eq = a.find('=');
    // split at the first '='; a word without one re-applies the previous
    // assignment in the reference (src/functions.cpp:24-42), i.e. changes nothing
    if (eq == std::string::npos) continue;
    const std::string key = a.substr(0, eq), val = a.substr(eq + 1);
    const double v = std::atof(val.c_str());
    if (key == "l") p.sites = (int)v;
    if (key == "m") p.cutoff = (int)v;
    if (key == "n") p.sweep = (int)v;
    if (key == "s") p.sector = (int)v;
    for (size_t j = 0; j < names.size(); ++j)
      if (key == names[j]) p.values[j] = v;
  }
  std::stringstream ss;
  ss << "_L_" << p.sites << "_m_" << p.cutoff << "_n_" << p.sweep << "_s_" << p.sector;
  for (size_t i = 0; i < names.size(); ++i) ss << "_" << names[i] << "_" << p.values[i];
  p.suffix = ss.str() + "_.txt";
  return p;
}

}  // namespace dmrgx
