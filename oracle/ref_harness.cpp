// Harness around the reference's own sources (test infrastructure only).
//
// Built by oracle/Makefile from /root/reference/src/{qtensor,tensor,useful,
// operators}.cpp with the reference's shipped global.h (SYMMETRY 2, FERMION)
// into oracle/_ref/ref_harness, and once more with the build switches of
// its 4e main (SYMMETRY 8, S4E, PBC; 4e.cpp:28-32) into
// oracle/_ref/ref_harness_s4e, which adds 4e.cpp's builders; see the
// Makefile for why the numerical wrappers are left unlinked.  It reads one
// case on stdin and prints the reference's results on stdout, for
// tests/golden/make_ref_vectors.py:
//
//   input:   model <kitaev|xyz|ising|potts2> p0 p1 p2 p3
//            (S4E build: model <4e1|4e2> t1 t3 f0 f1 p1 frac)
//            lstore <tensor>      the left environment store(l-1) [a, w, a']
//            rstore <tensor>      the right environment store(r+1) [b, w, b']
//            sector <s>
//   tensor:  nleg, per leg (nsec, dims...), nblk, nblk x nleg sectors,
//            nleg dirs, then the values of every block in block order
//            (real, or re im pairs for complex models)
//
//   output:  W <tensor>           the bulk MPO tensor (src/operators.cpp)
//            Wleft / Wright       mpo edges (include/mpo.h:45-48, src/qtensor.cpp:449-463)
//            lenv / renv          update_two's environments (src/core.cpp:188-192)
//            dim, sym             generate_dim_sym (src/useful.cpp:151-194)
//            rmap, dim_ret, sym_ret   renv.get_map(..., r_num=2, back=false)
//            lmap, ldim            lenv.get_map(..., l_num=3, back=true)
//                                  (src/core.cpp:236-240, src/qtensor.cpp:997-1119)
//            ledge / redge         (PBC build) the edge environments of
//                                  update_two with lstore / rstore as the
//                                  stored edges: the identity operator
//                                  id(W) (src/qtensor.cpp:157-182) and the
//                                  contractions of src/core.cpp:193-199
//            ops                   (S4E build) the bond-1 operators of the
//                                  4e main, c_up1 .. c_pair_dagger2 and
//                                  c_pair1().id() (4e.cpp:587-779, 93)
#include <cstdio>
#include <cstdlib>
#include <sstream>
#include <string>

#include "global.h"
#include "mpo.h"
#include "operators.h"
#include "qtensor.h"
#include "useful.h"

#ifdef S4E
// 4e.cpp:10-25 (compiled from 4e.cpp with its main renamed)
qtensor<complex<double> > H_4e1(double t1, double t3, double f0, double f1, double p1, double frac);
qtensor<complex<double> > H_4e2(double t1, double t3, double f0, double f1, double p1, double frac);
qtensor<complex<double> > c_up1();
qtensor<complex<double> > c_down1();
qtensor<complex<double> > c_up2();
qtensor<complex<double> > c_down2();
qtensor<complex<double> > c_pair1();
qtensor<complex<double> > c_pair2();
qtensor<complex<double> > c_pairpair();
qtensor<complex<double> > c_pair_dagger1();
qtensor<complex<double> > c_pair_dagger2();
#endif

namespace {

template <typename T>
struct Elem;
template <>
struct Elem<double> {
  static double read() {
    double x;
    if (std::scanf("%lf", &x) != 1) throw std::string("bad value");
    return x;
  }
  static void write(double x) { std::printf(" %.17g", x); }
};
template <>
struct Elem<complex<double> > {
  static complex<double> read() {
    double a, b;
    if (std::scanf("%lf %lf", &a, &b) != 2) throw std::string("bad value");
    return complex<double>(a, b);
  }
  static void write(complex<double> x) { std::printf(" %.17g %.17g", x.real(), x.imag()); }
};

int read_int() {
  int x;
  if (std::scanf("%d", &x) != 1) throw std::string("bad integer");
  return x;
}

std::string read_word() {
  char buf[64];
  if (std::scanf("%63s", buf) != 1) throw std::string("missing keyword");
  return buf;
}

template <typename T>
qtensor<T> read_tensor() {
  const int nleg = read_int();
  vector<vector<int> > dim(nleg), sym;
  for (int i = 0; i < nleg; ++i) {
    const int ns = read_int();
    for (int k = 0; k < ns; ++k) dim[i].push_back(read_int());
  }
  const int nb = read_int();
  int total = 0;
  for (int b = 0; b < nb; ++b) {
    vector<int> s(nleg);
    int n = 1;
    for (int i = 0; i < nleg; ++i) {
      s[i] = read_int();
      n *= dim[i][s[i]];
    }
    sym.push_back(s);
    total += n;
  }
  int dir[8] = {2, 2, 2, 2, 2, 2, 2, 2};
  if (nleg > 8) throw std::string("at most 8 legs");
  for (int i = 0; i < nleg; ++i) dir[i] = read_int();
  vector<T> val(total > 0 ? total : 1);
  for (int i = 0; i < total; ++i) val[i] = Elem<T>::read();
  // the 3-argument constructor (src/qtensor.cpp:25-48), then the sign
  // directions through add_sign (src/qtensor.cpp:189-205)
  qtensor<T> t(val.data(), dim, sym);
  t.add_sign(dir[0], dir[1], dir[2], dir[3], dir[4], dir[5], dir[6], dir[7]);
  return t;
}

void print_dimsym(const char* name, const vector<vector<int> >& v) {
  std::printf("%s %d", name, (int)v.size());
  for (size_t i = 0; i < v.size(); ++i) {
    std::printf(" %d", (int)v[i].size());
    for (size_t k = 0; k < v[i].size(); ++k) std::printf(" %d", v[i][k]);
  }
  std::printf("\n");
}

// the sign directions of a tensor, read from the first line of
// qtensor::print (the only public view of them)
template <typename T>
vector<int> dirs_of(const qtensor<T>& t) {
  std::ostringstream os;
  std::streambuf* old = cout.rdbuf(os.rdbuf());
  t.print();
  cout.rdbuf(old);
  const std::string s = os.str();
  vector<int> d;
  const size_t a = s.find('('), b = s.find(')');
  if (a == std::string::npos || b == std::string::npos) return d;
  std::istringstream is(s.substr(a + 1, b - a - 1));
  std::string tok;
  while (std::getline(is, tok, ',')) d.push_back(std::atoi(tok.c_str()));
  return d;
}

// the tensor in the input format: dims, sectors, dirs, then the values
// through qtensor::val, which concatenates the blocks in block order
template <typename T>
void print_tensor(const char* name, const qtensor<T>& t) {
  const vector<int> dir = dirs_of(t);
  vector<vector<int> > dim, sym;
  const int n = t.dimension();
  vector<T> v(n > 0 ? n : 1);
  const int k = t.val(v.data(), dim, sym);
  std::printf("%s", name);
  print_dimsym(" dim", dim);
  print_dimsym("sym", sym);
  std::printf("dir %d", (int)dir.size());
  for (size_t i = 0; i < dir.size(); ++i) std::printf(" %d", dir[i]);
  std::printf("\nval %d", k);
  for (int i = 0; i < k; ++i) Elem<T>::write(v[i]);
  std::printf("\n");
}

void print_map(const char* name, const vector<vector<int> >& m) {
  std::printf("%s %d", name, (int)m.size());
  for (size_t i = 0; i < m.size(); ++i)
    for (size_t k = 0; k < m[i].size(); ++k) std::printf(" %d", m[i][k]);
  std::printf("\n");
}

template <typename T>
int run(const qtensor<T>& W) {
  if (read_word() != "lstore") throw std::string("expected lstore");
  qtensor<T> ls = read_tensor<T>();
  if (read_word() != "rstore") throw std::string("expected rstore");
  qtensor<T> rs = read_tensor<T>();
  if (read_word() != "sector") throw std::string("expected sector");
  const int sector = read_int();

  mpo<T> H(8, W);
  H.edge();
  print_tensor("W", W);
  print_tensor("Wleft", H[0]);
  print_tensor("Wright", H[7]);

  // update_two's environments, src/core.cpp:188-192
  qtensor<T> lenv, renv;
  lenv.contract(ls, 1, W, 0);
  lenv = lenv.exchange(3, 4);
  qtensor<T> ope = W.exchange(0, 2);
  renv.contract(rs, 1, ope, 0, true);
  renv = renv.exchange(0, 1);
  print_tensor("lenv", lenv);
  print_tensor("renv", renv);

  // src/core.cpp:236-240
  vector<vector<int> > dim, sym, dim_ret, sym_ret, rmap, lmap;
  generate_dim_sym(dim, sym, lenv, 2, renv, 2, sector);
  print_dimsym("dim", dim);
  print_dimsym("sym", sym);
  renv.get_map(rmap, dim_ret, sym_ret, dim, sym, 'N', 'T', 2, 0);
  print_map("rmap", rmap);
  print_dimsym("dim_ret", dim_ret);
  print_dimsym("sym_ret", sym_ret);
  vector<vector<int> > ldim = dim, lsym = sym;
  lenv.get_map(lmap, ldim, lsym, dim_ret, sym_ret, 'N', 'T', 3, 0, true);
  print_map("lmap", lmap);
  print_dimsym("ldim", ldim);
#ifdef PBC
  // src/core.cpp:193-199: both edges through the identity built from the left one
  qtensor<T> id = ls.id(W), ledge, redge;
  ledge.contract(ls, 1, id, 0);
  ledge = ledge.exchange(3, 4);
  redge.contract(rs, 1, id, 0, true);
  redge = redge.exchange(0, 1);
  print_tensor("ledge", ledge);
  print_tensor("redge", redge);
#endif
#ifdef S4E
  print_tensor("op_up1", c_up1());
  print_tensor("op_down1", c_down1());
  print_tensor("op_up2", c_up2());
  print_tensor("op_down2", c_down2());
  print_tensor("op_pair1", c_pair1());
  print_tensor("op_pair2", c_pair2());
  print_tensor("op_pairpair", c_pairpair());
  print_tensor("op_pair_dagger1", c_pair_dagger1());
  print_tensor("op_pair_dagger2", c_pair_dagger2());
  print_tensor("op_id", c_pair1().id());
#endif
  return 0;
}

}  // namespace

int main() {
  try {
    if (read_word() != "model") throw std::string("expected model");
    const std::string m = read_word();
#ifdef S4E
    double q[6];
    for (int i = 0; i < 6; ++i)
      if (std::scanf("%lf", &q[i]) != 1) throw std::string("bad parameter");
    if (m == "4e1") return run<complex<double> >(H_4e1(q[0], q[1], q[2], q[3], q[4], q[5]));
    if (m == "4e2") return run<complex<double> >(H_4e2(q[0], q[1], q[2], q[3], q[4], q[5]));
    throw std::string("the S4E build knows the models 4e1 and 4e2 only");
#endif
    double p[4];
    for (int i = 0; i < 4; ++i)
      if (std::scanf("%lf", &p[i]) != 1) throw std::string("bad parameter");
    if (m == "kitaev") return run<double>(H_Kitaev(p[0], p[1], p[2]));
    if (m == "xyz") return run<complex<double> >(H_XYZ(p[0], p[1], p[2]));
    if (m == "ising") return run<double>(H_Ising(p[0], p[1]));
    if (m == "potts2") return run<double>(H_Potts2(p[0], p[1]));
    throw std::string("unknown model ") + m;
  } catch (const std::string& e) {
    std::fprintf(stderr, "ref_harness: %s\n", e.c_str());
    return 2;
  }
}
