// Host-side mirror of the reference's tensor / qtensor / mps / mpo API
// (include/tensor.h, include/qtensor.h, include/mps.h, include/mpo.h,
// include/val_for_lanczos.h) with block data resident in device memory.
// Integer bookkeeping (sector maps, charge fusion, block order, get_map
// tuples) runs on the host and follows the reference exactly; every
// floating-point operation on block data runs on the GPU.
#pragma once

#include <complex>
#include <memory>
#include <string>
#include <vector>

#include "common.hpp"

namespace dmrgx {

using std::vector;

constexpr double TOL = 1e-10;  // global.h:19
constexpr int NEV = 6;         // global.h:21
constexpr int NSI = 20;        // global.h:22
constexpr double PI_REF = 3.14159265359;  // global.h:25

// Charge group: Z_N (the reference's compile-time SYMMETRY, src/useful.cpp:29-31)
// or, for SYM_S4E, the Z2 x Z2 x Z2 group of its S4E build (src/useful.cpp:4-27):
// eight sectors 0..7 (three bits), fusion a XOR |b|, fermion parity by table
constexpr int SYM_S4E = -8;
void set_symmetry(int n);
int symmetry();  // number of sectors
bool s4e();
int s4e_parity(int a);  // src/useful.cpp:22-26: {0,0,1,1,1,1,0,0}
inline int ss(int a, int b) {  // src/useful.cpp:6-20, 30
  return s4e() ? (a ^ (b < 0 ? -b : b)) : (a + b + symmetry()) % symmetry();
}
inline int ff(int a) { return s4e() ? s4e_parity(a) : a % 2; }  // src/useful.cpp:22-26, 32

void generate_map(vector<vector<int>>& map, const vector<int>& index);  // src/useful.cpp:119-134
int64_t get_dimension(const vector<vector<int>>& dim, const vector<vector<int>>& sym);  // :137-148

template <typename T>
struct dtype_of;
template <>
struct dtype_of<double> { static constexpr int value = DT_R64; };
template <>
struct dtype_of<cplx> { static constexpr int value = DT_C64; };

// Small host dense tensor used to build MPOs (tensor<T>, include/tensor.h).
template <typename T>
struct tensor {
  vector<int> index_;
  vector<T> val_;
  tensor(int i0 = 0, int i1 = 0, int i2 = 0, int i3 = 0, int i4 = 0, int i5 = 0, int i6 = 0, int i7 = 0);
  void update(T v, int i0, int i1 = -1, int i2 = -1, int i3 = -1, int i4 = -1, int i5 = -1, int i6 = -1,
              int i7 = -1);
  bool empty() const;
};

template <typename T>
class qtensor {
 public:
  vector<vector<int>> dim_;  // dim_[leg][sector]
  vector<vector<int>> sym_;  // sym_[block][leg]
  vector<int> dir_;          // fermion sign directions
  vector<int64_t> off_;      // element offset of each block
  int64_t total_ = 0;
  std::shared_ptr<DevBuf> buf_;

  qtensor(int i0 = 0, int i1 = 0, int i2 = 0, int i3 = 0, int i4 = 0, int i5 = 0, int i6 = 0, int i7 = 0);
  // qtensor(T* val, dim, sym) (src/qtensor.cpp:24-48); `val` is a device pointer
  qtensor(const T* dev_val, const vector<vector<int>>& dim, const vector<vector<int>>& sym);

  T* data() const { return buf_ ? (T*)buf_->ptr : nullptr; }
  T* block(int b) const { return data() + off_[b]; }
  int nblocks() const { return (int)sym_.size(); }
  vector<int> shape(int b) const;
  int64_t bsize(int b) const;
  void layout();        // recompute off_/total_ from dim_/sym_
  void allocate();      // fresh (uninitialised) buffer for the current layout
  void own();           // copy-on-write before an in-place change

  void update(const tensor<T>& val, int i0, int i1 = -1, int i2 = -1, int i3 = -1, int i4 = -1, int i5 = -1,
              int i6 = -1, int i7 = -1);
  void add_sign(int i0 = 2, int i1 = 2, int i2 = 2, int i3 = 2, int i4 = 2, int i5 = 2, int i6 = 2, int i7 = 2);
  void clear() { dim_.clear(); sym_.clear(); off_.clear(); total_ = 0; buf_.reset(); }

  vector<vector<int>> dimension_all() const { return dim_; }
  int dimension(int n) const;
  int64_t dimension() const;
  bool empty() const { return sym_.empty(); }
  // copy all block values to the host in block order (qtensor::val)
  vector<T> values() const;

  qtensor conjugate();             // flips dir_, conjugates in place; returns a copy
  qtensor<cplx> comp() const;
  qtensor simplify() const;
  qtensor exchange(int a = 0, int b = 1) const;
  qtensor take(int n, int k) const;
  qtensor left() const;
  qtensor right() const;
  // qtensor::id(bulk_mpo), src/qtensor.cpp:157-182: the identity operator
  // with the legs of the bulk MPO, its bond legs 0 and 2 given this tensor's
  // leg-1 sectors (carries a periodic edge term through a site)
  qtensor id(const qtensor& bulk_mpo) const;
  // qtensor::id(), src/qtensor.cpp:128-154: the identity on a 4-leg operator's
  // own legs (legs 0, 2 and 1, 3 paired)
  qtensor id() const;
  // qtensor::u_dim, src/qtensor.cpp:90-124: every sector of leg i gets dimension d_i
  void u_dim(int i0, int i1 = 0, int i2 = 0, int i3 = 0, int i4 = 0, int i5 = 0, int i6 = 0, int i7 = 0);
  qtensor combine(int min = 0, int max = 1) const;
  qtensor split(int index, const vector<vector<int>>& dim) const;
  qtensor remove_symmetry() const;

  qtensor& contract(const qtensor& A, const qtensor& B, char transa = 'N', char transb = 'N', int num = 1);
  qtensor& contract(const qtensor& A, int a, const qtensor& B, int b, bool change = false, bool fermion = true);
  // task-list GEMM (src/qtensor.cpp:986-994); out / in are device pointers
  void contract(T* out, const T* in, const vector<vector<int>>& map, char transa = 'N', char transb = 'T',
                int num = 1) const;
  // src/qtensor.cpp:997-1119; `fermion=false` forces alpha = 1 (bosonic sign rule)
  bool get_map(vector<vector<int>>& map_ret, vector<vector<int>>& dim_ret, vector<vector<int>>& sym_ret,
               const vector<vector<int>>& dim, const vector<vector<int>>& sym, char transa = 'N',
               char transb = 'N', int num = 1, double beta = 0, bool back = false, bool fermion = true) const;
  vector<double> svd(qtensor& U, qtensor<double>& S, qtensor& V, int num = 1, int cutoff = 0);
  int eig(double* val, vector<T>& vec, int sector = -1) const;
  T trace(const qtensor& A, bool fermion = true) const;
  qtensor& plus(const qtensor& A, const qtensor& B, T alpha = T(1), T beta = T(1));
};

// qtensor::svd of several tensors with one block SVD call over all their
// blocks; results as t[c]->svd(U[c], S[c], V[c], num, cutoff)
template <typename T>
vector<vector<double>> svd_batched(const vector<const qtensor<T>*>& t, vector<qtensor<T>>& U,
                                   vector<qtensor<double>>& S, vector<qtensor<T>>& V, int num, int cutoff);

template <typename T>
void generate_dim_sym(vector<vector<int>>& dim, vector<vector<int>>& sym, const qtensor<T>& lenv, int lnum,
                      const qtensor<T>& renv, int rnum, int sector);

// mpo<T>, include/mpo.h:9-61
template <typename T>
class mpo {
 public:
  int sites_;
  vector<qtensor<T>> operators_;
  qtensor<T> l_edge_, r_edge_;
  mpo(int sites, const qtensor<T>& ham) : sites_(sites) { operators_.push_back(ham); }
  explicit mpo(int sites) : sites_(sites) { operators_.resize(sites); }
  int size() const { return sites_; }
  void edge() {
    l_edge_ = operators_[0].left();
    r_edge_ = operators_[(sites_ - 1) % operators_.size()].right();
  }
  void edge(const qtensor<T>& l, const qtensor<T>& r) { l_edge_ = l; r_edge_ = r; }
  qtensor<T>& operator[](int n) {
    if (n == 0) return l_edge_;
    if (n == sites_ - 1) return r_edge_;
    return operators_[n % operators_.size()];
  }
};

// mps<T>, include/mps.h:9-75
template <typename T>
class mps {
 public:
  int cut_ = 0;
  vector<qtensor<T>> state_, store_;
  // the periodic edge environments (the reference's PBC build,
  // include/mps.h:18-50): edge_[0] / edge_[size-1] start as the two halves
  // of the boundary term and are carried through the chain like store_
  vector<qtensor<T>> edge_;
  bool pbc_ = false;
  qtensor<double> singular_;
  mps() = default;
  explicit mps(const mpo<T>& m) { state_.resize(m.size()); store_.resize(m.size()); }
  int size() const { return (int)state_.size(); }
  void clear(int i) {
    state_[i].clear();
    store_[i].clear();
    if (pbc_) edge_[i].clear();
  }
  // mps::add_edge (include/mps.h:47-50)
  void add_edge(const qtensor<T>& left, const qtensor<T>& right) {
    edge_.assign(state_.size(), qtensor<T>());
    edge_.front() = left;
    edge_.back() = right;
    pbc_ = true;
  }
  bool pbc() const { return pbc_; }
  qtensor<T>& edge(int n) { return edge_[n]; }
  qtensor<T>& operator[](int n) { return state_[n]; }
  qtensor<T>& operator()(int n) { return store_[n]; }
  qtensor<double>& center() { return singular_; }
  qtensor<double>& center(int n) { cut_ = n; return singular_; }
  int position() const { return cut_; }
  // src/mps.cpp:56-128: the chain reversed (site i -> size-1-i, bond legs
  // exchanged; no environments), and the centre moved by one bond through a
  // full (untruncated) SVD of the two sites beside it
  mps inverse() const;
  void move_cut_left();
  void move_cut_right();
};

// Model MPOs (src/operators.cpp)
qtensor<cplx> H_XYZ(double Jx, double Jy, double Jz);
qtensor<double> H_Kitaev(double t, double p, double u);
qtensor<cplx> H_Potts3(double f, double j, double p, double t);
qtensor<double> H_Potts2(double h, double j);
qtensor<double> H_Ising(double h, double j);

}  // namespace dmrgx
