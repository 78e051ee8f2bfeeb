// qtensor<T> with device-resident blocks.  Sector bookkeeping follows
// src/qtensor.cpp line by line (block order, first-appearance output order,
// get_map tuples, split's TOL pruning, the SVD selection tie-break); the
// block arithmetic goes to the HIP kernels of gemm.hip, tensor_ops.hip and
// svd.hip.
#include "qtensor.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>

#include "gemm.hpp"
#include "lanczos.hpp"
#include "svd.hpp"
#include "tensor_ops.hpp"

namespace dmrgx {

static int g_symmetry = 2;
static bool g_s4e = false;
void set_symmetry(int n) {
  DMRGX_CHECK(n >= 1 || n == SYM_S4E, "symmetry: a positive Z_N order or the Z2^3 group");
  g_s4e = n == SYM_S4E;
  g_symmetry = g_s4e ? 8 : n;
}
int symmetry() { return g_symmetry; }
bool s4e() { return g_s4e; }
int s4e_parity(int a) {
  static const int f[8] = {0, 0, 1, 1, 1, 1, 0, 0};
  DMRGX_CHECK(a >= 0 && a < 8, "Z2^3 sector out of range");
  return f[a];
}

void generate_map(vector<vector<int>>& map, const vector<int>& index) {
  int64_t dim = 1;
  for (int x : index) dim *= x;
  map.assign(dim, vector<int>(index.size(), 0));
  for (int64_t i = 0; i < dim; ++i) {
    int64_t r = i;
    for (int k = (int)index.size() - 1; k >= 0; --k) {
      map[i][k] = (int)(r % index[k]);
      r /= index[k];
    }
  }
}

int64_t get_dimension(const vector<vector<int>>& dim, const vector<vector<int>>& sym) {
  int64_t d = 0;
  for (auto& s : sym) {
    int64_t t = 1;
    for (size_t j = 0; j < dim.size(); ++j) t *= dim[j][s[j]];
    d += t;
  }
  return d;
}

static vector<int> pack_args(int i0, int i1, int i2, int i3, int i4, int i5, int i6, int i7, int stop_below) {
  int a[8] = {i0, i1, i2, i3, i4, i5, i6, i7};
  vector<int> r;
  for (int k = 0; k < 8; ++k) {
    if (a[k] < stop_below) break;
    r.push_back(a[k]);
  }
  return r;
}

// ---------------------------------------------------------------- tensor<T>
template <typename T>
tensor<T>::tensor(int i0, int i1, int i2, int i3, int i4, int i5, int i6, int i7) {
  index_ = pack_args(i0, i1, i2, i3, i4, i5, i6, i7, 1);
  int64_t n = index_.empty() ? 0 : 1;
  for (int x : index_) n *= x;
  val_.assign(n, T(0));
}

// src/tensor.cpp:68-102: adds v at the index; |v| < TOL is ignored
template <typename T>
void tensor<T>::update(T v, int i0, int i1, int i2, int i3, int i4, int i5, int i6, int i7) {
  if (std::abs(v) < TOL) return;
  vector<int> loc = pack_args(i0, i1, i2, i3, i4, i5, i6, i7, 0);
  DMRGX_CHECK(loc.size() == index_.size(), "tensor::update: wrong number of indices");
  int64_t e = 0;
  for (size_t i = 0; i < loc.size(); ++i) {
    DMRGX_CHECK(loc[i] >= 0 && loc[i] < index_[i], "tensor::update: index out of range");
    e = e * index_[i] + loc[i];
  }
  val_[e] += v;
}

// src/tensor.cpp:176-182: the sum of |values| is below TOL
template <typename T>
bool tensor<T>::empty() const {
  double sum = 0;
  for (auto& x : val_) sum += std::abs(x);
  return sum < TOL;
}

// ---------------------------------------------------------------- qtensor<T>
template <typename T>
qtensor<T>::qtensor(int i0, int i1, int i2, int i3, int i4, int i5, int i6, int i7) {
  vector<int> idx = pack_args(i0, i1, i2, i3, i4, i5, i6, i7, 1);
  dim_.resize(idx.size());
  for (size_t i = 0; i < idx.size(); ++i) dim_[i].assign(idx[i], 0);
  dir_.assign(idx.size(), 0);
}

template <typename T>
qtensor<T>::qtensor(const T* dev_val, const vector<vector<int>>& dim, const vector<vector<int>>& sym) {
  dim_ = dim;
  sym_ = sym;
  dir_.assign(dim.size(), 0);
  layout();
  allocate();
  if (total_) copy_d2d(data(), dev_val, total_ * sizeof(T), rt().stream);
}

template <typename T>
vector<int> qtensor<T>::shape(int b) const {
  vector<int> s(dim_.size());
  for (size_t i = 0; i < dim_.size(); ++i) s[i] = dim_[i][sym_[b][i]];
  return s;
}

template <typename T>
int64_t qtensor<T>::bsize(int b) const {
  int64_t n = 1;
  for (size_t i = 0; i < dim_.size(); ++i) n *= dim_[i][sym_[b][i]];
  return n;
}

template <typename T>
void qtensor<T>::layout() {
  off_.resize(sym_.size());
  total_ = 0;
  for (size_t b = 0; b < sym_.size(); ++b) {
    off_[b] = total_;
    total_ += bsize((int)b);
  }
}

template <typename T>
void qtensor<T>::allocate() {
  buf_ = std::make_shared<DevBuf>(std::max<int64_t>(total_, 1) * sizeof(T));
}

template <typename T>
void qtensor<T>::own() {
  if (buf_ && buf_.use_count() > 1) {
    auto nb = std::make_shared<DevBuf>(buf_->bytes);
    copy_d2d(nb->ptr, buf_->ptr, buf_->bytes, rt().stream);
    buf_ = nb;
  }
}

template <typename T>
int qtensor<T>::dimension(int n) const {
  int r = 0;
  for (int x : dim_[n]) r += x;
  return r;
}

template <typename T>
int64_t qtensor<T>::dimension() const {
  int64_t r = 1;
  for (size_t i = 0; i < dim_.size(); ++i) r *= dimension((int)i);
  return r;
}

template <typename T>
vector<T> qtensor<T>::values() const {
  vector<T> h(total_);
  if (total_) {
    DMRGX_HIP(hipMemcpyAsync(h.data(), data(), total_ * sizeof(T), hipMemcpyDeviceToHost, rt().stream));
    DMRGX_HIP(hipStreamSynchronize(rt().stream));
  }
  return h;
}

// src/qtensor.cpp:51-85 (check_dim_: include/qtensor.h:34-45)
template <typename T>
void qtensor<T>::update(const tensor<T>& val, int i0, int i1, int i2, int i3, int i4, int i5, int i6, int i7) {
  if (val.empty()) return;
  vector<int> loc = pack_args(i0, i1, i2, i3, i4, i5, i6, i7, 0);
  DMRGX_CHECK(loc.size() == dim_.size(), "qtensor::update: wrong number of indices");
  for (size_t i = 0; i < loc.size(); ++i) DMRGX_CHECK(loc[i] < (int)dim_[i].size(), "qtensor::update: sector out of range");
  for (size_t i = 0; i < loc.size(); ++i) {
    int& d = dim_[i][loc[i]];
    if (d == 0) d = val.index_[i];
    else DMRGX_CHECK(d == val.index_[i], "qtensor::update: dimension mismatch");
  }
  vector<T> old = values();
  sym_.push_back(loc);
  layout();
  old.insert(old.end(), val.val_.begin(), val.val_.end());
  allocate();
  upload(data(), old.data(), old.size() * sizeof(T), rt().stream);
}

template <typename T>
void qtensor<T>::add_sign(int i0, int i1, int i2, int i3, int i4, int i5, int i6, int i7) {
  int a[8] = {i0, i1, i2, i3, i4, i5, i6, i7};
  dir_.clear();
  for (int k = 0; k < 8 && a[k] != 2; ++k) dir_.push_back(a[k]);
  DMRGX_CHECK(dir_.size() == dim_.size(), "add_sign: wrong number of directions");
}

// src/qtensor.cpp:297-303
template <typename T>
qtensor<T> qtensor<T>::conjugate() {
  for (auto& d : dir_) d = -d;
  if (dtype_of<T>::value == DT_C64 && total_) {
    own();
    conj_inplace(data(), total_, rt().stream);
  }
  return *this;
}

template <typename T>
qtensor<cplx> qtensor<T>::comp() const {
  qtensor<cplx> r;
  r.dim_ = dim_;
  r.sym_ = sym_;
  r.dir_ = dir_;
  r.layout();
  r.allocate();
  if (dtype_of<T>::value == DT_C64) {
    if (total_) copy_d2d(r.data(), data(), total_ * sizeof(T), rt().stream);
  } else if (total_) {
    fill_zero(r.data(), total_ * sizeof(cplx), rt().stream);
    vector<BoxTask> t{BoxTask{0, 0, 0, 0, 0, 0, 1, 1, 1, 0}};
    // real -> complex: copy every real value into the real slot (stride 2)
    vector<T> h = values();
    vector<cplx> hc(h.begin(), h.end());
    upload(r.data(), hc.data(), hc.size() * sizeof(cplx), rt().stream);
  }
  return r;
}

// helper: new tensor with a subset of blocks copied over
template <typename T>
static qtensor<T> select_blocks(const qtensor<T>& a, const vector<int>& keep) {
  qtensor<T> r;
  r.dim_ = a.dim_;
  r.dir_ = a.dir_;
  for (int b : keep) r.sym_.push_back(a.sym_[b]);
  r.layout();
  r.allocate();
  vector<BoxTask> t;
  for (size_t q = 0; q < keep.size(); ++q) {
    int64_t n = a.bsize(keep[q]);
    if (n) t.push_back(BoxTask{a.off_[keep[q]], 0, 0, r.off_[q], 0, 0, 1, 1, (int)n, 0});
  }
  copy_boxes(dtype_of<T>::value, a.data(), r.data(), t, rt().stream);
  return r;
}

// src/qtensor.cpp:327-346
template <typename T>
qtensor<T> qtensor<T>::simplify() const {
  vector<BoxTask> t;
  for (int b = 0; b < nblocks(); ++b) t.push_back(BoxTask{off_[b], 0, 0, 0, 0, 0, 1, 1, (int)bsize(b), 0});
  vector<double> mx = box_maxabs(dtype_of<T>::value, data(), t, rt().stream);
  vector<int> keep;
  for (int b = 0; b < nblocks(); ++b)
    if (mx[b] > TOL) keep.push_back(b);
  return select_blocks(*this, keep);
}

// src/qtensor.cpp:349-373 over tensor::exchange (src/tensor.cpp:335-366)
template <typename T>
qtensor<T> qtensor<T>::exchange(int a, int b) const {
  if (dim_.empty()) return *this;
  DMRGX_CHECK(a >= 0 && b >= 0 && a < (int)dim_.size() && b < (int)dim_.size(), "exchange: leg out of range");
  qtensor<T> r = *this;
  std::swap(r.dim_[a], r.dim_[b]);
  std::swap(r.dir_[a], r.dir_[b]);
  for (auto& s : r.sym_) std::swap(s[a], s[b]);
  if (a == b) return r;
  r.layout();
  r.allocate();
  vector<PermTask> t;
  for (int q = 0; q < nblocks(); ++q) {
    if (!bsize(q)) continue;
    PermTask p{};
    p.in_off = off_[q];
    p.out_off = r.off_[q];
    p.nleg = (int)dim_.size();
    vector<int> sh = shape(q);
    for (int i = 0; i < p.nleg; ++i) { p.dims[i] = sh[i]; p.perm[i] = i; }
    std::swap(p.perm[a], p.perm[b]);
    t.push_back(p);
  }
  permute(dtype_of<T>::value, data(), r.data(), t, rt().stream);
  return r;
}

// src/qtensor.cpp:401-446 (+ tensor::take src/tensor.cpp:402-429)
template <typename T>
qtensor<T> qtensor<T>::take(int n, int k) const {
  DMRGX_CHECK(n >= 0 && n < (int)dim_.size() && dim_[n].size() == 1, "take: leg must exist and carry no symmetry");
  DMRGX_CHECK(k >= 0 && k < dim_[n][0], "take: index out of range");
  qtensor<T> r;
  for (size_t i = 0; i < dim_.size(); ++i)
    if ((int)i != n) { r.dim_.push_back(dim_[i]); r.dir_.push_back(dir_[i]); }
  for (auto& s : sym_) {
    vector<int> x;
    for (size_t i = 0; i < s.size(); ++i)
      if ((int)i != n) x.push_back(s[i]);
    r.sym_.push_back(x);
  }
  r.layout();
  r.allocate();
  vector<BoxTask> t;
  for (int b = 0; b < nblocks(); ++b) {
    vector<int> sh = shape(b);
    int64_t l = 1, rr = 1;
    for (int i = 0; i < n; ++i) l *= sh[i];
    for (size_t i = n + 1; i < sh.size(); ++i) rr *= sh[i];
    if (!l || !rr) continue;
    t.push_back(BoxTask{off_[b] + (int64_t)k * rr, (int64_t)sh[n] * rr, 0, r.off_[b], rr, 0, (int)l, 1, (int)rr, 0});
  }
  copy_boxes(dtype_of<T>::value, data(), r.data(), t, rt().stream);
  return r.simplify();
}

template <typename T>
qtensor<T> qtensor<T>::left() const { return take(0, dim_[0][0] - 1); }

template <typename T>
qtensor<T> qtensor<T>::right() const { return take(2, 0).exchange(0, 1); }

// identity blocks of a 4-leg operator with legs (0, 2) and (1, 3) paired:
// one block per (s0, s1) sector pair, sym (s0, s1, s0, s1)
template <typename T>
static qtensor<T> identity_blocks(qtensor<T>& r) {
  vector<T> h;
  for (size_t s0 = 0; s0 < r.dim_[0].size(); ++s0)
    for (size_t s1 = 0; s1 < r.dim_[1].size(); ++s1) {
      r.sym_.push_back({(int)s0, (int)s1, (int)s0, (int)s1});
      const int d0 = r.dim_[0][s0], d1 = r.dim_[1][s1];
      for (int i0 = 0; i0 < d0; ++i0)
        for (int i1 = 0; i1 < d1; ++i1)
          for (int j0 = 0; j0 < d0; ++j0)
            for (int j1 = 0; j1 < d1; ++j1) h.push_back(i0 == j0 && i1 == j1 ? T(1) : T(0));
    }
  r.layout();
  DMRGX_CHECK(r.total_ == (int64_t)h.size(), "id: block sizes");
  r.allocate();
  if (r.total_) upload(r.data(), h.data(), h.size() * sizeof(T), rt().stream);
  return r;
}

// src/qtensor.cpp:157-182: the bulk MPO's legs, the bond legs 0 and 2 with
// this tensor's leg-1 sectors
template <typename T>
qtensor<T> qtensor<T>::id(const qtensor<T>& bulk) const {
  DMRGX_CHECK(dim_.size() >= 2 && bulk.dim_.size() == 4 && bulk.dir_.size() == 4, "id: wrong number of legs");
  qtensor<T> r;
  r.dim_ = bulk.dim_;
  r.dim_[0] = dim_[1];
  r.dim_[2] = dim_[1];
  r.dir_ = bulk.dir_;
  r.dir_[0] = dir_.empty() ? 0 : dir_[1];
  r.dir_[2] = r.dir_[0];
  return identity_blocks(r);
}

// src/qtensor.cpp:128-154
template <typename T>
qtensor<T> qtensor<T>::id() const {
  DMRGX_CHECK(dim_.size() == 4 && dim_[0] == dim_[2] && dim_[1] == dim_[3], "id: wrong dim");
  qtensor<T> r;
  r.dim_ = dim_;
  r.dir_ = dir_;
  return identity_blocks(r);
}

template <typename T>
void qtensor<T>::u_dim(int i0, int i1, int i2, int i3, int i4, int i5, int i6, int i7) {
  vector<int> idx = pack_args(i0, i1, i2, i3, i4, i5, i6, i7, 1);
  DMRGX_CHECK(idx.size() == dim_.size(), "u_dim: the number of index doesn't match");
  for (size_t i = 0; i < idx.size(); ++i)
    for (int& d : dim_[i]) {
      DMRGX_CHECK(d == 0 || d == idx[i], "u_dim: a sector already has another dimension");
      d = idx[i];
    }
  layout();
}

// src/qtensor.cpp:466-585
template <typename T>
qtensor<T> qtensor<T>::combine(int mn, int mx) const {
  if (mn == mx) return *this;
  DMRGX_CHECK(mn >= 0 && mn < mx && mx < (int)dim_.size(), "combine: legs out of range");
  for (int i = mn; i < mx; ++i) DMRGX_CHECK(dir_[i] == dir_[mx], "combine: legs in different directions");
  qtensor<T> r;
  for (int i = 0; i < mn; ++i) r.dir_.push_back(dir_[i]);
  for (size_t i = mx; i < dir_.size(); ++i) r.dir_.push_back(dir_[i]);
  vector<vector<int>> cmap;
  vector<int> nsec;
  for (int s = mn; s <= mx; ++s) nsec.push_back((int)dim_[s].size());
  generate_map(cmap, nsec);
  vector<int64_t> sym_dim(cmap.size(), 1);
  vector<int> sym_sum(cmap.size(), 0);
  for (size_t i = 0; i < cmap.size(); ++i)
    for (int s = 0; s <= mx - mn; ++s) {
      sym_dim[i] *= dim_[mn + s][cmap[i][s]];
      sym_sum[i] = ss(sym_sum[i], cmap[i][s]);
    }
  int top = 0;
  for (int x : sym_sum) top = std::max(top, x);
  for (int i = 0; i < mn; ++i) r.dim_.push_back(dim_[i]);
  r.dim_.push_back(vector<int>(top + 1, 0));
  for (size_t i = mx + 1; i < dim_.size(); ++i) r.dim_.push_back(dim_[i]);
  for (size_t i = 0; i < sym_dim.size(); ++i) r.dim_[mn][sym_sum[i]] += (int)sym_dim[i];
  vector<int> index(r.dim_.size());
  for (size_t i = 0; i < index.size(); ++i) index[i] = (int)r.dim_[i].size();
  vector<vector<int>> rmap;
  generate_map(rmap, index);
  vector<char> check(rmap.size(), 0);
  for (auto& s : sym_) {
    int64_t k = 0;
    vector<int> loc;
    for (int i = 0; i < mn; ++i) loc.push_back(s[i]);
    int c = 0;
    for (int i = mn; i <= mx; ++i) c = ss(c, s[i]);
    loc.push_back(c);
    for (size_t i = mx + 1; i < dim_.size(); ++i) loc.push_back(s[i]);
    for (size_t i = 0; i < loc.size(); ++i) k = k * index[i] + loc[i];
    check[k] = 1;
  }
  for (size_t t = 0; t < check.size(); ++t)
    if (check[t]) r.sym_.push_back(rmap[t]);
  r.layout();
  r.allocate();
  if (r.total_) fill_zero(r.data(), r.total_ * sizeof(T), rt().stream);
  vector<int64_t> begs(sym_dim.size());
  vector<int64_t> store(index[mn], 0);
  for (size_t i = 0; i < sym_dim.size(); ++i) {
    begs[i] = store[sym_sum[i]];
    store[sym_sum[i]] += sym_dim[i];
  }
  vector<BoxTask> tasks;
  for (int t = 0; t < nblocks(); ++t) {
    const auto& s = sym_[t];
    int c = 0;
    for (int i = mn; i <= mx; ++i) c = ss(c, s[i]);
    int loc = -1;
    for (int q = 0; q < r.nblocks(); ++q) {
      if (c != r.sym_[q][mn]) continue;
      bool ok = true;
      for (int i = 0; i < mn; ++i) ok = ok && s[i] == r.sym_[q][i];
      for (size_t i = mx + 1; i < dim_.size(); ++i) ok = ok && s[i] == r.sym_[q][i - (mx - mn)];
      if (ok) loc = q;
    }
    DMRGX_CHECK(loc >= 0, "combine: cannot find sector");
    int64_t b = 0;
    for (int i = mn; i <= mx; ++i) b = b * dim_[i].size() + s[i];
    int64_t beg = begs[b];
    vector<int> sh = shape(t);
    int64_t l = 1, m = 1, rr = 1;
    for (int i = 0; i < mn; ++i) l *= sh[i];
    for (int i = mn; i <= mx; ++i) m *= sh[i];
    for (size_t i = mx + 1; i < sh.size(); ++i) rr *= sh[i];
    if (!l || !m || !rr) continue;
    int64_t Dc = r.dim_[mn][c];
    tasks.push_back(BoxTask{off_[t], m * rr, rr, r.off_[loc] + beg * rr, Dc * rr, rr, (int)l, (int)m, (int)rr, 0});
  }
  copy_boxes(dtype_of<T>::value, data(), r.data(), tasks, rt().stream);
  return r;
}

// src/qtensor.cpp:588-656
template <typename T>
qtensor<T> qtensor<T>::split(int n, const vector<vector<int>>& dim) const {
  qtensor<T> r;
  for (int i = 0; i < n; ++i) r.dir_.push_back(dir_[i]);
  for (size_t i = 0; i < dim.size(); ++i) r.dir_.push_back(dir_[n]);
  for (size_t i = n + 1; i < dir_.size(); ++i) r.dir_.push_back(dir_[i]);
  for (int i = 0; i < n; ++i) r.dim_.push_back(dim_[i]);
  for (auto& d : dim) r.dim_.push_back(d);
  for (size_t i = n + 1; i < dim_.size(); ++i) r.dim_.push_back(dim_[i]);
  vector<int> idx;
  for (auto& d : dim) idx.push_back((int)d.size());
  vector<vector<int>> smap;
  generate_map(smap, idx);
  vector<int64_t> begs(dim_[n].size(), 0);
  struct SD { int c; int64_t d, beg; };
  vector<SD> sd;
  for (auto& m : smap) {
    int c = 0;
    int64_t d = 1;
    for (size_t s = 0; s < dim.size(); ++s) { c = ss(c, m[s]); d *= dim[s][m[s]]; }
    sd.push_back(SD{c, d, begs[c]});
    begs[c] += d;
  }
  struct Cand { int t, q; BoxTask box; };
  vector<Cand> cand;
  vector<BoxTask> check;
  for (int t = 0; t < nblocks(); ++t)
    for (size_t q = 0; q < sd.size(); ++q) {
      if (sd[q].c != sym_[t][n]) continue;
      vector<int> sh = shape(t);
      int64_t l = 1, rr = 1, m = sh[n];
      for (int i = 0; i < n; ++i) l *= sh[i];
      for (size_t i = n + 1; i < sh.size(); ++i) rr *= sh[i];
      BoxTask b{off_[t] + sd[q].beg * rr, m * rr, rr, 0, sd[q].d * rr, rr, (int)l, (int)sd[q].d, (int)rr, 0};
      cand.push_back(Cand{t, (int)q, b});
      check.push_back(b);
    }
  // blocks whose values are all <= TOL are dropped (src/qtensor.cpp:640-653)
  vector<BoxTask> nonempty;
  vector<int> map_ne;
  for (size_t i = 0; i < check.size(); ++i)
    if ((int64_t)check[i].n0 * check[i].n1 * check[i].n2 > 0) { nonempty.push_back(check[i]); map_ne.push_back((int)i); }
  vector<double> mx(check.size(), 0.0);
  {
    vector<double> m2 = box_maxabs(dtype_of<T>::value, data(), nonempty, rt().stream);
    for (size_t i = 0; i < map_ne.size(); ++i) mx[map_ne[i]] = m2[i];
  }
  vector<BoxTask> copies;
  for (size_t i = 0; i < cand.size(); ++i) {
    if (!(mx[i] > TOL)) continue;
    const auto& s = sym_[cand[i].t];
    vector<int> ns;
    for (int k = 0; k < n; ++k) ns.push_back(s[k]);
    for (int x : smap[cand[i].q]) ns.push_back(x);
    for (size_t k = n + 1; k < s.size(); ++k) ns.push_back(s[k]);
    r.sym_.push_back(ns);
    copies.push_back(cand[i].box);
  }
  r.layout();
  r.allocate();
  for (size_t i = 0; i < copies.size(); ++i) copies[i].dst_off = r.off_[i];
  copy_boxes(dtype_of<T>::value, data(), r.data(), copies, rt().stream);
  return r;
}

// src/qtensor.cpp:659-689 (host-side; only the tiny exact-diagonalisation
// branch and the checks use it)
template <typename T>
qtensor<T> qtensor<T>::remove_symmetry() const {
  qtensor<T> r;
  r.dir_ = dir_;
  for (size_t i = 0; i < dim_.size(); ++i) r.dim_.push_back(vector<int>(1, dimension((int)i)));
  r.sym_.push_back(vector<int>(dim_.size(), 0));
  r.layout();
  vector<T> h = values(), out(r.total_, T(0));
  vector<int64_t> full(dim_.size());
  for (size_t i = 0; i < dim_.size(); ++i) full[i] = dimension((int)i);
  for (int b = 0; b < nblocks(); ++b) {
    vector<int> sh = shape(b);
    vector<int64_t> base(dim_.size(), 0);
    for (size_t i = 0; i < dim_.size(); ++i)
      for (int k = 0; k < sym_[b][i]; ++k) base[i] += dim_[i][k];
    int64_t n = bsize(b);
    for (int64_t e = 0; e < n; ++e) {
      int64_t rem = e, loc = 0;
      vector<int64_t> c(dim_.size());
      for (int i = (int)dim_.size() - 1; i >= 0; --i) { c[i] = rem % sh[i]; rem /= sh[i]; }
      for (size_t i = 0; i < dim_.size(); ++i) loc = loc * full[i] + base[i] + c[i];
      out[loc] = h[off_[b] + e];
    }
  }
  r.allocate();
  upload(r.data(), out.data(), out.size() * sizeof(T), rt().stream);
  return r;
}

// src/qtensor.cpp:824-899 over tensor::contract(A,B,...) src/tensor.cpp:603-649
template <typename T>
qtensor<T>& qtensor<T>::contract(const qtensor<T>& A, const qtensor<T>& B, char ta, char tb, int num) {
  const bool an = ta == 'N' || ta == 'n', bn = tb == 'N' || tb == 'n';
  int ia = an ? (int)A.dim_.size() - num : 0;
  int ib = bn ? 0 : (int)B.dim_.size() - num;
  for (int i = 0; i < num; ++i) DMRGX_CHECK(A.dim_[ia + i] == B.dim_[ib + i], "contract: dimensions differ");
  vector<int> ka, kb;
  for (int i = 0; i < (int)A.dim_.size(); ++i)
    if (i < ia || i >= ia + num) ka.push_back(i);
  for (int i = 0; i < (int)B.dim_.size(); ++i)
    if (i < ib || i >= ib + num) kb.push_back(i);
  vector<vector<int>> dim;
  vector<int> dir;
  for (int i : ka) { dim.push_back(A.dim_[i]); dir.push_back(A.dir_[i]); }
  for (int i : kb) { dim.push_back(B.dim_[i]); dir.push_back(B.dir_[i]); }
  vector<vector<int>> sym;
  vector<int64_t> locs;
  struct P { int i, j, o; int64_t M, N, K; };
  vector<P> pairs;
  for (int i = 0; i < A.nblocks(); ++i)
    for (int j = 0; j < B.nblocks(); ++j) {
      bool ok = true;
      for (int k = 0; k < num; ++k) ok = ok && A.sym_[i][ia + k] == B.sym_[j][ib + k];
      if (!ok) continue;
      vector<int> s;
      for (int k : ka) s.push_back(A.sym_[i][k]);
      for (int k : kb) s.push_back(B.sym_[j][k]);
      int64_t loc = 0;
      for (size_t q = 0; q < dim.size(); ++q) loc = loc * dim[q].size() + s[q];
      int o = -1;
      for (size_t t = 0; t < locs.size(); ++t)
        if (locs[t] == loc) o = (int)t;
      if (o < 0) {
        o = (int)locs.size();
        locs.push_back(loc);
        sym.push_back(s);
      }
      int64_t K = 1;
      vector<int> sa = A.shape(i);
      for (int k = 0; k < num; ++k) K *= sa[ia + k];
      pairs.push_back(P{i, j, o, A.bsize(i) / std::max<int64_t>(K, 1), B.bsize(j) / std::max<int64_t>(K, 1), K});
    }
  qtensor<T> r;
  r.dim_ = dim;
  r.dir_ = dir;
  r.sym_ = sym;
  // check_dim_ on the result
  for (auto& p : pairs) {
    (void)p;
  }
  r.layout();
  r.allocate();
  if (r.total_) fill_zero(r.data(), r.total_ * sizeof(T), rt().stream);
  vector<GemmTask> tasks;
  vector<char> seen(sym.size(), 0);
  for (auto& p : pairs) {
    GemmTask g{};
    g.a_off = A.off_[p.i];
    g.b_off = B.off_[p.j];
    g.c_off = r.off_[p.o];
    g.M = (int)p.M;
    g.N = (int)p.N;
    g.K = (int)p.K;
    g.alpha = 1;
    g.beta = seen[p.o] ? 1 : 0;
    seen[p.o] = 1;
    DMRGX_CHECK(g.M * (int64_t)g.N == r.bsize(p.o), "contract: output block size mismatch");
    tasks.push_back(g);
  }
  if (!tasks.empty()) {
    auto plan = make_gemm_plan(tasks, dtype_of<T>::value, PREC_FP64, an ? 'N' : 'T', bn ? 'N' : 'T', {});
    run_gemm_plan(*plan, A.data(), B.data(), r.data(), rt().stream);
  }
  *this = r;
  return *this;
}

// src/qtensor.cpp:902-983 over tensor::contract(A,a,B,b,cover,sign) src/tensor.cpp:652-695
template <typename T>
qtensor<T>& qtensor<T>::contract(const qtensor<T>& A, int a, const qtensor<T>& B, int b, bool change, bool fermion) {
  DMRGX_CHECK(A.dim_[a] == B.dim_[b], "contract: dimension mismatch");
  const int nb = (int)B.dim_.size();
  vector<vector<int>> dim;
  vector<int> dir;
  auto order = [&](const vector<int>& sa, const vector<int>& sb, vector<int>& out) {
    out.clear();
    for (int i = 0; i < a; ++i) out.push_back(sa[i]);
    for (int i = 0; i < b; ++i) out.push_back(sb[i]);
    for (int i = b + 1; i < nb; ++i) out.push_back(sb[i]);
    for (int i = a + 1; i < (int)sa.size(); ++i) out.push_back(sa[i]);
  };
  for (int i = 0; i < a; ++i) { dim.push_back(A.dim_[i]); dir.push_back(A.dir_[i]); }
  for (int i = 0; i < b; ++i) { dim.push_back(B.dim_[i]); dir.push_back(B.dir_[i]); }
  for (int i = b + 1; i < nb; ++i) { dim.push_back(B.dim_[i]); dir.push_back(B.dir_[i]); }
  for (int i = a + 1; i < (int)A.dim_.size(); ++i) { dim.push_back(A.dim_[i]); dir.push_back(A.dir_[i]); }
  vector<vector<int>> sym;
  vector<int64_t> locs;
  struct P { int l, r, o, sign; };
  vector<P> pairs;
  vector<int> s;
  for (int l = 0; l < A.nblocks(); ++l)
    for (int q = 0; q < B.nblocks(); ++q) {
      if (A.sym_[l][a] != B.sym_[q][b]) continue;
      order(A.sym_[l], B.sym_[q], s);
      int sign = 1;
      if (fermion) {
        if (change) {
          int d_a = 0, b_p = 0;
          for (size_t i = 0; i < A.dir_.size(); ++i) d_a = ss(d_a, A.sym_[l][i] * A.dir_[i]);
          for (size_t i = 0; i < B.dir_.size(); ++i)
            if (B.dir_[i] == 1) b_p = ss(b_p, B.sym_[q][i]);
          sign = 1 - 2 * ((ff(d_a) * ff(b_p)) % 2);
        } else {
          int d_b = 0, a_p = 0;
          for (size_t i = 0; i < B.dir_.size(); ++i) d_b = ss(d_b, B.sym_[q][i] * B.dir_[i]);
          for (size_t i = 0; i < A.dir_.size(); ++i)
            if (A.dir_[i] == 1) a_p = ss(a_p, A.sym_[l][i]);
          sign = 1 - 2 * ((ff(d_b) * ff(a_p)) % 2);
        }
      }
      int64_t loc = 0;
      for (size_t k = 0; k < s.size(); ++k) loc = loc * dim[k].size() + s[k];
      int o = -1;
      for (size_t t = 0; t < locs.size(); ++t)
        if (locs[t] == loc) o = (int)t;
      if (o < 0) {
        o = (int)locs.size();
        locs.push_back(loc);
        sym.push_back(s);
      }
      pairs.push_back(P{l, q, o, sign});
    }
  qtensor<T> r;
  r.dim_ = dim;
  r.dir_ = dir;
  r.sym_ = sym;
  r.layout();
  r.allocate();
  // group pairs by output block (first-appearance order of outputs)
  vector<LegOut> outs;
  vector<LegPair> lp;
  for (int o = 0; o < (int)sym.size(); ++o) {
    LegOut lo{};
    lo.c_off = r.off_[o];
    lo.pair_begin = (int)lp.size();
    bool first = true;
    for (auto& p : pairs) {
      if (p.o != o) continue;
      vector<int> sa = A.shape(p.l), sb = B.shape(p.r);
      int64_t al = 1, ar = 1, bl = 1, br = 1;
      for (int i = 0; i < a; ++i) al *= sa[i];
      for (size_t i = a + 1; i < sa.size(); ++i) ar *= sa[i];
      for (int i = 0; i < b; ++i) bl *= sb[i];
      for (size_t i = b + 1; i < sb.size(); ++i) br *= sb[i];
      if (first) {
        lo.al = (int)al; lo.ar = (int)ar; lo.bl = (int)bl; lo.br = (int)br;
        first = false;
      }
      lp.push_back(LegPair{A.off_[p.l], B.off_[p.r], sa[a], p.sign});
    }
    lo.pair_end = (int)lp.size();
    lo.accumulate = 0;
    if ((int64_t)lo.al * lo.bl * lo.br * lo.ar > 0) outs.push_back(lo);
  }
  leg_contract(dtype_of<T>::value, A.data(), B.data(), r.data(), outs, lp, rt().stream);
  *this = r;
  return *this;
}

template <typename T>
void qtensor<T>::contract(T* out, const T* in, const vector<vector<int>>& map, char ta, char tb, int num) const {
  DMRGX_CHECK((ta == 'N' || ta == 'n') && (tb == 'T' || tb == 't'), "task contract supports ('N','T') only");
  (void)num;
  vector<GemmTask> t;
  for (auto& m : map) {
    GemmTask g{};
    g.a_off = off_[m[0]];
    g.b_off = m[1];
    g.c_off = m[2];
    g.K = m[4];
    g.N = m[3];
    g.M = (int)(bsize(m[0]) / m[4]);
    g.beta = m[5];
    g.alpha = m[6];
    t.push_back(g);
  }
  auto plan = make_gemm_plan(t, dtype_of<T>::value, PREC_FP64, 'N', 'T', {});
  run_gemm_plan(*plan, data(), in, out, rt().stream);
}

// src/qtensor.cpp:997-1119
template <typename T>
bool qtensor<T>::get_map(vector<vector<int>>& map_ret, vector<vector<int>>& dim_ret, vector<vector<int>>& sym_ret,
                         const vector<vector<int>>& dim, const vector<vector<int>>& sym, char ta, char tb, int num,
                         double beta, bool back, bool fermion) const {
  int ia = (ta == 'N' || ta == 'n') ? (int)dim_.size() - num : 0;
  int ib = (tb == 'N' || tb == 'n') ? 0 : (int)dim.size() - num;
  for (int i = 0; i < num; ++i)
    if (dim_[ia + i] != dim[ib + i]) {
      set_last_error("get_map: dimensions differ");
      return false;
    }
  dim_ret.clear();
  for (int i = 0; i < (int)dim_.size(); ++i)
    if (i < ia || i >= ia + num) dim_ret.push_back(dim_[i]);
  for (int i = 0; i < (int)dim.size(); ++i)
    if (i < ib || i >= ib + num) dim_ret.push_back(dim[i]);
  vector<std::array<int64_t, 3>> sym_dim(sym.size());
  for (size_t i = 0; i < sym.size(); ++i) {
    int64_t tot = 1, row = 1;
    for (size_t s = 0; s < dim.size(); ++s) tot *= dim[s][sym[i][s]];
    if (ib)
      for (int s = 0; s < ib; ++s) row *= dim[s][sym[i][s]];
    else
      for (int s = 0; s < num; ++s) row *= dim[s][sym[i][s]];
    sym_dim[i] = {tot, row, 0};
  }
  for (size_t i = 1; i < sym.size(); ++i) sym_dim[i][2] = sym_dim[i - 1][2] + sym_dim[i - 1][0];
  map_ret.clear();
  if (!back) sym_ret.clear();
  vector<int64_t> locs, map_dim(1, 0);
  if (back) {
    locs.assign(sym_ret.size(), 0);
    for (size_t i = 0; i < sym_ret.size(); ++i)
      for (size_t j = 0; j < dim_ret.size(); ++j) locs[i] = locs[i] * dim_ret[j].size() + sym_ret[i][j];
    map_dim.assign(sym_ret.size(), 0);
    for (size_t i = 0; i + 1 < map_dim.size(); ++i) {
      int64_t d = 1;
      for (size_t j = 0; j < dim_ret.size(); ++j) d *= dim_ret[j][sym_ret[i][j]];
      map_dim[i + 1] = map_dim[i] + d;
    }
  }
  vector<char> first(locs.size(), 1);
  vector<int> st(dim_ret.size());
  for (int i = 0; i < nblocks(); ++i)
    for (size_t j = 0; j < sym.size(); ++j) {
      bool ok = true;
      for (int k = 0; k < num; ++k) ok = ok && sym_[i][ia + k] == sym[j][ib + k];
      if (!ok || sym_dim[j][0] == 0) continue;
      DMRGX_CHECK(sym_dim[j][2] <= INT32_MAX && map_dim.back() <= INT32_MAX, "get_map: offsets exceed int32");
      vector<int> ms = {i, (int)sym_dim[j][2], back && map_dim.empty() ? 0 : (int)map_dim.back(),
                        (int)sym_dim[j][1], (int)(sym_dim[j][0] / sym_dim[j][1]), (int)beta, 1};
      if (!back) {
        int d_b = 0, a_p = 0;
        for (size_t k = 0; k < dir_.size(); ++k) d_b = ss(d_b, sym_[i][k] * dir_[k]);
        for (size_t k = 0; k + num < dim.size(); ++k) a_p = ss(a_p, sym[j][k]);
        ms[6] = fermion ? 1 - 2 * ((ff(d_b) * ff(a_p)) % 2) : 1;
      }
      int ad = 0;
      for (int k = 0; k < (int)dim_.size(); ++k)
        if (k < ia || k >= ia + num) st[ad++] = sym_[i][k];
      for (int k = 0; k < (int)dim.size(); ++k)
        if (k < ib || k >= ib + num) st[ad++] = sym[j][k];
      int64_t loc = 0;
      for (size_t s = 0; s < dim_ret.size(); ++s) loc = loc * dim_ret[s].size() + st[s];
      bool same = false;
      for (size_t t = 0; t < locs.size(); ++t)
        if (loc == locs[t]) {
          ms[2] = (int)map_dim[t];
          if (back && first[t]) first[t] = 0;
          else ms[5] = 1;
          map_ret.push_back(ms);
          same = true;
        }
      if (!same && !back) {
        int64_t d = 1;
        for (size_t s = 0; s < dim_ret.size(); ++s) d *= dim_ret[s][st[s]];
        map_dim.push_back(map_dim.back() + d);
        locs.push_back(loc);
        first.push_back(1);
        sym_ret.push_back(st);
        map_ret.push_back(ms);
      }
    }
  return true;
}

// qtensor::svd, src/qtensor.cpp:1122-1262, in two halves around the block
// SVD so that several tensors can share one block_svd call (svd_batched).
namespace {

template <typename T>
struct SvdJob {
  qtensor<T> comb;
  vector<SvdBlock> blocks;  // offsets relative to comb and to this job's U, S, V
  vector<int> rows, cols, need_u, need_v;
  int64_t uo = 0, vo = 0, so = 0;
};

template <typename T>
SvdJob<T> svd_prepare(const qtensor<T>& t, int num, int cutoff) {
  DMRGX_CHECK(num > 0 && num < (int)t.dim_.size(), "svd: number of left legs out of range");
  SvdJob<T> j;
  j.comb = t.combine(num, (int)t.dim_.size() - 1).combine(0, num - 1);
  const qtensor<T>& comb = j.comb;
  DMRGX_CHECK(comb.nblocks() > 0, "svd: empty tensor");
  int sym0 = ss(comb.sym_[0][0], comb.sym_[0][1]);
  for (auto& s : comb.sym_) DMRGX_CHECK(ss(s[0], s[1]) == sym0, "svd: symmetry sectors do not have constraint");
  const int nsym = comb.nblocks();
  j.blocks.resize(nsym);
  j.rows.resize(nsym);
  j.cols.resize(nsym);
  for (int i = 0; i < nsym; ++i) {
    vector<int> sh = comb.shape(i);
    j.rows[i] = sh[0];
    j.cols[i] = sh[1];
    j.blocks[i] = SvdBlock{sh[0], sh[1], comb.off_[i], j.uo, j.so, j.vo};
    j.uo += (int64_t)sh[0] * sh[0];
    j.vo += (int64_t)sh[1] * sh[1];
    j.so += std::min(sh[0], sh[1]);
  }
  j.need_u.resize(nsym);
  j.need_v.resize(nsym);
  for (int i = 0; i < nsym; ++i) {
    j.need_u[i] = cutoff ? std::min(j.rows[i], cutoff) : j.rows[i];
    j.need_v[i] = cutoff ? std::min(j.cols[i], cutoff) : j.cols[i];
  }
  return j;
}

// truncation and U, S, V from the blocks' factors: Ub / Vb (device) and sall
// (host) hold this job's full U, V and singular values
template <typename T>
vector<double> svd_finish(const vector<vector<int>>& dim, const SvdJob<T>& j, const void* Ub, const void* Vb,
                          const double* sall, qtensor<T>& U, qtensor<double>& S, qtensor<T>& V, int num,
                          int cutoff) {
  const qtensor<T>& comb = j.comb;
  const vector<SvdBlock>& blocks = j.blocks;
  const vector<int>& rows = j.rows;
  const vector<int>& cols = j.cols;
  const int nsym = comb.nblocks();
  vector<vector<double>> Sv(nsym);
  int total = 0;
  for (int i = 0; i < nsym; ++i) {
    int mn = std::min(rows[i], cols[i]);
    Sv[i].assign(sall + blocks[i].s_off, sall + blocks[i].s_off + mn);
    total += mn;
  }
  // global selection by repeated max scan (ties -> lowest sector, strict <)
  vector<double> s(total);
  int keep = total;
  if (cutoff != 0 && cutoff < total) keep = cutoff;
  vector<int> u_num(nsym, 0), v_num;
  for (int t = 0; t < total; ++t) {
    double mx = -1;
    int loc = -1;
    for (int i = 0; i < nsym; ++i)
      if (u_num[i] < (int)Sv[i].size() && mx < Sv[i][u_num[i]]) { mx = Sv[i][u_num[i]]; loc = i; }
    s[t] = mx;
    u_num[loc] += 1;
    if (t == keep - 1) v_num = u_num;
  }
  if (keep == 0) v_num.assign(nsym, 0);
  u_num = v_num;
  // U, V, S metadata (U.dim_[1] = U.dim_[0], V.dim_[0] = V.dim_[1])
  U = qtensor<T>();
  V = qtensor<T>();
  S = qtensor<double>();
  U.dir_ = {1, 1};
  V.dir_ = {1, 1};
  S.dir_ = {1, 1};
  U.dim_ = comb.dim_;
  U.dim_[1] = U.dim_[0];
  V.dim_ = comb.dim_;
  V.dim_[0] = V.dim_[1];
  S.dim_ = comb.dim_;
  vector<int> ucols(nsym), vrows(nsym), srows(nsym), scols(nsym);
  for (int i = 0; i < nsym; ++i) {
    vector<int> su = comb.sym_[i], sv = comb.sym_[i];
    su[1] = su[0];
    sv[0] = sv[1];
    U.sym_.push_back(su);
    V.sym_.push_back(sv);
    S.sym_.push_back(comb.sym_[i]);
    ucols[i] = rows[i];
    vrows[i] = cols[i];
    srows[i] = rows[i];
    scols[i] = cols[i];
  }
  if (cutoff != 0) {
    // zero-weight padding up to min(dimension, cutoff) columns of U / rows of
    // V.  Each block's bound is the leg dimension of the sector numbered like
    // the block, comb.dim_[.][i] (src/qtensor.cpp:1225,1233), not the block's
    // own sector; the box copies below zero-pad a U wider than its rows.
    DMRGX_CHECK(nsym <= (int)comb.dim_[0].size() && nsym <= (int)comb.dim_[1].size(),
                "svd: more blocks than sectors in the padding");
    int diff = (int)std::min<int64_t>(comb.dimension(0), cutoff) - keep;
    while (diff > 0)
      for (int i = 0; i < nsym; ++i)
        if (u_num[i] < comb.dim_[0][i] && diff > 0) { u_num[i] += 1; --diff; }
    diff = (int)std::min<int64_t>(comb.dimension(1), cutoff) - keep;
    while (diff > 0)
      for (int i = 0; i < nsym; ++i)
        if (v_num[i] < comb.dim_[1][i] && diff > 0) { v_num[i] += 1; --diff; }
    for (int i = 0; i < nsym; ++i) {
      ucols[i] = u_num[i];
      U.dim_[1][U.sym_[i][1]] = u_num[i];
      vrows[i] = v_num[i];
      V.dim_[0][V.sym_[i][0]] = v_num[i];
      srows[i] = u_num[i];
      S.dim_[0][S.sym_[i][0]] = u_num[i];
      scols[i] = v_num[i];
      S.dim_[1][S.sym_[i][1]] = v_num[i];
    }
  }
  U.layout();
  V.layout();
  S.layout();
  U.allocate();
  V.allocate();
  S.allocate();
  if (U.total_) fill_zero(U.data(), U.total_ * sizeof(T), rt().stream);
  if (V.total_) fill_zero(V.data(), V.total_ * sizeof(T), rt().stream);
  vector<BoxTask> tu, tv;
  for (int i = 0; i < nsym; ++i) {
    // first ucols[i] columns of the full rows x rows U (zero-padded if wider)
    int uc = std::min(ucols[i], rows[i]);
    if (rows[i] && uc)
      tu.push_back(BoxTask{blocks[i].u_off, rows[i], 0, U.off_[i], ucols[i], 0, rows[i], 1, uc, 0});
    int vr = std::min(vrows[i], cols[i]);
    if (vr && cols[i]) tv.push_back(BoxTask{blocks[i].v_off, cols[i], 0, V.off_[i], cols[i], 0, vr, 1, cols[i], 0});
  }
  copy_boxes(dtype_of<T>::value, Ub, U.data(), tu, rt().stream);
  copy_boxes(dtype_of<T>::value, Vb, V.data(), tv, rt().stream);
  // S blocks: diagonal singular values, resized like tensor::diag(2) + resize
  vector<double> hs(S.total_, 0.0);
  for (int i = 0; i < nsym; ++i) {
    int n = std::min({(int)Sv[i].size(), srows[i], scols[i]});
    for (int k = 0; k < n; ++k) hs[S.off_[i] + (int64_t)k * scols[i] + k] = Sv[i][k];
  }
  if (S.total_) {
    upload(S.data(), hs.data(), hs.size() * 8, rt().stream);
  }
  vector<vector<int>> dl(dim.begin(), dim.begin() + num), dr(dim.begin() + num, dim.end());
  U = U.split(0, dl);
  V = V.split(1, dr);
  return s;
}

}  // namespace

template <typename T>
vector<double> qtensor<T>::svd(qtensor<T>& U, qtensor<double>& S, qtensor<T>& V, int num, int cutoff) {
  SvdJob<T> j = svd_prepare(*this, num, cutoff);
  DevBuf Ub(std::max<int64_t>(j.uo, 1) * sizeof(T)), Vb(std::max<int64_t>(j.vo, 1) * sizeof(T)),
      Sb(std::max<int64_t>(j.so, 1) * 8);
  block_svd(dtype_of<T>::value, j.comb.data(), Ub.ptr, (double*)Sb.ptr, Vb.ptr, j.blocks, j.need_u, j.need_v,
            rt().stream);
  vector<double> sall(j.so);
  if (j.so) {
    DMRGX_HIP(hipMemcpyAsync(sall.data(), Sb.ptr, j.so * 8, hipMemcpyDeviceToHost, rt().stream));
    DMRGX_HIP(hipStreamSynchronize(rt().stream));
  }
  return svd_finish(dim_, j, Ub.ptr, Vb.ptr, sall.data(), U, S, V, num, cutoff);
}

template <typename T>
vector<vector<double>> svd_batched(const vector<const qtensor<T>*>& t, vector<qtensor<T>>& U,
                                   vector<qtensor<double>>& S, vector<qtensor<T>>& V, int num, int cutoff) {
  const int n = (int)t.size();
  hipStream_t st = rt().stream;
  vector<SvdJob<T>> jobs;
  jobs.reserve(n);
  vector<int64_t> ab(n), ub(n), vb(n), sb(n);
  int64_t at = 0, ut = 0, vt = 0, stt = 0;
  for (int c = 0; c < n; ++c) {
    jobs.push_back(svd_prepare(*t[c], num, cutoff));
    ab[c] = at;
    ub[c] = ut;
    vb[c] = vt;
    sb[c] = stt;
    at += jobs[c].comb.total_;
    ut += jobs[c].uo;
    vt += jobs[c].vo;
    stt += jobs[c].so;
  }
  // every tensor's blocks in one arena and one block_svd call
  DevBuf Ab(std::max<int64_t>(at, 1) * sizeof(T)), Ub(std::max<int64_t>(ut, 1) * sizeof(T)),
      Vb(std::max<int64_t>(vt, 1) * sizeof(T)), Sb(std::max<int64_t>(stt, 1) * 8);
  vector<SvdBlock> blocks;
  vector<int> need_u, need_v;
  for (int c = 0; c < n; ++c) {
    const SvdJob<T>& j = jobs[c];
    if (j.comb.total_) copy_d2d((T*)Ab.ptr + ab[c], j.comb.data(), j.comb.total_ * sizeof(T), st);
    for (auto b : j.blocks) {
      b.a_off += ab[c];
      b.u_off += ub[c];
      b.v_off += vb[c];
      b.s_off += sb[c];
      blocks.push_back(b);
    }
    need_u.insert(need_u.end(), j.need_u.begin(), j.need_u.end());
    need_v.insert(need_v.end(), j.need_v.begin(), j.need_v.end());
  }
  block_svd(dtype_of<T>::value, Ab.ptr, Ub.ptr, (double*)Sb.ptr, Vb.ptr, blocks, need_u, need_v, st);
  vector<double> sall(stt);
  if (stt) {
    DMRGX_HIP(hipMemcpyAsync(sall.data(), Sb.ptr, stt * 8, hipMemcpyDeviceToHost, st));
    DMRGX_HIP(hipStreamSynchronize(st));
  }
  U.resize(n);
  S.resize(n);
  V.resize(n);
  vector<vector<double>> out(n);
  for (int c = 0; c < n; ++c)
    out[c] = svd_finish(t[c]->dim_, jobs[c], (const T*)Ub.ptr + ub[c], (const T*)Vb.ptr + vb[c],
                        sall.data() + sb[c], U[c], S[c], V[c], num, cutoff);
  return out;
}

template vector<vector<double>> svd_batched(const vector<const qtensor<double>*>&, vector<qtensor<double>>&,
                                            vector<qtensor<double>>&, vector<qtensor<double>>&, int, int);
template vector<vector<double>> svd_batched(const vector<const qtensor<cplx>*>&, vector<qtensor<cplx>>&,
                                            vector<qtensor<double>>&, vector<qtensor<cplx>>&, int, int);

// qtensor::eig (src/qtensor.cpp:1265-1313) for the tiny exact-diagonalisation
// branch: eigenvalues ascending, lowest eigenvector; returns the block size.
template <typename T>
int qtensor<T>::eig(double* val, vector<T>& vec, int sector) const {
  DMRGX_CHECK(dim_.size() == 2 && dim_[0] == dim_[1], "eig: needs a square two-leg tensor");
  bool offdiag = false;
  for (auto& s : sym_) offdiag = offdiag || s[0] != s[1];
  if (offdiag || (sector == -1 && nblocks() != 1)) return remove_symmetry().eig(val, vec, -1);
  int loc = sector == -1 ? 0 : -1;
  if (sector != -1)
    for (int i = 0; i < nblocks(); ++i)
      if (sym_[i][0] == sector) loc = i;
  if (loc < 0) {
    int d = dim_[0][sector];
    for (int i = 0; i < d; ++i) val[i] = 0;
    vec.assign(d, T(0));
    if (d) vec[0] = T(1);
    return 0;
  }
  vector<int> sh = shape(loc);
  int n = sh[0];
  vector<T> h = values();
  vector<cplx> a((size_t)n * n);
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < n; ++j) a[(size_t)i * n + j] = cplx(h[off_[loc] + (int64_t)i * n + j]);
  // the upper triangle defines the matrix (?syev / ?heev with UPLO='U')
  for (int i = 0; i < n; ++i)
    for (int j = 0; j < i; ++j) a[(size_t)i * n + j] = std::conj(a[(size_t)j * n + i]);
  vector<double> w;
  vector<cplx> v;
  host_heeig(n, a, w, v);
  for (int i = 0; i < n; ++i) val[i] = w[i];
  vec.resize(n);
  // real symmetric input: rotate the lowest vector to a real representative
  cplx ph(1, 0);
  if (dtype_of<T>::value == DT_R64) {
    int big = 0;
    for (int i = 0; i < n; ++i)
      if (std::abs(v[(size_t)i * n]) > std::abs(v[(size_t)big * n])) big = i;
    ph = std::abs(v[(size_t)big * n]) > 0 ? std::conj(v[(size_t)big * n]) / std::abs(v[(size_t)big * n]) : cplx(1);
  }
  for (int i = 0; i < n; ++i) {
    cplx z = v[(size_t)i * n] * ph;
    if constexpr (std::is_same<T, double>::value) vec[i] = z.real();
    else vec[i] = z;
  }
  return n;
}

// qtensor::trace, src/qtensor.cpp:789-821 (host reduction of device blocks)
template <typename T>
T qtensor<T>::trace(const qtensor<T>& A, bool fermion) const {
  DMRGX_CHECK(A.dim_ == dim_ && A.dir_ == dir_, "trace: tensors do not match");
  vector<T> x = values(), y = A.values();
  T ret = T(0);
  for (int l = 0; l < nblocks(); ++l)
    for (int r = 0; r < A.nblocks(); ++r) {
      if (sym_[l] != A.sym_[r]) continue;
      double sign = 1;
      if (fermion) {
        int d_b = 0, a_p = 0;
        for (size_t i = 0; i < A.dir_.size(); ++i) d_b = ss(d_b, A.sym_[r][i] * A.dir_[i]);
        for (size_t i = 0; i < dir_.size(); ++i)
          if (dir_[i] == 1) a_p = ss(a_p, sym_[l][i]);
        sign = 1 - 2 * ((ff(d_b) * ff(a_p)) % 2);
      }
      T acc = T(0);
      int64_t n = bsize(l);
      for (int64_t e = 0; e < n; ++e) acc += x[off_[l] + e] * y[A.off_[r] + e];
      ret += acc * sign;
    }
  return ret;
}

// qtensor::plus, src/qtensor.cpp:692-786 (host; small tensors only)
template <typename T>
qtensor<T>& qtensor<T>::plus(const qtensor<T>& A, const qtensor<T>& B, T alpha, T beta) {
  DMRGX_CHECK(A.dim_ == B.dim_ && A.dir_ == B.dir_, "plus: tensors do not match");
  qtensor<T> r;
  r.dim_ = A.dim_;
  r.dir_ = A.dir_;
  auto code = [&](const vector<int>& s) {
    int64_t c = 0;
    for (size_t j = 0; j < A.dim_.size(); ++j) c = c * A.dim_[j].size() + s[j];
    return c;
  };
  std::map<int64_t, std::pair<int, int>> m;
  for (int i = 0; i < A.nblocks(); ++i) m[code(A.sym_[i])].first = i + 1;
  for (int i = 0; i < B.nblocks(); ++i) m[code(B.sym_[i])].second = i + 1;
  vector<AxpbySeg> segs;
  for (auto& kv : m) {
    int ia = kv.second.first - 1, ib = kv.second.second - 1;
    r.sym_.push_back(ia >= 0 ? A.sym_[ia] : B.sym_[ib]);
  }
  r.layout();
  r.allocate();
  int k = 0;
  for (auto& kv : m) {
    int ia = kv.second.first - 1, ib = kv.second.second - 1;
    segs.push_back(AxpbySeg{r.off_[k], ia >= 0 ? A.off_[ia] : -1, ib >= 0 ? B.off_[ib] : -1, r.bsize(k)});
    ++k;
  }
  const cplx a(alpha), b(beta);
  blocks_axpby(dtype_of<T>::value == DT_C64, A.data(), B.data(), r.data(), segs, r.total_, a.real(), a.imag(),
               b.real(), b.imag(), rt().stream);
  *this = r;
  return *this;
}

// src/useful.cpp:151-194
template <typename T>
void generate_dim_sym(vector<vector<int>>& dim, vector<vector<int>>& sym, const qtensor<T>& lenv, int lnum,
                      const qtensor<T>& renv, int rnum, int sector) {
  const auto& ld = lenv.dim_;
  const auto& rd = renv.dim_;
  dim.clear();
  for (int i = 0; i < lnum; ++i) dim.push_back(ld[ld.size() - lnum + i]);
  for (int i = 0; i < rnum; ++i) dim.push_back(rd[rd.size() - rnum + i]);
  vector<int> li, ri;
  for (int i = 0; i < lnum; ++i) li.push_back((int)dim[i].size());
  for (int i = 0; i < rnum; ++i) ri.push_back((int)dim[lnum + i].size());
  vector<vector<int>> lmap, rmap;
  generate_map(lmap, li);
  generate_map(rmap, ri);
  sym.clear();
  if (sector != -1) {
    for (int s = 0; s < symmetry(); ++s) {
      int t = ss(sector, -s);
      for (auto& l : lmap)
        for (auto& r : rmap) {
          int a = 0, b = 0;
          for (int x : l) a = ss(a, x);
          for (int x : r) b = ss(b, x);
          if (a == s && b == t) {
            vector<int> v(l);
            v.insert(v.end(), r.begin(), r.end());
            sym.push_back(v);
          }
        }
    }
  } else {
    sym.assign(1, vector<int>(dim.size(), 0));
  }
}

template struct tensor<double>;
template struct tensor<cplx>;
template class qtensor<double>;
template class qtensor<cplx>;
template void generate_dim_sym(vector<vector<int>>&, vector<vector<int>>&, const qtensor<double>&, int,
                               const qtensor<double>&, int, int);
template void generate_dim_sym(vector<vector<int>>&, vector<vector<int>>&, const qtensor<cplx>&, int,
                               const qtensor<cplx>&, int, int);

}  // namespace dmrgx
