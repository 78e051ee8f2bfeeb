// Observables of a finished MPS: norm, overlap and MPO expectation value,
// contracted site by site on the device through the qtensor mirror.
#include "observables.hpp"

#include <map>

#include "heff.hpp"
#include "lanczos.hpp"

namespace dmrgx {

template <>
qtensor<double> centre_as<double>(const qtensor<double>& s) {
  return s;
}
template <>
qtensor<cplx> centre_as<cplx>(const qtensor<double>& s) {
  return s.comp();
}

// mps::inverse, src/mps.cpp:56-68
template <typename T>
mps<T> mps<T>::inverse() const {
  mps<T> ret;
  const int n = (int)state_.size();
  DMRGX_CHECK(n >= 2, "mps inverse: empty chain");
  ret.cut_ = n - cut_;
  ret.state_.resize(n);
  ret.store_.resize(n);
  for (int i = 1; i < n - 1; ++i) ret.state_[n - 1 - i] = state_[i].exchange(0, 2);
  ret.state_[0] = state_[n - 1].exchange(0, 1);
  ret.state_[n - 1] = state_[0].exchange(0, 1);
  ret.singular_ = singular_;
  return ret;
}

// mps::move_cut_left, src/mps.cpp:71-99: A_{c-2} A_{c-1} S split again with
// every singular value kept (cutoff 0); site 0 has one leg before the bond
template <typename T>
void mps<T>::move_cut_left() {
  DMRGX_CHECK(cut_ >= 2, "mps move_cut_left: cut=" + std::to_string(cut_));
  qtensor<T> two, temp;
  two.contract(state_[cut_ - 2], state_[cut_ - 1]);
  temp.contract(two, centre_as<T>(singular_));
  temp.svd(state_[cut_ - 2], singular_, state_[cut_ - 1], cut_ > 2 ? 2 : 1);
  --cut_;
}

// mps::move_cut_right, src/mps.cpp:101-125: S A_c A_{c+1} split again
template <typename T>
void mps<T>::move_cut_right() {
  DMRGX_CHECK(cut_ <= (int)state_.size() - 2, "mps move_cut_right: cut=" + std::to_string(cut_));
  qtensor<T> one, temp;
  one.contract(centre_as<T>(singular_), state_[cut_]);
  temp.contract(one, state_[cut_ + 1]);
  temp.svd(state_[cut_], singular_, state_[cut_ + 1], 2);
  ++cut_;
}

template class mps<double>;
template class mps<cplx>;

// mps::prep_calc (src/mps.cpp:33-54) on a copy: the singular values at the
// cut are absorbed into the site to their right, leaving one tensor network
// with no centre; the environments are not carried over.
template <typename T>
mps<T> prepared(const mps<T>& m) {
  mps<T> r;
  r.cut_ = m.cut_;
  r.state_ = m.state_;
  r.store_.assign(m.state_.size(), qtensor<T>());
  DMRGX_CHECK(m.cut_ >= 0 && m.cut_ < (int)m.state_.size(), "prepared: cut outside the chain");
  qtensor<T> s = centre_as<T>(m.singular_);
  qtensor<T> t;
  t.contract(s, m.state_[m.cut_]);
  r.state_[m.cut_] = t;
  return r;
}

// <m|m> without the centre (calc(mps), src/functions.cpp:395-422)
template <typename T>
T calc(const mps<T>& m, bool fermion) {
  return calc(m, m, fermion);
}

// <r|l> (calc(mps, mps), src/functions.cpp:428-460): the transfer matrix runs
// from the left end, contracting each site of l with the conjugate of r's
template <typename T>
T calc(const mps<T>& l, const mps<T>& r, bool fermion) {
  const int n = l.size();
  DMRGX_CHECK(n == r.size(), "calc(mps, mps): chain lengths differ");
  DMRGX_CHECK(n >= 2, "calc(mps, mps): empty MPS");
  qtensor<T> left, end, bra;
  bra = r.state_[0];
  bra.conjugate();
  left.contract(bra, l.state_[0], 'T', 'N', 1);
  for (int i = 1; i < n - 1; ++i) {
    end.contract(left, l.state_[i], 'N', 'N', 1);
    bra = r.state_[i];
    bra.conjugate();
    left.contract(bra, end, 'T', 'N', 2);
  }
  bra = r.state_[n - 1];
  bra.conjugate();
  end.contract(bra, l.state_[n - 1], 'N', 'T', 1);
  return left.trace(end, fermion);
}

// <m|O|m> (calc(mps, mpo), src/functions.cpp:467-511): the MPO leg rides along
// the transfer matrix; leg exchanges carry the graded signs of fermionic MPOs
template <typename T>
T calc(const mps<T>& m, mpo<T>& o, bool fermion) {
  const int n = m.size();
  DMRGX_CHECK(n == o.size(), "calc(mps, mpo): chain lengths differ");
  DMRGX_CHECK(n >= 2, "calc(mps, mpo): empty MPS");
  qtensor<T> left, end, tmp;
  end.contract(o[0], m.state_[0], 'N', 'N', 1);
  tmp = m.state_[0];
  tmp.conjugate();
  left.contract(tmp, end, 'T', 'N', 1);
  for (int i = 1; i < n - 1; ++i) {
    tmp.contract(left, 1, o[i], 0, true, fermion);
    left = tmp.exchange(3, 4);
    end.contract(left, m.state_[i], 'N', 'N', 2);
    tmp = m.state_[i];
    tmp.conjugate();
    left.contract(tmp, end, 'T', 'N', 2);
    left = left.simplify();
  }
  qtensor<T> store;
  store.contract(o[n - 1], m.state_[n - 1], 'N', 'T', 1);
  tmp = m.state_[n - 1];
  tmp.conjugate();
  end.contract(tmp, store, 'N', 'N', 1);
  end = end.simplify();
  return left.trace(end, fermion);
}

namespace {

// left store of site k from the one of k-1 and the site tensor (the
// env update of update_two / move2right, src/core.cpp:266-268)
template <typename T>
qtensor<T> lenv_of(const qtensor<T>& store, qtensor<T>& W) {
  qtensor<T> t;
  t.contract(store, 1, W, 0);
  return t.exchange(3, 4);
}
template <typename T>
qtensor<T> renv_of(const qtensor<T>& store, qtensor<T>& W) {
  qtensor<T> ope = W.exchange(0, 2);
  qtensor<T> t;
  t.contract(store, 1, ope, 0, true);
  return t.exchange(0, 1);
}
template <typename T>
qtensor<T> left_sandwich(const qtensor<T>& env, const qtensor<T>& U, int num) {
  qtensor<T> vec, c = U, out;
  vec.contract(env, U, 'N', 'N', num);
  c.conjugate();
  out.contract(c, vec, 'T', 'N', num);
  return out;
}
template <typename T>
qtensor<T> right_sandwich(const qtensor<T>& env, const qtensor<T>& V, int num) {
  qtensor<T> vec, c = V, out;
  vec.contract(env, V, 'N', 'T', num);
  c.conjugate();
  out.contract(c, vec, 'N', 'N', num);
  return out;
}
template <typename T>
qtensor<T> add(const qtensor<T>& a, const qtensor<T>& b) {
  qtensor<T> r;
  r.plus(a, b, T(1), T(1));
  return r;
}

// Re sum conj(g) x over the blocks of x, g laid out as the blocks of ref
template <typename T>
double block_dot(const qtensor<T>& ref, const void* g, const qtensor<T>& x) {
  auto code = [&](const vector<int>& s) {
    int64_t c = 0;
    for (size_t j = 0; j < ref.dim_.size(); ++j) c = c * ref.dim_[j].size() + s[j];
    return c;
  };
  std::map<int64_t, int> at;
  for (int i = 0; i < ref.nblocks(); ++i) at[code(ref.sym_[i])] = i;
  const bool cx = dtype_of<T>::value == DT_C64;
  double acc = 0;
  for (int j = 0; j < x.nblocks(); ++j) {
    auto it = at.find(code(x.sym_[j]));
    if (it == at.end()) continue;  // the energy does not depend on a block the environment lacks
    double o[2];
    dev_dot((const double*)g + (cx ? 2 : 1) * ref.off_[it->second], (const double*)x.block(j), x.bsize(j), cx, o,
            rt().stream);
    acc += o[0];
  }
  return acc;
}

}  // namespace

// `two` (a two-site state [a, s, s', b]) copied into the block order
// (dim, sym) of the eigensolver: blocks matched by their sectors, blocks
// `two` lacks are zero.  *matched (optional) counts the copied blocks.
template <typename T>
qtensor<T> embed_state(const qtensor<T>& two, const vector<vector<int>>& dim, const vector<vector<int>>& sym,
                       int* matched) {
  qtensor<T> psi;
  psi.dim_ = dim;
  psi.sym_ = sym;
  psi.dir_.assign(dim.size(), 0);
  psi.layout();
  psi.allocate();
  DMRGX_CHECK(two.dim_.size() == dim.size(), "embed_state: leg counts differ");
  auto code = [&](const vector<int>& s) {
    int64_t c = 0;
    for (size_t j = 0; j < dim.size(); ++j) c = c * dim[j].size() + s[j];
    return c;
  };
  std::map<int64_t, int> at;
  for (int i = 0; i < two.nblocks(); ++i) at[code(two.sym_[i])] = i;
  int n = 0;
  for (int b = 0; b < psi.nblocks(); ++b) {
    auto it = at.find(code(psi.sym_[b]));
    if (it == at.end()) fill_zero(psi.block(b), psi.bsize(b) * sizeof(T), rt().stream);
    else {
      DMRGX_CHECK(two.bsize(it->second) == psi.bsize(b), "embed_state: two-site block shapes differ");
      copy_d2d(psi.block(b), two.block(it->second), psi.bsize(b) * sizeof(T), rt().stream);
      ++n;
    }
  }
  if (matched) *matched = n;
  return psi;
}

// the two-site state U_l S V_r of a finished MPS in the block order
// (dim, sym) of the eigensolver
template <typename T>
qtensor<T> centre_psi(const mps<T>& m, const vector<vector<int>>& dim, const vector<vector<int>>& sym) {
  const int l = m.cut_ - 1, r = m.cut_;
  qtensor<T> us, two;
  us.contract(m.state_[l], centre_as<T>(m.singular_));
  two.contract(us, m.state_[r]);
  return embed_state(two, dim, sym, nullptr);
}

template <typename T>
std::unique_ptr<HeffPlan> make_centre_plan(const qtensor<T>& Lv, const qtensor<T>& Rv, const vector<vector<int>>& dim,
                                           const vector<vector<int>>& sym, bool alpha, int prec) {
  vector<vector<int>> rmap, lmap, dim_ret, sym_ret, dim2;
  DMRGX_CHECK(Rv.get_map(rmap, dim_ret, sym_ret, dim, sym, 'N', 'T', 2, 0, false, alpha), "get_map (rmap)");
  const int64_t d_store = get_dimension(dim_ret, sym_ret);
  vector<vector<int>> psym = sym;
  DMRGX_CHECK(Lv.get_map(lmap, dim2, psym, dim_ret, sym_ret, 'N', 'T', 3, 0, true, alpha), "get_map (lmap)");
  vector<int32_t> rm, lm;
  for (auto& t : rmap) rm.insert(rm.end(), t.begin(), t.end());
  for (auto& t : lmap) lm.insert(lm.end(), t.begin(), t.end());
  vector<int64_t> rsz(Rv.nblocks()), lsz(Lv.nblocks());
  for (int b = 0; b < Rv.nblocks(); ++b) rsz[b] = Rv.bsize(b);
  for (int b = 0; b < Lv.nblocks(); ++b) lsz[b] = Lv.bsize(b);
  return make_heff(dtype_of<T>::value, prec, Rv.data(), Rv.off_.data(), rsz.data(), Rv.nblocks(), Rv.total_,
                   rm.data(), (int)rmap.size(), Lv.data(), Lv.off_.data(), lsz.data(), Lv.nblocks(), Lv.total_,
                   lm.data(), (int)lmap.size(), get_dimension(dim, sym), d_store);
}

template <typename T>
CentreOp<T> centre_operator(const mps<T>& m, mpo<T>& W, int sector, bool alpha, int prec) {
  const int n = m.size();
  const int l = m.cut_ - 1, r = m.cut_;
  DMRGX_CHECK(l >= 1 && r <= n - 2, "centre operator: the centre must have a site on each side");
  CentreOp<T> c;
  qtensor<T> S = left_sandwich(W[0], m.state_[0], 1);
  for (int k = 1; k < l; ++k) S = left_sandwich(lenv_of(S, W[k]), m.state_[k], 2);
  c.lenv = lenv_of(S, W[l]);
  qtensor<T> Q = right_sandwich(W[n - 1], m.state_[n - 1], 1);
  for (int k = n - 2; k > r; --k) Q = right_sandwich(renv_of(Q, W[k]), m.state_[k], 2);
  c.renv = renv_of(Q, W[r]);
  vector<vector<int>> dim, sym;
  generate_dim_sym(dim, sym, c.lenv, 2, c.renv, 2, sector);
  c.psi = centre_psi(m, dim, sym);
  c.plan = make_centre_plan(c.lenv, c.renv, dim, sym, alpha, prec);
  return c;
}

template <typename T>
double centre_derivative(const mps<T>& m, mpo<T>& W, mpo<T>& dW, int sector, bool alpha, int prec, double* energy) {
  const int n = m.size();
  const int l = m.cut_ - 1, r = m.cut_;
  DMRGX_CHECK(l >= 1 && r <= n - 2, "centre_derivative: the centre must have a site on each side");
  // stores from the two ends and their tangents along dW
  qtensor<T> S = left_sandwich(W[0], m.state_[0], 1), dS = left_sandwich(dW[0], m.state_[0], 1);
  for (int k = 1; k < l; ++k) {
    qtensor<T> e = lenv_of(S, W[k]);
    qtensor<T> de = add(lenv_of(dS, W[k]), lenv_of(S, dW[k]));
    S = left_sandwich(e, m.state_[k], 2);
    dS = left_sandwich(de, m.state_[k], 2);
  }
  qtensor<T> Lv = lenv_of(S, W[l]), dL = add(lenv_of(dS, W[l]), lenv_of(S, dW[l]));
  qtensor<T> Q = right_sandwich(W[n - 1], m.state_[n - 1], 1), dQ = right_sandwich(dW[n - 1], m.state_[n - 1], 1);
  for (int k = n - 2; k > r; --k) {
    qtensor<T> e = renv_of(Q, W[k]);
    qtensor<T> de = add(renv_of(dQ, W[k]), renv_of(Q, dW[k]));
    Q = right_sandwich(e, m.state_[k], 2);
    dQ = right_sandwich(de, m.state_[k], 2);
  }
  qtensor<T> Rv = renv_of(Q, W[r]), dR = add(renv_of(dQ, W[r]), renv_of(Q, dW[r]));
  vector<vector<int>> dim, sym;
  generate_dim_sym(dim, sym, Lv, 2, Rv, 2, sector);
  const int64_t d = get_dimension(dim, sym);
  qtensor<T> psi = centre_psi(m, dim, sym);
  // H_eff plan of (Lv, Rv) and the adjoint for the output cotangent psi
  auto plan = make_centre_plan(Lv, Rv, dim, sym, alpha, prec);
  hipStream_t s = rt().stream;
  const bool cx = dtype_of<T>::value == DT_C64;
  DevBuf out(std::max<int64_t>(d, 1) * sizeof(T)), gR(std::max<int64_t>(Rv.total_, 1) * sizeof(T)),
      gL(std::max<int64_t>(Lv.total_, 1) * sizeof(T));
  plan->apply((const double*)psi.data(), (double*)out.ptr, s);
  double e[2], nn[2];
  dev_dot((const double*)psi.data(), (const double*)out.ptr, d, cx, e, s);
  dev_dot((const double*)psi.data(), (const double*)psi.data(), d, cx, nn, s);
  plan->adjoint((const double*)psi.data(), (const double*)psi.data(), nullptr, (double*)gR.ptr, (double*)gL.ptr, s);
  const double der = block_dot(Lv, gL.ptr, dL) + block_dot(Rv, gR.ptr, dR);
  if (energy) *energy = e[0] / nn[0];
  return der / nn[0];
}

template qtensor<double> embed_state(const qtensor<double>&, const vector<vector<int>>&, const vector<vector<int>>&,
                                     int*);
template qtensor<cplx> embed_state(const qtensor<cplx>&, const vector<vector<int>>&, const vector<vector<int>>&, int*);
template CentreOp<double> centre_operator(const mps<double>&, mpo<double>&, int, bool, int);
template CentreOp<cplx> centre_operator(const mps<cplx>&, mpo<cplx>&, int, bool, int);
template double centre_derivative(const mps<double>&, mpo<double>&, mpo<double>&, int, bool, int, double*);
template double centre_derivative(const mps<cplx>&, mpo<cplx>&, mpo<cplx>&, int, bool, int, double*);

template mps<double> prepared(const mps<double>&);
template mps<cplx> prepared(const mps<cplx>&);
template double calc(const mps<double>&, bool);
template cplx calc(const mps<cplx>&, bool);
template double calc(const mps<double>&, const mps<double>&, bool);
template cplx calc(const mps<cplx>&, const mps<cplx>&, bool);
template double calc(const mps<double>&, mpo<double>&, bool);
template cplx calc(const mps<cplx>&, mpo<cplx>&, bool);

}  // namespace dmrgx
