// Observables of a finished MPS: norm, overlap and MPO expectation value,
// contracted site by site on the device through the qtensor mirror.
#include "observables.hpp"

namespace dmrgx {

template <>
qtensor<double> centre_as<double>(const qtensor<double>& s) {
  return s;
}
template <>
qtensor<cplx> centre_as<cplx>(const qtensor<double>& s) {
  return s.comp();
}

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

template mps<double> prepared(const mps<double>&);
template mps<cplx> prepared(const mps<cplx>&);
template double calc(const mps<double>&, bool);
template cplx calc(const mps<cplx>&, bool);
template double calc(const mps<double>&, const mps<double>&, bool);
template cplx calc(const mps<cplx>&, const mps<cplx>&, bool);
template double calc(const mps<double>&, mpo<double>&, bool);
template cplx calc(const mps<cplx>&, mpo<cplx>&, bool);

}  // namespace dmrgx
