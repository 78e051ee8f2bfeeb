// Observables of a finished MPS (src/functions.cpp:394-511, src/mps.cpp:33-54).
#pragma once

#include <memory>

#include "heff.hpp"
#include "qtensor.hpp"

namespace dmrgx {

// the centre (singular values, real) in the element type of the MPS
template <typename T>
qtensor<T> centre_as(const qtensor<double>& s);

// `two` (a two-site state [a, s, s', b]) copied into the eigensolver's block
// order (dim, sym): blocks matched by their sectors, the others zero;
// *matched (optional) counts the copied blocks
template <typename T>
qtensor<T> embed_state(const qtensor<T>& two, const vector<vector<int>>& dim, const vector<vector<int>>& sym,
                       int* matched = nullptr);

// copy of m with the centre absorbed into site cut_ (mps::prep_calc)
template <typename T>
mps<T> prepared(const mps<T>& m);

// calc(mps): <m|m>; calc(mps, mps): <r|l>; calc(mps, mpo): <m|O|m>.  The
// arguments are prepared MPSs (no centre).  `fermion` switches the graded
// signs of the leg exchanges and the final trace: the reference applies them
// for every model; the physical rule only for fermionic ones (the same switch
// as get_map's alpha).
template <typename T>
T calc(const mps<T>& m, bool fermion = true);
template <typename T>
T calc(const mps<T>& l, const mps<T>& r, bool fermion = true);
template <typename T>
T calc(const mps<T>& m, mpo<T>& o, bool fermion = true);

// Derivative of the two-site energy E = <psi|H_eff(W)|psi> / <psi|psi> at the
// centre of a finished (unprepared) MPS along an MPO direction dW (the MPO
// of the coupling change; the model builders are linear in the couplings).
// The environments and their tangents along dW are recomputed from the MPS
// site by site (no stored environments), and the derivative is
// Re(<g_L, dL> + <g_R, dR>) with g_L, g_R the adjoint cotangents of the two
// environments for the output cotangent psi.  `alpha` is get_map's sign
// switch; `prec` the arithmetic of the adjoint (FP64, or FP32 MFMA).
// Also returns E in *energy.
template <typename T>
double centre_derivative(const mps<T>& m, mpo<T>& W, mpo<T>& dW, int sector, bool alpha, int prec,
                         double* energy);

// The two-site operator at the centre of a finished (unprepared) MPS: both
// environments recomputed from the MPS site by site, the two-site state U S V
// in the eigensolver's block order, and the H_eff plan (arithmetic `prec`)
// over them; the plan reads lenv / renv, which must outlive it.
template <typename T>
struct CentreOp {
  qtensor<T> lenv, renv, psi;
  std::unique_ptr<HeffPlan> plan;
};
template <typename T>
CentreOp<T> centre_operator(const mps<T>& m, mpo<T>& W, int sector, bool alpha, int prec);

}  // namespace dmrgx
