// The 4e superconductor chain of 4e.cpp (Z2 x Z2 x Z2 sectors, graded
// fermionic MPOs): the MPO builders, the operators of its observables and
// the correlation functions its main prints.
#pragma once

#include <string>

#include "qtensor.hpp"

namespace dmrgx {

// H_4e1 (second = false, 4e.cpp:397-487) and H_4e2 (second = true,
// 4e.cpp:490-580): bond 9, 16 local states in 8 sectors of dimension 2
qtensor<cplx> H_4e(bool second, double t1, double t3, double f0, double f1, double p1, double frac);

// c_up1 .. c_pair_dagger2 (4e.cpp:587-779): bond-1 operators on a two-site tensor
enum class Op4e { up1, down1, up2, down2, pair1, pair2, pairpair, pair_dagger1, pair_dagger2 };
qtensor<cplx> op_4e(Op4e which);

// the main's MPO (4e.cpp:49-59) over `sites` tensors (two physical sites each)
// with p = (t1, t3, f0, f1, frac), and its periodic edge terms (4e.cpp:64-68)
mpo<cplx> mpo_4e(int sites, const double* p);
void edges_4e(int sites, const double* p, qtensor<cplx>& first, qtensor<cplx>& last);

// the lines of the main's data file (4e.cpp:71-196): label, abscissa
// (sites, or the bond / site position), calc(mps, mpo) value; m prepared
struct Corr4e {
  std::string label;
  double x;
  cplx value;
};
vector<Corr4e> correlations_4e(const mps<cplx>& m, bool fermion);

}  // namespace dmrgx
