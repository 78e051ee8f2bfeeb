// Observables of a finished MPS (src/functions.cpp:394-511, src/mps.cpp:33-54).
#pragma once

#include "qtensor.hpp"

namespace dmrgx {

// the centre (singular values, real) in the element type of the MPS
template <typename T>
qtensor<T> centre_as(const qtensor<double>& s);

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

}  // namespace dmrgx
