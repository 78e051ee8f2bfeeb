// Small dense symmetric / Hermitian eigensolvers on the host (no device code,
// so this pair of files also builds with a plain C++ compiler).
#pragma once

#include <complex>
#include <vector>

namespace dmrgx {

using cplx = std::complex<double>;

// Symmetric eigen-decomposition of a small dense matrix: Householder
// tridiagonalisation and implicit QL.  a: n x n row-major (symmetrised, not
// modified); w ascending; v: eigenvectors as columns (row-major n x n).
void host_symeig(int n, std::vector<double>& a, std::vector<double>& w, std::vector<double>& v);
// Hermitian version (complex, row-major) through the real symmetric embedding
// [[X, -Y], [Y, X]] of X + iY; used by the tiny exact-diagonalisation branches.
void host_heeig(int n, std::vector<cplx>& a, std::vector<double>& w, std::vector<cplx>& v);
// host_heeig, or host_symeig on the real part when `real` (then the
// eigenvectors are real, never a real vector times a phase)
void host_hermitian_eig(int n, std::vector<cplx>& a, bool real, std::vector<double>& w, std::vector<cplx>& v);

}  // namespace dmrgx
