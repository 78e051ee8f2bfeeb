// FP64 thick-restart Lanczos on device-resident vectors.
#pragma once

#include <functional>
#include <vector>

#include "common.hpp"

namespace dmrgx {

// y = H x on device vectors of n elements (complex interleaved if cplx).
using MatVec = std::function<void(const double* x, double* y)>;

struct EigshResult {
  std::vector<double> evals;  // nev lowest, ascending
  int n_matvec = 0;
  int n_restart = 0;
  int n_reorth = 0;  // second Gram-Schmidt passes (one per step)
  bool converged = false;
  double max_resid = 0;
};

// Lowest `nev` eigenpairs of a Hermitian operator: thick-restart Lanczos with
// full reorthogonalisation (two classical Gram-Schmidt passes per step).  Converged when every wanted Ritz pair satisfies
// |beta * y_last| <= tol * max(|theta|, eps^(2/3)) (the ARPACK test).  `v0`
// (device, may be null) is the start vector; otherwise a seeded random one.
// `evec` (device, n elements) receives the lowest Ritz vector.
EigshResult eigsh(const MatVec& op, int64_t n, bool cplx, int nev, int ncv, double tol,
                  int max_restart, uint64_t seed, const double* v0, double* evec);

// Rayleigh-Ritz polish of an approximate lowest eigenvector in a small Krylov
// space K_k(H, x) built with `op`; x is replaced by the polished vector and
// the Ritz value returned.
double krylov_polish(const MatVec& op, int64_t n, bool cplx, int k, double* x, int* n_matvec);

// device vector helpers (n = number of elements, complex if cplx)
void dev_random(double* v, int64_t n, bool cplx, uint64_t seed, hipStream_t s);
double dev_norm2(const double* v, int64_t n, bool cplx, hipStream_t s);
void dev_scale(double* v, int64_t n, bool cplx, double a, hipStream_t s);
void dev_dot(const double* a, const double* b, int64_t n, bool cplx, double out[2], hipStream_t s);
// v -= V (V^H v), two passes; V: k columns of n elements at stride ld
void dev_orthogonalise(const double* V, int64_t ld, int k, double* v, int64_t n, bool cplx, hipStream_t s);

// The same solver on many independent operators at once (one concatenated
// vector space; chain c owns elements [off_c, off_c + len_c), off_0 = 0 and
// the chains contiguous).  `op` applies every chain's operator in one call.
// Each chain starts from dev_random(seed_c) (or its segment of v0), follows
// eigsh's algorithm and stopping test with the chains' reductions and vector
// updates segmented in single launches, and gets its own result; all chains
// take the same number of steps until the last one has converged.  Chains of
// at most 64 elements are diagonalised densely, as in eigsh.
std::vector<EigshResult> eigsh_batched(const MatVec& op, const std::vector<int64_t>& len, bool cplx, int nev,
                                       int ncv, double tol, int max_restart, const std::vector<uint64_t>& seed,
                                       const double* v0, double* evec);
// krylov_polish of every chain's segment of x at once; returns the Ritz values
std::vector<double> krylov_polish_batched(const MatVec& op, const std::vector<int64_t>& len, bool cplx, int k,
                                          double* x, int* n_matvec);

// Symmetric eigen-decomposition of a small dense matrix on the host
// (cyclic Jacobi).  a: n x n row-major, overwritten; w ascending; v columns.
void host_symeig(int n, std::vector<double>& a, std::vector<double>& w, std::vector<double>& v);
// Hermitian version (complex, row-major), used by the tiny exact-diagonalisation branch.
void host_heeig(int n, std::vector<cplx>& a, std::vector<double>& w, std::vector<cplx>& v);

}  // namespace dmrgx
