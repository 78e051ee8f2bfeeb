// FP64 thick-restart Lanczos on device-resident vectors.
#pragma once

#include <functional>
#include <vector>

#include "common.hpp"
#include "host_eig.hpp"

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
// `evec` (device, n_vec x n elements) receives the n_vec lowest Ritz vectors
// (n_vec <= ncv), the lowest normalised.  `early`: the test also runs after
// every step, not only when the basis is full (ARPACK tests at the end of a
// cycle only), so a start vector close to an eigenvector stops within a few
// matvecs.  `fp32_basis`: keep the Krylov basis in FP32 (operator input and
// output, Gram-Schmidt sums and Ritz vectors stay FP64), which halves the
// Gram-Schmidt traffic of a low-precision solve; used only when tol >= 1e-6,
// above the basis' orthogonality level.
EigshResult eigsh(const MatVec& op, int64_t n, bool cplx, int nev, int ncv, double tol,
                  int max_restart, uint64_t seed, const double* v0, double* evec, int n_vec = 1,
                  bool early = false, bool fp32_basis = false);

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
// When `retire` is given, a chain whose Ritz pairs pass the test at the end of
// a restart cycle is finished there (its result and vectors are final) and
// retire(active) is called with the chains still iterating, so that the
// caller can shrink `op` to them; the retired chains' segments of later
// vectors are ignored (they may read as zero).
using RetireFn = std::function<void(const std::vector<char>& active)>;
std::vector<EigshResult> eigsh_batched(const MatVec& op, const std::vector<int64_t>& len, bool cplx, int nev,
                                       int ncv, double tol, int max_restart, const std::vector<uint64_t>& seed,
                                       const double* v0, double* evec, int n_vec = 1,
                                       const RetireFn& retire = nullptr);
// Whether eigsh_batched replays its Lanczos steps as hipGraphs (default off:
// a DMRG step's plan lives for a few restart cycles only, and capturing plus
// instantiating the graphs costs more than the replays save; the results are
// identical either way).
bool lanczos_graphs();
void set_lanczos_graphs(bool on);
// Jacobi-Davidson refinement of the lowest eigenpair of every chain: the FP64
// Rayleigh-Ritz basis starts as the p approximate eigenvectors X (p columns
// of the concatenated length, as eigsh writes them) and grows by one vector
// per outer step, the correction s from a few conjugate-gradient steps on
// P (H_lo - theta) P s = r with the cheap operator op_lo (P: the projector
// orthogonal to the basis, r = H64 u - theta u), so each outer step costs one
// FP64 matvec and at most `inner` low-precision ones.  Stops when every
// chain's FP64 residual passes tol max(|theta|, eps^(2/3)) or after max_outer
// steps.  X column 0 receives the refined lowest vector (normalised); returns
// the Ritz values; *n64 / *nlo count the matvecs of each operator.
// res_out / conv_out (optional) receive each chain's final FP64 residual norm
// |H64 u - theta u| and whether it passed the test.
std::vector<double> jd_polish_batched(const MatVec& op_lo, const MatVec& op64, const std::vector<int64_t>& len,
                                      bool cplx, int p, double* X, double tol, int max_outer, int inner, int* n64,
                                      int* nlo, std::vector<double>* res_out = nullptr,
                                      std::vector<char>* conv_out = nullptr);
// |H x_c - theta_c x_c| for every chain c of the concatenated vector x (one
// call of `op`, computed explicitly, not from Lanczos coefficients).
std::vector<double> residual_batched(const MatVec& op, const std::vector<int64_t>& len, bool cplx, const double* x,
                                     const std::vector<double>& theta);
}  // namespace dmrgx
