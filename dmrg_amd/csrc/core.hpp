// DMRG step functions and sweep schedules over device-resident qtensors
// (src/core.cpp, src/functions.cpp).
#pragma once

#include <fstream>
#include <functional>
#include <memory>
#include <string>
#include <vector>

#include "heff.hpp"
#include "qtensor.hpp"

namespace dmrgx {

struct RunConfig {
  int prec = PREC_FP64;        // arithmetic of the H_eff matvecs inside the eigensolver
  bool physical_signs = true;  // bosonic models: get_map alpha = 1 (reference rule when false)
  bool fermion = false;        // the model is fermionic (the reference's FERMION switch)
  int sector = 0;
  uint64_t seed = 1234;
  double tol = 1e-10;          // Lanczos tolerance (FP64 eigensolve)
  double tol_lo = 1e-6;        // Lanczos tolerance of a reduced-precision eigensolve
  int refine_k = 8;            // refinement of a reduced-precision eigensolve: 2 refine_k inner steps
  int polish_cycles = 12;      // FP64 refinement steps at most, until the FP64 residual passes `tol`
  // start each sweep step's eigensolve from the state the previous step left
  // (White's prediction) and converge the lowest pair only (E1..E5 read 0)
  bool warm_start = false;
  int nev = NEV;
  int nev_lo = 1;              // eigenpairs whose residual a reduced-precision solve tests
  int max_restart = 5000;
  int verbose = 0;             // 1: a line per step on stdout, 2: on stderr
  // when set, every step appends to "energy" + output_suffix and "singular" +
  // output_suffix in the working directory, as the reference's drivers do
  std::string output_suffix;
};

struct StepLog {
  int l = 0, r = 0;
  double e[NEV] = {0, 0, 0, 0, 0, 0};
  double entropy = 0, trunc = 0;
  int64_t d = 0;
  int n_matvec = 0, n_polish = 0;
  bool warm = false;  // the eigensolve started from the predicted state
  double flops = 0;  // algorithmic matvec flops of this step (all matvecs)
  double t_total = 0, t_eig = 0, t_svd = 0, t_env = 0;
  double t_plan = 0, t_lanczos = 0, t_polish = 0;  // inside t_eig: maps and plans, eigsh, FP64 refinement
  int svd_sweeps = 0;     // Jacobi sweeps of the block SVD
  double t_block_svd = 0;  // the block SVD alone (inside t_svd)
  std::vector<double> s;  // leading singular values
  int n_eval = NEV;       // eigenvalues the step reports (d when it diagonalised exactly)
};

template <typename T>
class BatchEngine;

// The two-site sweep schedule of dmrg2 (src/functions.cpp:234-299): `first`
// is the infinite-size start on the two edge sites, `step(l, r, cutoff)` an
// update_two; ladder=false skips the pre-sweep cutoff ladder 100, 200, ...
// Returns the last step's value.
double schedule_dmrg2(int L, int cutoff, int sweep, bool ladder, const std::function<double()>& first,
                      const std::function<double(int, int, int)>& step);

template <typename T>
class Engine {
 public:
  Engine(const mpo<T>& m, const RunConfig& cfg);
  double two_sites(int l, int r, int cutoff);
  double update_two(int l, int r, int cutoff);
  double move2right(int l, int r, int cutoff);
  double move2left(int l, int r, int cutoff);
  double dmrg(int cutoff, int sweep);
  double dmrg2(int cutoff, int sweep, bool ladder = true);

  mpo<T> mpo_;
  mps<T> mps_;
  RunConfig cfg;
  std::vector<StepLog> log;
  std::function<void(const StepLog&)> on_step;

 private:
  friend class BatchEngine<T>;
  bool alpha_on() const { return !cfg.physical_signs || cfg.fermion; }
  qtensor<T> make_lenv(int l);
  qtensor<T> make_renv(int r);
  // Lanczos on (lenv, renv) with r_num / l_num contracted legs; fills vals
  // and psi.  `guess` (may be null): a two-site state to start from.
  // With ledge / redge (periodic chains) the operator is the bulk one plus
  // the edge term L_edge W_id W_id R_edge (src/core.cpp:27-30).
  void solve(const qtensor<T>& lenv, const qtensor<T>& renv, int lnum, int rnum, int r_num, int l_num,
             vector<vector<int>>& dim, vector<vector<int>>& sym, double* vals, qtensor<T>& psi, StepLog& lg,
             bool allow_ed, int ed_which, const qtensor<T>* guess = nullptr, const qtensor<T>* ledge = nullptr,
             const qtensor<T>* redge = nullptr);
  // the edge environments of a step: from the stored ones through the
  // identity operator (src/core.cpp:193-199, 302-308, 398-404)
  qtensor<T> make_ledge(int l);
  qtensor<T> make_redge(int r, int from);
  // the edge environment updates after the SVD (src/core.cpp:272-281)
  void update_ledge(int l, const qtensor<T>& ledge, qtensor<T> U, int num);
  void update_redge(int r, const qtensor<T>& redge, qtensor<T> V, int num);
  // the two-site state on (l, r) predicted from the previous step's
  // U S V when that step was the neighbouring bond; false otherwise
  bool predict(int l, int r, qtensor<T>& guess);
  void finish(StepLog& lg, const vector<double>& s, int cutoff);
  std::unique_ptr<std::ofstream> f_energy_, f_singular_;
};

}  // namespace dmrgx
