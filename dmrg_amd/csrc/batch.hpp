// Many independent chains of one model swept in lockstep: every step builds
// one H_eff plan over all chains' get_map tasks (one grouped launch per pass),
// runs the segmented batched Lanczos over the concatenated two-site states and
// one block SVD over all chains' blocks.  The reference runs one chain per
// process (ARPACK is not re-entrant); each chain here follows the two-site
// schedule of src/functions.cpp:234-299 and the step src/core.cpp:182-283.
#pragma once

#include <memory>
#include <vector>

#include "core.hpp"

namespace dmrgx {

struct BatchStepLog {
  int l = 0, r = 0;
  int64_t d = 0;  // summed two-site dimension of all chains
  int n_matvec = 0, n_polish = 0;  // batched matvec calls
  double flops = 0;  // algorithmic flops of all chains' matvecs
  double t_total = 0, t_eig = 0, t_svd = 0, t_env = 0, t_plan = 0;
};

template <typename T>
class BatchEngine {
 public:
  // one chain per MPO; cfg as for Engine (every chain uses cfg.seed)
  BatchEngine(const std::vector<mpo<T>>& m, const RunConfig& cfg);
  // the two-site schedule on every chain; returns chain 0's last E0
  double dmrg2(int cutoff, int sweep, bool ladder = true);
  double update_two(int l, int r, int cutoff);

  std::vector<std::unique_ptr<Engine<T>>> chains;  // each chain's MPO, MPS and step log
  RunConfig cfg;
  std::vector<BatchStepLog> log;

 private:
  void solve(std::vector<qtensor<T>>& lenv, std::vector<qtensor<T>>& renv, std::vector<qtensor<T>>& psi,
             std::vector<StepLog>& lg, BatchStepLog& bl);
};

}  // namespace dmrgx
