// Effective-Hamiltonian plan: the two task-list passes of av().
#pragma once

#include <memory>
#include <vector>

#include "common.hpp"
#include "gemm.hpp"

namespace dmrgx {

struct HeffPlan {
  int dtype = DT_R64;
  int prec = PREC_FP64;
  int64_t n_psi = 0, n_store = 0;
  const void* renv = nullptr;  // FP64 env storage (caller-owned)
  const void* lenv = nullptr;
  std::unique_ptr<DevBuf> renv_lo, lenv_lo;  // FP32 copies for reduced precision
  std::unique_ptr<DevBuf> store, in_lo, out_lo;
  // BF16x3: bfloat16 planes of the environments, psi and store
  std::unique_ptr<DevBuf> renv_pl, lenv_pl, in_pl, store_pl;
  int64_t renv_ps = 0, lenv_ps = 0, psi_ps = 0, store_ps = 0;
  // padded plane layouts of the four operands (rows rounded up to 8 elements)
  std::unique_ptr<PadLayout> renv_lay, lenv_lay, psi_lay, store_lay;
  bool store_direct = false;  // the R pass writes store's planes itself
  std::unique_ptr<GemmPlan> rplan, lplan;
  double flops = 0;
  // ev: optional {R begin, R end, L begin, L end} events around the two GEMM launches
  void apply(const double* in, double* out, hipStream_t s, hipEvent_t* ev = nullptr);
};

// get_map tuples -> GEMM tasks (M = blk_size / col, N = row, K = col).
std::vector<GemmTask> tasks_from_map(const int32_t* map, int n_task, const int64_t* blk_off,
                                     const int64_t* blk_size);
// C ranges in [0, n_out) that no task writes (they must read as zero).
std::vector<std::pair<int64_t, int64_t>> uncovered(const std::vector<GemmTask>& t, int64_t n_out);

std::unique_ptr<HeffPlan> make_heff(int dtype, int prec, const void* renv, const int64_t* rblk_off,
                                    const int64_t* rblk_size, int n_rblk, int64_t renv_len,
                                    const int32_t* rmap, int n_rtask, const void* lenv,
                                    const int64_t* lblk_off, const int64_t* lblk_size, int n_lblk,
                                    int64_t lenv_len, const int32_t* lmap, int n_ltask, int64_t n_psi,
                                    int64_t n_store);

}  // namespace dmrgx

// opaque C handle: the plan plus any device buffers it owns (synthetic envs)
struct dmrgx_heff {
  std::unique_ptr<dmrgx::HeffPlan> p;
  std::vector<std::shared_ptr<dmrgx::DevBuf>> keep;
};
