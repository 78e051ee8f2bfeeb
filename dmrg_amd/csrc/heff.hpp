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
  double flops = 0;       // algorithmic flops of one apply (every task in full, as the reference)
  double flops_exec = 0;  // flops the pass plans execute (structurally zero slices skipped)
  int64_t renv_len = 0, lenv_len = 0;
  std::vector<GemmTask> rtasks, ltasks;  // the get_map tasks (element offsets of the plain layout)
  // the tasks as the pass plans execute them (padded-plane offsets for
  // BF16x3), the output ranges no task writes, and the chosen split targets
  std::vector<GemmTask> rexec, lexec;
  // the original task each executed task comes from, and the row indices
  // that the row-mapped R-pass tasks point into
  std::vector<int> rexec_src, lexec_src;
  std::vector<int32_t> row_maps;
  const std::vector<int32_t>* maps() const { return row_maps.empty() ? nullptr : &row_maps; }
  const PadLayout* r_out_layout() const;
  std::vector<std::pair<int64_t, int64_t>> rzero, lzero;
  int r_split = 0, l_split = 0;
  // rebuild the two pass plans with the given split-K targets (work items
  // per launch; 0: two per CU; negative: balanced over that many resident
  // work-groups)
  void replan(int r_target, int l_target);
  // rebuild the pass plans over the tasks with keep_r[i] / keep_l[i] set
  // (outputs of the dropped tasks read as zero); empty masks: all tasks
  void restrict_tasks(const std::vector<char>& keep_r, const std::vector<char>& keep_l);
  // pick each pass plan's split-K target by timing candidates once per
  // sector-shape histogram (cached for the process)
  void autotune(hipStream_t s);
  // adjoint plans, built on the first adjoint call: gradients of the env
  // blocks (gR, gL) and of the pass inputs (gP for psi, gS for store), plus an
  // FP32 R pass when the forward plan works on bf16 planes
  std::unique_ptr<GemmPlan> ad_gr, ad_gl, ad_gp, ad_gs, ad_fr;
  // the periodic edge term (the reference's PBC build, src/core.cpp:27-30):
  // a second operator over the edge environments whose result apply() adds
  // to the bulk one; edge_out holds its result
  std::unique_ptr<HeffPlan> edge;
  std::unique_ptr<DevBuf> edge_out;
  // ev: optional {R begin, R end, L begin, L end} events around the two GEMM
  // launches; coarse (BF16x3 plans): the hi planes' product alone, about
  // three significant digits, for the refinement's inner solves
  void apply(const double* in, double* out, hipStream_t s, hipEvent_t* ev = nullptr, bool coarse = false);
  // Per-launch time (ms) of each pass's GEMM kernel alone: one apply of `in`
  // (so that every intermediate holds its data), then `iters` back-to-back
  // launches of the R-pass kernel and of the L-pass kernel between events
  // (the launches rewrite the same outputs; out gets the apply's result again
  // by a final apply)
  void time_kernels(const double* in, double* out, int iters, hipStream_t s, double* ms_r, double* ms_l);
  // Adjoint of out = H_eff(psi; renv, lenv): for a cotangent g of out, the
  // cotangents of psi, renv and lenv (each may be null): for every task
  // C += alpha A B^T, gA += alpha g conj(B) and gB += alpha g^T conj(A), so
  // that Re<g, dout> = Re<g_psi, dpsi> + Re<g_renv, drenv> + Re<g_lenv, dlenv>.
  // FP64 for an FP64 plan, FP32 otherwise; outputs are FP64, overwritten.
  void adjoint(const double* psi, const double* g_out, double* g_psi, double* g_renv, double* g_lenv,
               hipStream_t s);
};

// Adds the edge operator e to h: h.apply then computes bulk + edge (the
// flop counts include the edge's).  Throws when the sizes differ.
void attach_edge(HeffPlan& h, std::unique_ptr<HeffPlan> e);

// get_map tuples -> GEMM tasks (M = blk_size / col, N = row, K = col); throws
// on a block index outside [0, n_blk), negative offsets or empty rows / cols.
std::vector<GemmTask> tasks_from_map(const int32_t* map, int n_task, const int64_t* blk_off,
                                     const int64_t* blk_size, int n_blk);
// C ranges in [0, n_out) that no task writes (they must read as zero).
std::vector<std::pair<int64_t, int64_t>> uncovered(const std::vector<GemmTask>& t, int64_t n_out);

// Whether make_heff tunes the pass plans (default on) and how many
// histogram keys the tuner has timed in this process.
void set_heff_autotune(bool on);
int heff_autotune_keys();
// Whether make_heff skips the environments' structurally zero slices (default
// on): R-pass rows that are zero in every task of an output block, and L-pass
// K columns that are zero in the environment or in the R pass's output.
void set_heff_skip_zeros(bool on);

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
