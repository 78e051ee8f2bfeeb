// Grouped block GEMM over get_map-style task lists.
#pragma once

#include <memory>
#include <vector>

#include "common.hpp"

namespace dmrgx {

// One GEMM of a task list: C(MxN) (+)= alpha * op(A)(MxK) * op(B)(KxN).
// A is row-major MxK (transA 'N') or KxM ('T'); B is row-major NxK (transB
// 'T') or KxN ('N'); C is row-major MxN.  Offsets are element offsets.
struct GemmTask {
  int64_t a_off, b_off, c_off;
  int M, N, K;
  int alpha;  // +1 / -1 (the reference's fermion sign)
  int beta;   // 0: first write overwrites, 1: accumulate onto C
};

// Device work decomposition (shared with the kernels).
struct GemmSeg {
  int64_t a_off, b_off;
  int K;        // the task's full K (leading dimension when k-contiguous)
  int k0, k1;   // k range handled by this segment
  int alpha;
};
struct GemmItem {
  int64_t c_off;
  int M, N;
  int m0, n0;
  int seg_begin, seg_end;
  int flags;  // bit0: shared tile (atomic accumulate), bit1: beta (exclusive accumulate)
  int pad;
};

enum : int { ITEM_ATOMIC = 1, ITEM_BETA = 2 };

struct GemmPlan {
  int dtype = DT_R64;
  int prec = PREC_FP64;
  char ta = 'N', tb = 'T';
  int n_items = 0, n_segs = 0, grid = 0;
  DevBuf* d_items = nullptr;
  DevBuf* d_segs = nullptr;
  std::vector<std::pair<int64_t, int64_t>> zero_ranges;  // [off, len) in C elements
  double flops = 0;        // algorithmic real flops (c*M*N*K summed)
  ~GemmPlan();
};

// Build a plan; `extra_zero` lists C ranges with no task that must read as 0.
std::unique_ptr<GemmPlan> make_gemm_plan(const std::vector<GemmTask>& tasks, int dtype, int prec,
                                         char ta, char tb,
                                         const std::vector<std::pair<int64_t, int64_t>>& extra_zero,
                                         int split_target = 0);

// Execute: A, B, C are device pointers with storage precision of the plan
// (double for PREC_FP64, float for PREC_FP32 / PREC_BF16X3; complex
// interleaved when dtype == DT_C64).
void run_gemm_plan(const GemmPlan& p, const void* A, const void* B, void* C, hipStream_t s,
                   hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr);

// Precision conversion helpers (n counts real scalars).
void convert_f64_to_f32(const double* src, float* dst, int64_t n, hipStream_t s);
void convert_f32_to_f64(const float* src, double* dst, int64_t n, hipStream_t s);

}  // namespace dmrgx
