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
  int lda = 0, ldb = 0;  // row strides of A and B ('N','T' pre-split plans); 0 means K
  // only the K range [k0, k1) is contracted (k1 == 0: the whole K); K stays
  // the row stride of k-contiguous operands
  int k0 = 0, k1 = 0;
  // >= 0: the task's M rows are rows row_map[0..M) (offset into the plan's
  // row-map table) of A and of the output block, whose full row count is not
  // needed ('N' A only); -1: rows 0..M
  int64_t row_map = -1;
};

// Where a plane-output block's rows land in a padded layout (PadLayout): the
// block (M x N, row-major at c_off) is rows [r0, r0 + M) of a run of rows of
// N elements, q of which make one row of a padded block of row stride kp at
// base.  Row `row` (after any row map) starts at
// base + ((r0 + row) / q) * kp + ((r0 + row) % q) * N.
struct OutPad {
  int64_t base;
  int r0, q, kp, pad;
};

// Device work decomposition (shared with the kernels).
struct GemmSeg {
  int64_t a_off, b_off;
  int K;        // the task's full K (leading dimension when k-contiguous)
  int k0, k1;   // k range handled by this segment
  int alpha;
  int lda, ldb;  // row strides of A and B in the pre-split kernel (>= K)
  int pad;
};
struct GemmItem {
  int64_t c_off;
  int64_t row_map;  // offset of the block's row indices in the row-map table (A rows, and output rows
                    // unless the item writes a compact partial tile), or -1
  int M, N;
  int m0, n0;
  int seg_begin, seg_end;
  int flags;  // bit0: shared tile (atomic accumulate), bit1: beta (exclusive accumulate)
  int pad;
  OutPad op;  // plane output into a padded layout (op.kp == 0: C row-major at c_off)
};

// ITEM_PARTIAL: the item writes a K-chunk partial tile into the plan's
// workspace (c_off is then a workspace offset); a reduction pass sums the
// chunks into C.  ITEM_BETA: single-writer tile accumulating onto C.
enum : int { ITEM_PARTIAL = 1, ITEM_BETA = 2 };

struct RedTask {
  int64_t c_off, w_off, len;  // len = M*N elements
  int64_t row_map;            // row-map offset of the output rows (-1: rows 0..M)
  int nch, beta;
  int ncol, pad;              // N: output elements per row
  OutPad op;                  // as GemmItem::op
};

// Output format of the pre-split BF16x3 kernel: FP32, bf16 hi/lo planes
// (re_hi, re_lo[, im_hi, im_lo], stride c_ps; outputs may not accumulate), or FP64.
enum { OUT_F32 = 0, OUT_PLANES = 1, OUT_F64 = 2 };

// The work decomposition of a plan as host arrays (what the plan uploads).
struct GemmTables {
  std::vector<GemmItem> items;
  std::vector<GemmSeg> segs;
  std::vector<RedTask> red;
  std::vector<int64_t> red_start;  // first reduction block of each RedTask
};

struct GemmPlan {
  int dtype = DT_R64;
  int prec = PREC_FP64;
  char ta = 'N', tb = 'T';
  int tile = 64;    // output tile edge (64, or 128 for the pre-split BF16x3 kernel)
  bool planes = false;  // operands are pre-split bfloat16 planes
  int n_items = 0, n_segs = 0, grid = 0;
  std::vector<std::pair<int64_t, int64_t>> zero_ranges;  // [off, len) in C elements
  int64_t ws_elems = 0;                 // split-K workspace, C elements
  int n_red = 0, red_blocks = 0;
  // device addresses the kernels read: the work items, their K segments, the
  // split-K workspace, the reduction tasks and their first blocks; they point
  // into the owned buffers below, or into caller-provided memory for a plan
  // built with host tables only
  const GemmItem* items = nullptr;
  const GemmSeg* segs = nullptr;
  void* ws = nullptr;
  const RedTask* red = nullptr;
  const int64_t* red_start = nullptr;
  const int32_t* row_maps = nullptr;  // the row-map table of row-mapped tasks
  std::unique_ptr<DevBuf> own_items, own_segs, own_ws, own_red, own_red_start, own_row_maps;
  double flops = 0;        // algorithmic real flops (c*M*N*K summed)
  bool any_beta = false;   // some output block accumulates onto existing C
  // built with an output layout: every output block maps into it (plane
  // output then lands in that padded layout, which the caller zeroes once:
  // the plan writes neither the padding nor the structurally zero ranges)
  bool out_padded = false;
};

// Build a plan; `extra_zero` lists C ranges with no task that must read as 0.
// split_target > 0: cut K so that the launch has about that many work items;
// 0: two per CU; < 0: balanced over -split_target resident work-groups (whole
// output blocks cut into two K halves where that shortens the launch).
// With `host_only` the plan allocates and uploads nothing: its decomposition
// is returned in *host_only and the caller sets the device addresses.
// `row_maps` holds the row indices that row-mapped tasks point into (tasks of
// one output block with different row maps form separate groups).
// `out_pad` (plane output only): the padded layout C's planes are stored in;
// when some output block does not map into it as OutPad describes, the plan
// is built without it (out_padded false).
struct PadLayout;
std::unique_ptr<GemmPlan> make_gemm_plan(const std::vector<GemmTask>& tasks, int dtype, int prec,
                                         char ta, char tb,
                                         const std::vector<std::pair<int64_t, int64_t>>& extra_zero,
                                         int split_target = 0, bool planes = false,
                                         GemmTables* host_only = nullptr,
                                         const std::vector<int32_t>* row_maps = nullptr,
                                         const PadLayout* out_pad = nullptr);
// Bytes of the workspace of a plan (split-K partial tiles).
inline size_t gemm_ws_bytes(const GemmPlan& p) {
  return (size_t)p.ws_elems * (p.prec == PREC_FP64 ? 8 : 4) * (p.dtype == DT_C64 ? 2 : 1);
}

// Execute: A, B, C are device pointers with storage precision of the plan
// (double for PREC_FP64, float for PREC_FP32 / PREC_BF16X3; complex
// interleaved when dtype == DT_C64).
void run_gemm_plan(const GemmPlan& p, const void* A, const void* B, void* C, hipStream_t s,
                   hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr);

// Pre-split BF16x3 execution ('N','T' only): A and B are bfloat16 planes
// (re_hi, re_lo[, im_hi, im_lo]) of a_stride / b_stride elements each (both
// multiples of 8); C is FP32 or FP64 (complex interleaved when dtype ==
// DT_C64) or, for OUT_PLANES, bf16 planes of stride c_ps.
// hi: the coarse product of the hi planes alone (about 3 significant digits)
void run_gemm_planes(const GemmPlan& p, const uint16_t* A, int64_t a_stride, const uint16_t* B, int64_t b_stride,
                     void* C, int out, int64_t c_ps, hipStream_t s, hipEvent_t ev_begin = nullptr,
                     hipEvent_t ev_end = nullptr, bool hi = false);
// Stride of a bf16 plane of n elements (a multiple of 8: 16-byte rows).
inline int64_t plane_stride(int64_t n) { return (n + 7) / 8 * 8; }

// Padded plane layout of an operand: its blocks (rows x K, row-major at
// src_off) are stored with rows of Kp = K rounded up to a multiple of 8,
// zero-filled, so that every row starts on a 16-byte boundary of a bf16 plane.
struct PadBlk {
  int64_t src_off, dst_off;
  int rows, K, Kp, pad;
};
struct PadLayout {
  std::vector<PadBlk> blk;   // ascending src_off
  int64_t total = 0;         // padded elements
  bool identity = true;      // dst_off == src_off and Kp == K for every block
  std::unique_ptr<DevBuf> d_blk;
  int64_t map(int64_t src_off) const;  // padded offset of the block at src_off
  int kp(int64_t src_off) const;
};
// Layout of the distinct operand blocks {src_off, rows, K} (duplicates allowed).
std::unique_ptr<PadLayout> make_pad_layout(std::vector<PadBlk> blocks);
// src (FP64 or FP32, complex if cplx) -> padded hi/lo bf16 planes of stride ps.
void split_padded_f64(const double* src, uint16_t* dst, int64_t ps, const PadLayout& L, bool cplx, hipStream_t s);
void split_padded_f32(const float* src, uint16_t* dst, int64_t ps, const PadLayout& L, bool cplx, hipStream_t s);

// The GEMM kernel of a plan alone (no zero fills, no split-K reduction): what
// the per-launch timing of a pass repeats back to back.
void launch_plan_kernel(const GemmPlan& p, const void* A, const void* B, void* C, hipStream_t s);
void launch_planes_kernel(const GemmPlan& p, const uint16_t* A, int64_t a_stride, const uint16_t* B,
                          int64_t b_stride, void* C, int out, int64_t c_ps, hipStream_t s, bool hi = false);

// Precision conversion helpers (n counts real scalars).
void convert_f64_to_f32(const double* src, float* dst, int64_t n, hipStream_t s);
void convert_f32_to_f64(const float* src, double* dst, int64_t n, hipStream_t s);

}  // namespace dmrgx
