// Task-list GEMM that allocates nothing: the work decomposition and the
// split-K partial tiles live in a fixed device array of this file instead of
// buffers from the library's pool.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>

#include "dmrgx.h"
#include "common.hpp"
#include "gemm.hpp"
#include "heff.hpp"

namespace {

// 64 MiB per device, held from the library's first use on a device until the
// process ends.  Every call on a device shares it: calls that overlap on one
// device race.
constexpr size_t SCRATCH_BYTES = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
unsigned char* g_scratch[MAX_DEVICES] = {};  // device address of `scratch` per device
std::mutex g_mu;

size_t align256(size_t b) { return (b + 255) / 256 * 256; }

}  // namespace

namespace dmrgx {

// Called by dmrgx_init with `device` current: the device address of this
// device's copy of `scratch` (the array's name in host code is only a handle).
void noalloc_setup(int device) {
  DMRGX_CHECK(device >= 0 && device < MAX_DEVICES, "no-allocation scratch: device index out of range");
  std::lock_guard<std::mutex> lk(g_mu);
  if (g_scratch[device]) return;
  void* p = nullptr;
  DMRGX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  DMRGX_CHECK(p != nullptr, "no-allocation scratch: no device address");
  g_scratch[device] = (unsigned char*)p;
}

}  // namespace dmrgx

using namespace dmrgx;

extern "C" int dmrgx_task_gemm_noalloc(int dtype, const void* A, const int64_t* blk_off, const int64_t* blk_size,
                                       int n_blk, const int32_t* map, int n_task, const void* in, void* out,
                                       int64_t n_out) {
  try {
    DMRGX_CHECK(dtype == DT_R64 || dtype == DT_C64, "task_gemm_noalloc: dtype must be R64 or C64");
    DMRGX_CHECK(rt().ready, "task_gemm_noalloc: call dmrgx_init first (initialisation allocates)");
    auto t = tasks_from_map(map, n_task, blk_off, blk_size, n_blk);
    for (auto& g : t)
      DMRGX_CHECK(g.c_off + (int64_t)g.M * g.N <= n_out, "task_gemm_noalloc: output past n_out");
    unsigned char* base = nullptr;
    {
      std::lock_guard<std::mutex> lk(g_mu);
      base = g_scratch[rt().device];
    }
    DMRGX_CHECK(base != nullptr, "task_gemm_noalloc: scratch not set up on this device");
    hipStream_t s = rt().stream;
    // the default decomposition (as dmrgx_task_gemm), or fewer K splits
    // until tables and partial tiles fit; one work item per output tile
    // needs no partial tiles at all
    GemmTables tab;
    std::unique_ptr<GemmPlan> plan;
    size_t o_segs = 0, o_red = 0, o_start = 0, o_ws = 0, total = 0;
    for (int target = 0;; target = target == 0 ? rt().num_cu : target / 2) {
      tab = GemmTables();
      plan = make_gemm_plan(t, dtype, PREC_FP64, 'N', 'T', {}, target, false, &tab);
      o_segs = align256(tab.items.size() * sizeof(GemmItem));
      o_red = o_segs + align256(tab.segs.size() * sizeof(GemmSeg));
      o_start = o_red + align256(tab.red.size() * sizeof(RedTask));
      o_ws = o_start + align256(tab.red_start.size() * sizeof(int64_t));
      total = o_ws + gemm_ws_bytes(*plan);
      if (total <= SCRATCH_BYTES || target == 1) break;
    }
    DMRGX_CHECK(total <= SCRATCH_BYTES, "task_gemm_noalloc: the task list needs " + std::to_string(total) +
                                            " bytes of scratch, above the 64 MiB maximum");
    // nothing in the scratch needs to start at zero: every table is written
    // here and every partial tile is written by its work item before the
    // reduction reads it
    if (!tab.items.empty()) upload(base, tab.items.data(), tab.items.size() * sizeof(GemmItem), s);
    if (!tab.segs.empty()) upload(base + o_segs, tab.segs.data(), tab.segs.size() * sizeof(GemmSeg), s);
    if (!tab.red.empty()) {
      upload(base + o_red, tab.red.data(), tab.red.size() * sizeof(RedTask), s);
      upload(base + o_start, tab.red_start.data(), tab.red_start.size() * sizeof(int64_t), s);
    }
    plan->items = (const GemmItem*)base;
    plan->segs = (const GemmSeg*)(base + o_segs);
    plan->red = (const RedTask*)(base + o_red);
    plan->red_start = (const int64_t*)(base + o_start);
    plan->ws = base + o_ws;
    // outputs without any task are left untouched, as the reference's loop does
    run_gemm_plan(*plan, A, in, out, s);
    DMRGX_HIP(hipStreamSynchronize(s));
    return 0;
  } catch (const dmrgx::Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return -2;
  }
}
