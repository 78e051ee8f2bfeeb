// H_eff plans and the C entry points of the task GEMM, the matvec and the
// eigensolver.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <map>
#include <mutex>
#include <string>

#include "dmrgx.h"
#include "common.hpp"
#include "gemm.hpp"
#include "heff.hpp"
#include "lanczos.hpp"
#include "tensor_ops.hpp"

namespace dmrgx {

std::vector<GemmTask> tasks_from_map(const int32_t* map, int n_task, const int64_t* blk_off,
                                     const int64_t* blk_size, int n_blk) {
  DMRGX_CHECK(n_task >= 0 && (n_task == 0 || map != nullptr), "task list: null map");
  std::vector<GemmTask> t;
  t.reserve(n_task);
  for (int i = 0; i < n_task; ++i) {
    const int32_t* m = map + 7 * i;
    GemmTask g{};
    int col = m[4], row = m[3];
    DMRGX_CHECK(m[0] >= 0 && m[0] < n_blk, "task block index out of range");
    DMRGX_CHECK(m[1] >= 0 && m[2] >= 0, "task with a negative input or output offset");
    DMRGX_CHECK(blk_off[m[0]] >= 0 && blk_size[m[0]] >= 0, "task block with a negative offset or size");
    DMRGX_CHECK(col > 0 && row > 0, "task with empty row/col");
    DMRGX_CHECK(blk_size[m[0]] % col == 0, "task col does not divide the block size");
    g.a_off = blk_off[m[0]];
    g.b_off = m[1];
    g.c_off = m[2];
    g.M = (int)(blk_size[m[0]] / col);
    g.N = row;
    g.K = col;
    g.beta = m[5];
    g.alpha = m[6];
    t.push_back(g);
  }
  return t;
}

std::vector<std::pair<int64_t, int64_t>> uncovered(const std::vector<GemmTask>& t, int64_t n_out) {
  std::vector<std::pair<int64_t, int64_t>> iv;
  for (auto& g : t) iv.push_back({g.c_off, g.c_off + (int64_t)g.M * g.N});
  std::sort(iv.begin(), iv.end());
  std::vector<std::pair<int64_t, int64_t>> gaps;
  int64_t pos = 0;
  for (auto& r : iv) {
    if (r.first > pos) gaps.push_back({pos, r.first - pos});
    pos = std::max(pos, r.second);
  }
  if (pos < n_out) gaps.push_back({pos, n_out - pos});
  return gaps;
}

static std::atomic<bool> g_autotune{true};
static std::atomic<bool> g_skip_zeros{true};

void set_heff_skip_zeros(bool on) { g_skip_zeros.store(on); }

namespace {

// A block scanned for zero rows or columns: rows x cols elements row-major at
// `off`; its flags start at flag_off, its scan units (rows or columns) at
// `first` in the launch's numbering.
struct NzBlk {
  int64_t off, flag_off, first;
  int rows, cols;
};

__device__ __forceinline__ int find_blk(const NzBlk* __restrict__ b, int nb, int64_t u) {
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (b[mid].first <= u) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// flags[row] = 1 when the row holds a nonzero element: one wavefront per row
__global__ __launch_bounds__(256) void k_nonzero_rows(const double* __restrict__ A, const NzBlk* __restrict__ b,
                                                      int nb, int64_t units, int cplx, int* __restrict__ flags) {
  const int64_t u = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (u >= units) return;  // uniform over the wavefront
  const int lane = threadIdx.x & 63;
  const NzBlk q = b[find_blk(b, nb, u)];
  const int64_t row = u - q.first;
  const int64_t len = (int64_t)q.cols << cplx;
  const double* p = A + ((q.off + row * q.cols) << cplx);
  bool nz = false;
  for (int64_t i = lane; i < len; i += 64) nz = nz || p[i] != 0.0;
  const unsigned long long any = __ballot(nz);
  if (lane == 0) flags[q.flag_off + row] = any != 0ull;
}

// flags[col] = 1 when the column holds a nonzero element: one thread per column
__global__ __launch_bounds__(256) void k_nonzero_cols(const double* __restrict__ A, const NzBlk* __restrict__ b,
                                                      int nb, int64_t units, int cplx, int* __restrict__ flags) {
  const int64_t u = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (u >= units) return;
  const NzBlk q = b[find_blk(b, nb, u)];
  const int64_t col = u - q.first;
  bool nz = false;
  for (int r = 0; r < q.rows && !nz; ++r) {
    const double* p = A + ((q.off + (int64_t)r * q.cols + col) << cplx);
    nz = p[0] != 0.0 || (cplx && p[1] != 0.0);
  }
  flags[q.flag_off + col] = nz;
}

// Zero rows (rows = true) or columns of the distinct A blocks of a task list
// (keyed by a_off): one flag per row / column, in the order of `blocks`.
std::vector<int> scan_nonzero(const void* A, const std::vector<GemmTask>& blocks, bool rows, bool cplx,
                              hipStream_t s) {
  std::vector<NzBlk> nb;
  int64_t nflags = 0;
  for (auto& g : blocks) {
    NzBlk q{g.a_off, nflags, nflags, g.M, g.K};
    nflags += rows ? g.M : g.K;
    nb.push_back(q);
  }
  std::vector<int> out(nflags, 1);
  if (nflags == 0) return out;
  DevBuf db(nb.size() * sizeof(NzBlk)), df(nflags * sizeof(int));
  upload(db.ptr, nb.data(), nb.size() * sizeof(NzBlk), s);
  if (rows)
    hipLaunchKernelGGL(k_nonzero_rows, dim3((unsigned)((nflags + 3) / 4)), dim3(256), 0, s, (const double*)A,
                       (const NzBlk*)db.ptr, (int)nb.size(), nflags, cplx ? 1 : 0, (int*)df.ptr);
  else
    hipLaunchKernelGGL(k_nonzero_cols, dim3((unsigned)((nflags + 255) / 256)), dim3(256), 0, s, (const double*)A,
                       (const NzBlk*)db.ptr, (int)nb.size(), nflags, cplx ? 1 : 0, (int*)df.ptr);
  DMRGX_HIP(hipGetLastError());
  DMRGX_HIP(hipMemcpyAsync(out.data(), df.ptr, nflags * sizeof(int), hipMemcpyDeviceToHost, s));
  DMRGX_HIP(hipStreamSynchronize(s));
  return out;
}

// The executed task lists of a two-site H_eff without the environments'
// structurally zero slices (exact zeros: the MPO bond index w of an
// environment block carries only the operators its sectors allow).
//   R pass (store = renv x psi, renv block rows (s, b, w)): the rows of an
//   output block are grouped by the set of tasks in which they are nonzero;
//   each group becomes row-mapped tasks over those rows and tasks only, and
//   rows that are zero in every task are not computed (their store elements
//   are zero and stay so: the plan zero-fills its store once).
//   L pass (out = lenv x store, lenv block columns (w, a', s')): each task
//   contracts only the K columns that are nonzero in lenv and computed in
//   store, as contiguous K ranges (8-aligned for the bf16 plane kernels).
// The sums keep their order over the remaining terms.
void skip_zero_slices(HeffPlan& h, std::vector<GemmTask>& rx, std::vector<GemmTask>& lx, bool align8,
                      hipStream_t s) {
  const bool cx = h.dtype == DT_C64;
  const std::vector<GemmTask> rt = rx, lt = lx;
  // distinct A blocks of each pass
  auto distinct = [](const std::vector<GemmTask>& t) {
    std::vector<GemmTask> b;
    std::map<int64_t, int> idx;
    std::vector<int> of(t.size());
    for (size_t i = 0; i < t.size(); ++i) {
      auto f = idx.find(t[i].a_off);
      if (f == idx.end()) {
        idx[t[i].a_off] = (int)b.size();
        b.push_back(t[i]);
        of[i] = (int)b.size() - 1;
      } else {
        DMRGX_CHECK(b[f->second].M == t[i].M && b[f->second].K == t[i].K, "skip: one block with two shapes");
        of[i] = f->second;
      }
    }
    return std::make_pair(b, of);
  };
  auto [rb, rof] = distinct(rt);
  auto [lb, lof] = distinct(lt);
  const std::vector<int> rflag = scan_nonzero(h.renv, rb, true, cx, s);
  const std::vector<int> lflag = scan_nonzero(h.lenv, lb, false, cx, s);
  std::vector<int64_t> rbase(rb.size()), lbase(lb.size());
  for (size_t i = 0, a = 0; i < rb.size(); a += rb[i].M, ++i) rbase[i] = (int64_t)a;
  for (size_t i = 0, a = 0; i < lb.size(); a += lb[i].K, ++i) lbase[i] = (int64_t)a;

  // ---- R pass: output blocks (first-appearance order) and their rows
  struct OutBlk { int64_t c_off; int M, N; std::vector<int> t; std::vector<char> comp; };
  std::vector<OutBlk> ob;
  std::map<int64_t, int> oidx;
  for (int i = 0; i < (int)rt.size(); ++i) {
    auto f = oidx.find(rt[i].c_off);
    if (f == oidx.end()) {
      oidx[rt[i].c_off] = (int)ob.size();
      ob.push_back(OutBlk{rt[i].c_off, rt[i].M, rt[i].N, {}, {}});
      f = oidx.find(rt[i].c_off);
    }
    DMRGX_CHECK(rt[i].M == ob[f->second].M && rt[i].N == ob[f->second].N,
                "skip: tasks into one output block disagree on shape");
    ob[f->second].t.push_back(i);
  }
  rx.clear();
  h.rexec_src.clear();
  h.row_maps.clear();
  for (auto& o : ob) {
    const int nt = (int)o.t.size();
    // row signature: which of the block's tasks are nonzero on the row
    std::map<std::vector<char>, std::vector<int32_t>> sig;
    std::vector<std::vector<char>> order;
    o.comp.assign(o.M, 0);
    for (int r = 0; r < o.M; ++r) {
      std::vector<char> key(nt);
      bool any = false;
      for (int j = 0; j < nt; ++j) {
        const int i = o.t[j];
        key[j] = (char)rflag[rbase[rof[i]] + r];
        any = any || key[j];
      }
      if (!any) continue;
      o.comp[r] = 1;
      auto f = sig.find(key);
      if (f == sig.end()) {
        order.push_back(key);
        sig[key].push_back(r);
      } else {
        f->second.push_back(r);
      }
    }
    const int beta0 = rt[o.t[0]].beta;
    const bool dense = sig.size() == 1 && (int)sig.begin()->second.size() == o.M &&
                       std::all_of(order[0].begin(), order[0].end(), [](char c) { return c != 0; });
    // every row zero: nothing to compute (the store block stays zero)
    if (sig.empty()) continue;
    // dense, or many small row groups (which would cost more in tile
    // efficiency than they save): the block's tasks as they are
    if (dense || sig.size() > 8) {
      for (int j = 0; j < nt; ++j) {
        rx.push_back(rt[o.t[j]]);
        h.rexec_src.push_back(o.t[j]);
      }
      std::fill(o.comp.begin(), o.comp.end(), 1);
      continue;
    }
    for (auto& key : order) {
      const std::vector<int32_t>& rows = sig[key];
      const int64_t off = (int64_t)h.row_maps.size();
      h.row_maps.insert(h.row_maps.end(), rows.begin(), rows.end());
      bool first = true;
      for (int j = 0; j < nt; ++j) {
        if (!key[j]) continue;
        GemmTask g = rt[o.t[j]];
        g.M = (int)rows.size();
        g.row_map = off;
        g.beta = first ? beta0 : 1;
        first = false;
        rx.push_back(g);
        h.rexec_src.push_back(o.t[j]);
      }
    }
  }

  // ---- L pass: per task, the K columns worth contracting
  lx.clear();
  h.lexec_src.clear();
  std::map<int64_t, int> lfirst;  // output block -> index in lx of its first executed task
  for (int i = 0; i < (int)lt.size(); ++i) {
    const GemmTask& t = lt[i];
    const int K = t.K;
    std::vector<char> col(K);
    const int64_t lb0 = lbase[lof[i]];
    for (int k = 0; k < K; ++k) col[k] = (char)lflag[lb0 + k];
    // store columns: the B block (N x K at b_off) inside one R output block
    // whose rows hold N_R = an exact divisor of K elements each
    auto ob_it = std::find_if(ob.begin(), ob.end(), [&](const OutBlk& o) {
      return o.c_off <= t.b_off && t.b_off + (int64_t)t.N * K <= o.c_off + (int64_t)o.M * o.N;
    });
    if (ob_it != ob.end() && K % ob_it->N == 0 && (t.b_off - ob_it->c_off) % ob_it->N == 0) {
      const int q = K / ob_it->N;
      const int64_t r0 = (t.b_off - ob_it->c_off) / ob_it->N;
      for (int c = 0; c < q; ++c) {
        bool any = false;
        for (int n = 0; n < t.N && !any; ++n) any = ob_it->comp[r0 + (int64_t)n * q + c] != 0;
        if (!any)
          for (int k = c * ob_it->N; k < (c + 1) * ob_it->N; ++k) col[k] = 0;
      }
    }
    // contiguous runs of columns (8-aligned for the plane kernels); gaps
    // shorter than 16 columns are contracted rather than cut
    std::vector<std::pair<int, int>> runs;
    for (int k = 0; k < K;) {
      if (!col[k]) { ++k; continue; }
      int e = k;
      while (e < K && col[e]) ++e;
      int b0 = k, e0 = e;
      if (align8) {
        b0 = b0 / 8 * 8;
        e0 = std::min(K, (e0 + 7) / 8 * 8);
      }
      if (!runs.empty() && b0 - runs.back().second < 16) runs.back().second = std::max(runs.back().second, e0);
      else runs.push_back({b0, e0});
      k = e;
    }
    for (auto& r : runs) {
      GemmTask g = t;
      g.k0 = r.first;
      g.k1 = r.second == K && r.first == 0 ? 0 : r.second;
      auto f = lfirst.find(t.c_off);
      if (f == lfirst.end()) {
        lfirst[t.c_off] = (int)lx.size();
        // the block's first write overwrites, whichever task it comes from
        int b0 = 0;
        for (auto& u : lt)
          if (u.c_off == t.c_off) { b0 = u.beta; break; }
        g.beta = b0;
      } else {
        g.beta = 1;
      }
      lx.push_back(g);
      h.lexec_src.push_back(i);
    }
  }
}

}  // namespace

static std::unique_ptr<DevBuf> to_f32(const void* src, int64_t nreal, hipStream_t s) {
  auto b = std::make_unique<DevBuf>(std::max<int64_t>(nreal, 1) * 4);
  convert_f64_to_f32((const double*)src, (float*)b->ptr, nreal, s);
  return b;
}

std::unique_ptr<HeffPlan> make_heff(int dtype, int prec, const void* renv, const int64_t* rblk_off,
                                    const int64_t* rblk_size, int n_rblk, int64_t renv_len,
                                    const int32_t* rmap, int n_rtask, const void* lenv,
                                    const int64_t* lblk_off, const int64_t* lblk_size, int n_lblk,
                                    int64_t lenv_len, const int32_t* lmap, int n_ltask, int64_t n_psi,
                                    int64_t n_store) {
  // every argument is validated before the device is touched
  DMRGX_CHECK(dtype == DT_R64 || dtype == DT_C64, "heff: dtype must be R64 or C64");
  DMRGX_CHECK(prec >= PREC_FP64 && prec <= PREC_BF16X3, "heff: unknown precision");
  DMRGX_CHECK(n_psi >= 0 && n_store >= 0, "heff: negative psi or store length");
  for (int i = 0; i < n_rtask; ++i)
    DMRGX_CHECK(rmap[7 * i] >= 0 && rmap[7 * i] < n_rblk, "heff: rmap block index out of range");
  for (int i = 0; i < n_ltask; ++i)
    DMRGX_CHECK(lmap[7 * i] >= 0 && lmap[7 * i] < n_lblk, "heff: lmap block index out of range");
  auto rt_ = tasks_from_map(rmap, n_rtask, rblk_off, rblk_size, n_rblk);
  auto lt_ = tasks_from_map(lmap, n_ltask, lblk_off, lblk_size, n_lblk);
  for (auto& g : rt_) {
    DMRGX_CHECK(g.b_off + (int64_t)g.N * g.K <= n_psi, "heff: rmap reads past the end of psi");
    DMRGX_CHECK(g.c_off + (int64_t)g.M * g.N <= n_store, "heff: rmap writes past the end of store");
    DMRGX_CHECK(g.a_off + (int64_t)g.M * g.K <= renv_len, "heff: rmap block past the end of renv");
  }
  for (auto& g : lt_) {
    DMRGX_CHECK(g.b_off + (int64_t)g.N * g.K <= n_store, "heff: lmap reads past the end of store");
    DMRGX_CHECK(g.c_off + (int64_t)g.M * g.N <= n_psi, "heff: lmap writes past the end of psi");
    DMRGX_CHECK(g.a_off + (int64_t)g.M * g.K <= lenv_len, "heff: lmap block past the end of lenv");
  }
  ensure_init();
  auto h = std::make_unique<HeffPlan>();
  h->dtype = dtype;
  h->prec = prec;
  h->n_psi = n_psi;
  h->n_store = n_store;
  h->renv = renv;
  h->lenv = lenv;
  const bool planes = prec == PREC_BF16X3;
  h->rtasks = rt_;
  h->ltasks = lt_;
  h->renv_len = renv_len;
  h->lenv_len = lenv_len;
  const double cf = dtype == DT_C64 ? 8.0 : 2.0;
  for (auto* v : {&rt_, &lt_})
    for (auto& g : *v) h->flops += cf * (double)g.M * g.N * g.K;
  // store ranges no task writes read as zero at every apply; with the zero
  // slices skipped, rows that no executed task writes are zero-filled once
  const auto rzero = uncovered(rt_, n_store);
  std::vector<GemmTask> rx = rt_, lx = lt_;
  h->rexec_src.resize(rx.size());
  h->lexec_src.resize(lx.size());
  for (size_t i = 0; i < rx.size(); ++i) h->rexec_src[i] = (int)i;
  for (size_t i = 0; i < lx.size(); ++i) h->lexec_src[i] = (int)i;
  hipStream_t s = rt().stream;
  if (g_skip_zeros.load() && renv && lenv) skip_zero_slices(*h, rx, lx, planes, s);
  const auto lzero = uncovered(lx, n_psi);
  if (planes) {
    // every operand block gets rows of a multiple of 8 bf16 elements
    auto lay = [](const std::vector<GemmTask>& t, bool a) {
      std::vector<PadBlk> b;
      for (auto& g : t) b.push_back(PadBlk{a ? g.a_off : g.b_off, 0, a ? g.M : g.N, g.K, 0, 0});
      return make_pad_layout(std::move(b));
    };
    h->renv_lay = lay(rt_, true);
    h->psi_lay = lay(rt_, false);
    h->lenv_lay = lay(lt_, true);
    h->store_lay = lay(lt_, false);
    auto remap = [](std::vector<GemmTask>& t, const PadLayout& A, const PadLayout& B) {
      for (auto& g : t) {
        g.lda = A.kp(g.a_off);
        g.ldb = B.kp(g.b_off);
        g.a_off = A.map(g.a_off);
        g.b_off = B.map(g.b_off);
      }
    };
    remap(rx, *h->renv_lay, *h->psi_lay);
    remap(lx, *h->lenv_lay, *h->store_lay);
  }
  h->rexec = rx;
  h->lexec = lx;
  h->rzero = rzero;
  h->lzero = lzero;
  h->rplan = make_gemm_plan(rx, dtype, prec, 'N', 'T', rzero, 0, planes, nullptr, h->maps(),
                            planes ? h->store_lay.get() : nullptr);
  h->lplan = make_gemm_plan(lx, dtype, prec, 'N', 'T', lzero, 0, planes, nullptr, h->maps());
  h->flops_exec = h->rplan->flops + h->lplan->flops;
  const int c = dtype == DT_C64 ? 2 : 1;
  if (prec == PREC_FP64) {
    h->store = std::make_unique<DevBuf>(std::max<int64_t>(n_store, 1) * 8 * c);
  } else if (planes) {
    const int npl = 2 * c;
    // the R pass writes `store` as planes when no output accumulates and its
    // blocks map into the L pass's padded layout (row by row, zeroed once
    // below); otherwise it writes FP32 and the planes are split from it
    h->store_direct = !h->rplan->any_beta &&
                      (h->store_lay->identity ? h->store_lay->total == n_store : h->rplan->out_padded);
    h->renv_ps = plane_stride(h->renv_lay->total);
    h->lenv_ps = plane_stride(h->lenv_lay->total);
    h->psi_ps = plane_stride(h->psi_lay->total);
    h->store_ps = plane_stride(h->store_lay->total);  // n_store when the layout is the plain one
    h->renv_pl = std::make_unique<DevBuf>(std::max<int64_t>(h->renv_ps, 8) * npl * 2);
    h->lenv_pl = std::make_unique<DevBuf>(std::max<int64_t>(h->lenv_ps, 8) * npl * 2);
    h->in_pl = std::make_unique<DevBuf>(std::max<int64_t>(h->psi_ps, 8) * npl * 2);
    h->store_pl = std::make_unique<DevBuf>(std::max<int64_t>(h->store_ps, 8) * npl * 2);
    split_padded_f64((const double*)renv, (uint16_t*)h->renv_pl->ptr, h->renv_ps, *h->renv_lay, c == 2, s);
    split_padded_f64((const double*)lenv, (uint16_t*)h->lenv_pl->ptr, h->lenv_ps, *h->lenv_lay, c == 2, s);
    if (!h->store_direct) h->store = std::make_unique<DevBuf>(std::max<int64_t>(n_store, 1) * 4 * c);
  } else {
    h->renv_lo = to_f32(renv, renv_len * c, s);
    h->lenv_lo = to_f32(lenv, lenv_len * c, s);
    h->store = std::make_unique<DevBuf>(std::max<int64_t>(n_store, 1) * 4 * c);
    h->in_lo = std::make_unique<DevBuf>(std::max<int64_t>(n_psi, 1) * 4 * c);
    h->out_lo = std::make_unique<DevBuf>(std::max<int64_t>(n_psi, 1) * 4 * c);
  }
  // rows of store that no executed task writes hold zeros from here on
  if (h->store) fill_zero(h->store->ptr, h->store->bytes, s);
  if (h->store_pl) fill_zero(h->store_pl->ptr, h->store_pl->bytes, s);
  DMRGX_HIP(hipStreamSynchronize(s));
  if (g_autotune.load()) h->autotune(s);
  return h;
}

// the padded layout the R pass writes its plane output into, if any
const PadLayout* HeffPlan::r_out_layout() const {
  return prec == PREC_BF16X3 && store_direct && store_lay && !store_lay->identity ? store_lay.get() : nullptr;
}

void HeffPlan::replan(int r_target, int l_target) {
  const bool planes = prec == PREC_BF16X3;
  rplan = make_gemm_plan(rexec, dtype, prec, 'N', 'T', rzero, r_target, planes, nullptr, maps(), r_out_layout());
  DMRGX_CHECK(!r_out_layout() || rplan->out_padded, "replan: the R pass lost its padded output layout");
  lplan = make_gemm_plan(lexec, dtype, prec, 'N', 'T', lzero, l_target, planes, nullptr, maps());
  r_split = r_target;
  l_split = l_target;
}

void HeffPlan::restrict_tasks(const std::vector<char>& keep_r, const std::vector<char>& keep_l) {
  if (keep_r.empty() && keep_l.empty()) {
    replan(r_split, l_split);
    return;
  }
  DMRGX_CHECK(keep_r.size() == rtasks.size() && keep_l.size() == ltasks.size(), "restrict_tasks: mask sizes");
  // the masks select original tasks; the executed tasks follow their origin
  std::vector<GemmTask> r, l, ro;
  for (size_t i = 0; i < rexec.size(); ++i)
    if (keep_r[rexec_src[i]]) r.push_back(rexec[i]);
  for (size_t i = 0; i < lexec.size(); ++i)
    if (keep_l[lexec_src[i]]) l.push_back(lexec[i]);
  for (size_t i = 0; i < rtasks.size(); ++i)
    if (keep_r[i]) ro.push_back(rtasks[i]);
  const bool planes = prec == PREC_BF16X3;
  // dropped outputs read as zero: the ranges that no kept original R task
  // writes (rows skipped inside kept blocks stay zero) and that no kept
  // executed L task writes
  rplan = make_gemm_plan(r, dtype, prec, 'N', 'T', uncovered(ro, n_store), r_split, planes, nullptr, maps(),
                         r_out_layout());
  DMRGX_CHECK(!r_out_layout() || rplan->out_padded || rplan->n_items == 0,
              "restrict_tasks: the R pass lost its padded output layout");
  lplan = make_gemm_plan(l, dtype, prec, 'N', 'T', uncovered(l, n_psi), l_split, planes, nullptr, maps());
}

namespace {

// The tuner's key: precision, element type, and per pass the histogram of
// task shapes with each of M, N, K rounded to a power of two (so that the
// slightly different sector sizes of consecutive DMRG steps share a key).
std::string shape_key(const HeffPlan& h) {
  auto hist = [](const std::vector<GemmTask>& t) {
    std::map<std::array<int, 3>, int> m;
    auto lg = [](int x) { int b = 0; while ((1 << b) < x) ++b; return b; };
    for (auto& g : t) m[{lg(g.M), lg(g.N), lg(g.K)}]++;
    std::string k;
    for (auto& e : m) {
      int c = 0;
      while ((1 << c) < e.second) ++c;  // counts rounded to powers of two too
      k += std::to_string(e.first[0]) + "," + std::to_string(e.first[1]) + "," + std::to_string(e.first[2]) + "x" +
           std::to_string(c) + ";";
    }
    return k;
  };
  return std::to_string(h.dtype) + "/" + std::to_string(h.prec) + "/R" + hist(h.rexec) + "/L" + hist(h.lexec);
}

std::mutex g_tune_mu;
std::map<std::string, std::pair<int, int>> g_tune;  // key -> (R target, L target)

}  // namespace

void set_heff_autotune(bool on) { g_autotune.store(on); }
int heff_autotune_keys() {
  std::lock_guard<std::mutex> lk(g_tune_mu);
  return (int)g_tune.size();
}

void HeffPlan::autotune(hipStream_t s) {
  // small operators keep the default: timing them costs more than any choice saves
  if (rexec.empty() || lexec.empty() || n_psi <= 0 || flops < 1e9) return;
  const std::string key = shape_key(*this);
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto f = g_tune.find(key);
    if (f != g_tune.end()) {
      if (f->second.first != r_split || f->second.second != l_split) replan(f->second.first, f->second.second);
      return;
    }
  }
  // candidates: 2 (the default, first), 1, 4 and 8 work items per CU, and for
  // BF16x3 the balanced decomposition over the resident work-groups (one
  // 160 KB-LDS work-group per CU for complex operands, two for real).  The
  // whole matvec (both passes with their split-K reductions and zero fills)
  // is timed with events, two applies after one warm-up on a zero input:
  // the R pass's candidates first with the L pass at the default, then the L
  // pass's with the R pass at its best.  Another candidate replaces the
  // default only when the matvec is at least 3 % faster.
  const int cu = std::max(rt().num_cu, 1);
  std::vector<int> cand = {2 * cu, cu, 4 * cu, 8 * cu};
  if (prec == PREC_BF16X3) cand.push_back(-cu * (dtype == DT_C64 ? 1 : 2));
  const int nc = (int)cand.size();
  const int c = dtype == DT_C64 ? 2 : 1;
  DevBuf x(std::max<int64_t>(n_psi, 1) * 8 * c), y(std::max<int64_t>(n_psi, 1) * 8 * c);
  fill_zero(x.ptr, n_psi * 8 * c, s);
  hipEvent_t ev[2];
  for (auto& e : ev) DMRGX_HIP(hipEventCreate(&e));
  // the matvec with the current pass plans: one warm-up, then two timed applies
  auto time_now = [&]() {
    apply((const double*)x.ptr, (double*)y.ptr, s);
    DMRGX_HIP(hipEventRecord(ev[0], s));
    for (int rep = 0; rep < 2; ++rep) apply((const double*)x.ptr, (double*)y.ptr, s);
    DMRGX_HIP(hipEventRecord(ev[1], s));
    DMRGX_HIP(hipEventSynchronize(ev[1]));
    float ms = 0;
    DMRGX_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
    return (double)ms;
  };
  // the fastest candidate that beats the default by the margin, else the
  // default; a candidate whose plan repeats one already timed (the same work
  // items, segments and reductions) reuses that time
  auto pick = [&](bool r_pass, int other) {
    std::vector<std::pair<std::array<int64_t, 4>, double>> seen;
    double t0 = 0, bt = 0;
    int b = cand[0];
    for (int i = 0; i < nc; ++i) {
      if (r_pass) replan(cand[i], other);
      else replan(other, cand[i]);
      const GemmPlan& g = r_pass ? *rplan : *lplan;
      const std::array<int64_t, 4> sig = {g.n_items, g.n_segs, g.n_red, g.ws_elems};
      double t = -1;
      for (auto& e : seen)
        if (e.first == sig) t = e.second;
      if (t < 0) {
        t = time_now();
        seen.push_back({sig, t});
      }
      if (i == 0) {
        t0 = bt = t;
      } else if (t < 0.97 * t0 && t < bt) {
        bt = t;
        b = cand[i];
      }
    }
    return b;
  };
  const int br = pick(true, cand[0]);
  const int bl = pick(false, br);
  for (auto& e : ev) (void)hipEventDestroy(e);
  replan(br, bl);
  DMRGX_HIP(hipStreamSynchronize(s));
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune[key] = {br, bl};
}

namespace {

// gA = g conj(B) (M x K): C' = A'(M x N) B'(N x K), 'N','N'
std::vector<GemmTask> adj_tasks_a(const std::vector<GemmTask>& t) {
  std::vector<GemmTask> o;
  for (auto& g : t) {
    GemmTask a{};
    a.c_off = g.a_off;
    a.a_off = g.c_off;
    a.b_off = g.b_off;
    a.M = g.M;
    a.N = g.K;
    a.K = g.N;
    a.alpha = g.alpha;
    a.beta = 0;
    o.push_back(a);
  }
  return o;
}

// gB = g^T conj(A) (N x K): C' = A'^T B' with A' = g stored M x N ('T') and B' = conj(A) stored M x K
std::vector<GemmTask> adj_tasks_b(const std::vector<GemmTask>& t) {
  std::vector<GemmTask> o;
  for (auto& g : t) {
    GemmTask b{};
    b.c_off = g.b_off;
    b.a_off = g.c_off;
    b.b_off = g.a_off;
    b.M = g.N;
    b.N = g.K;
    b.K = g.M;
    b.alpha = g.alpha;
    b.beta = 0;
    o.push_back(b);
  }
  return o;
}

}  // namespace

void HeffPlan::adjoint(const double* psi, const double* g_out, double* g_psi, double* g_renv, double* g_lenv,
                       hipStream_t s) {
  DMRGX_CHECK(!edge, "adjoint: operators with periodic edge terms have no adjoint kernels");
  const bool cx = dtype == DT_C64;
  const int c = cx ? 2 : 1;
  const int aprec = prec == PREC_FP64 ? PREC_FP64 : PREC_FP32;
  const size_t eb = (aprec == PREC_FP64 ? 8 : 4) * c;
  if (!ad_gr) {
    auto ta = adj_tasks_a(rtasks), tb = adj_tasks_b(rtasks), la = adj_tasks_a(ltasks), lb = adj_tasks_b(ltasks);
    ad_gr = make_gemm_plan(ta, dtype, aprec, 'N', 'N', uncovered(ta, renv_len));
    ad_gp = make_gemm_plan(tb, dtype, aprec, 'T', 'N', uncovered(tb, n_psi));
    ad_gl = make_gemm_plan(la, dtype, aprec, 'N', 'N', uncovered(la, lenv_len));
    ad_gs = make_gemm_plan(lb, dtype, aprec, 'T', 'N', uncovered(lb, n_store));
    // store is recomputed over the original tasks (every row written)
    ad_fr = make_gemm_plan(rtasks, dtype, aprec, 'N', 'T', uncovered(rtasks, n_store));
  }
  // operands in the adjoint precision
  auto as_prec = [&](const double* x, int64_t n, std::unique_ptr<DevBuf>& keep) -> const void* {
    if (aprec == PREC_FP64) return x;
    keep = std::make_unique<DevBuf>(std::max<int64_t>(n, 1) * eb);
    convert_f64_to_f32(x, (float*)keep->ptr, n * c, s);
    return keep->ptr;
  };
  std::unique_ptr<DevBuf> kp, kg, kr, kl;
  const void* P = as_prec(psi, n_psi, kp);
  const void* G = as_prec(g_out, n_psi, kg);
  const void* R = aprec == PREC_FP64 ? renv : (renv_lo ? renv_lo->ptr : as_prec((const double*)renv, renv_len, kr));
  const void* Lv = aprec == PREC_FP64 ? lenv : (lenv_lo ? lenv_lo->ptr : as_prec((const double*)lenv, lenv_len, kl));
  // store = R pass of psi
  DevBuf st(std::max<int64_t>(n_store, 1) * eb);
  run_gemm_plan(*ad_fr, R, P, st.ptr, s);
  // conjugated copies of the B operands (complex only)
  auto conj_copy = [&](const void* x, int64_t n) {
    auto b = std::make_unique<DevBuf>(std::max<int64_t>(n, 1) * eb);
    copy_d2d(b->ptr, x, n * eb, s);
    if (cx) {
      if (aprec == PREC_FP64) conj_inplace(b->ptr, n, s);
      else conj_inplace_f32(b->ptr, n, s);
    }
    return b;
  };
  DevBuf gst(std::max<int64_t>(n_store, 1) * eb);
  {
    auto cl = conj_copy(Lv, lenv_len);
    run_gemm_plan(*ad_gs, G, cl->ptr, gst.ptr, s);  // g_store = g^T conj(lenv)
  }
  auto out64 = [&](const DevBuf& src, double* dst, int64_t n) {
    if (aprec == PREC_FP64) copy_d2d(dst, src.ptr, n * 8 * c, s);
    else convert_f32_to_f64((const float*)src.ptr, dst, n * c, s);
  };
  if (g_lenv) {
    auto cs = conj_copy(st.ptr, n_store);
    DevBuf o(std::max<int64_t>(lenv_len, 1) * eb);
    run_gemm_plan(*ad_gl, G, cs->ptr, o.ptr, s);  // g_lenv = g conj(store)
    out64(o, g_lenv, lenv_len);
    DMRGX_HIP(hipStreamSynchronize(s));
  }
  if (g_psi) {
    auto cr = conj_copy(R, renv_len);
    DevBuf o(std::max<int64_t>(n_psi, 1) * eb);
    run_gemm_plan(*ad_gp, gst.ptr, cr->ptr, o.ptr, s);  // g_psi = g_store^T conj(renv)
    out64(o, g_psi, n_psi);
    DMRGX_HIP(hipStreamSynchronize(s));
  }
  if (g_renv) {
    auto cp = conj_copy(P, n_psi);
    DevBuf o(std::max<int64_t>(renv_len, 1) * eb);
    run_gemm_plan(*ad_gr, gst.ptr, cp->ptr, o.ptr, s);  // g_renv = g_store conj(psi)
    out64(o, g_renv, renv_len);
  }
  DMRGX_HIP(hipStreamSynchronize(s));
}

namespace {

// y += x over n doubles
__global__ __launch_bounds__(256) void k_add_into(double* __restrict__ y, const double* __restrict__ x, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] += x[i];
}

}  // namespace

void attach_edge(HeffPlan& h, std::unique_ptr<HeffPlan> e) {
  DMRGX_CHECK(e && e->n_psi == h.n_psi && e->dtype == h.dtype && e->prec == h.prec && !e->edge,
              "attach_edge: the edge operator does not match the bulk one");
  const int c = h.dtype == DT_C64 ? 2 : 1;
  h.edge_out = std::make_unique<DevBuf>(std::max<int64_t>(h.n_psi, 1) * 8 * c);
  h.flops += e->flops;
  h.flops_exec += e->flops_exec;
  h.edge = std::move(e);
}

void HeffPlan::apply(const double* in, double* out, hipStream_t s, hipEvent_t* ev, bool coarse) {
  const int c = dtype == DT_C64 ? 2 : 1;
  hipEvent_t e0 = ev ? ev[0] : nullptr, e1 = ev ? ev[1] : nullptr, e2 = ev ? ev[2] : nullptr,
             e3 = ev ? ev[3] : nullptr;
  if (prec == PREC_FP64) {
    run_gemm_plan(*rplan, renv, in, store->ptr, s, e0, e1);
    run_gemm_plan(*lplan, lenv, store->ptr, out, s, e2, e3);
  } else if (prec == PREC_BF16X3) {
    // the R pass writes `store` straight as bf16 planes and the L pass writes
    // the FP64 result (split-K partial tiles are converted by the reduction)
    split_padded_f64(in, (uint16_t*)in_pl->ptr, psi_ps, *psi_lay, c == 2, s);
    if (store_direct) {
      run_gemm_planes(*rplan, (const uint16_t*)renv_pl->ptr, renv_ps, (const uint16_t*)in_pl->ptr, psi_ps,
                      store_pl->ptr, OUT_PLANES, store_ps, s, e0, e1, coarse);
    } else {
      run_gemm_planes(*rplan, (const uint16_t*)renv_pl->ptr, renv_ps, (const uint16_t*)in_pl->ptr, psi_ps,
                      store->ptr, OUT_F32, 0, s, e0, e1, coarse);
      split_padded_f32((const float*)store->ptr, (uint16_t*)store_pl->ptr, store_ps, *store_lay, c == 2, s);
    }
    run_gemm_planes(*lplan, (const uint16_t*)lenv_pl->ptr, lenv_ps, (const uint16_t*)store_pl->ptr, store_ps, out,
                    OUT_F64, 0, s, e2, e3, coarse);
  } else {
    convert_f64_to_f32(in, (float*)in_lo->ptr, n_psi * c, s);
    run_gemm_plan(*rplan, renv_lo->ptr, in_lo->ptr, store->ptr, s, e0, e1);
    run_gemm_plan(*lplan, lenv_lo->ptr, store->ptr, out_lo->ptr, s, e2, e3);
    convert_f32_to_f64((const float*)out_lo->ptr, out, n_psi * c, s);
  }
  if (edge) {
    // out += L_edge W_id W_id R_edge psi (src/core.cpp:27-30)
    edge->apply(in, (double*)edge_out->ptr, s);
    const int64_t n = n_psi * c;
    if (n > 0) {
      hipLaunchKernelGGL(k_add_into, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, out,
                         (const double*)edge_out->ptr, n);
      DMRGX_HIP(hipGetLastError());
    }
  }
}

void HeffPlan::time_kernels(const double* in, double* out, int iters, hipStream_t s, double* ms_r, double* ms_l) {
  DMRGX_CHECK(iters > 0, "time_kernels: iters must be positive");
  apply(in, out, s);
  // each pass's kernel with the operands apply() uses
  auto r_pass = [&]() {
    if (prec == PREC_FP64) launch_plan_kernel(*rplan, renv, in, store->ptr, s);
    else if (prec == PREC_BF16X3)
      launch_planes_kernel(*rplan, (const uint16_t*)renv_pl->ptr, renv_ps, (const uint16_t*)in_pl->ptr, psi_ps,
                           store_direct ? store_pl->ptr : store->ptr, store_direct ? OUT_PLANES : OUT_F32,
                           store_direct ? store_ps : 0, s);
    else launch_plan_kernel(*rplan, renv_lo->ptr, in_lo->ptr, store->ptr, s);
  };
  auto l_pass = [&]() {
    if (prec == PREC_FP64) launch_plan_kernel(*lplan, lenv, store->ptr, out, s);
    else if (prec == PREC_BF16X3)
      launch_planes_kernel(*lplan, (const uint16_t*)lenv_pl->ptr, lenv_ps, (const uint16_t*)store_pl->ptr,
                           store_ps, out, OUT_F64, 0, s);
    else launch_plan_kernel(*lplan, lenv_lo->ptr, store->ptr, out_lo->ptr, s);
  };
  hipEvent_t ev[3];
  for (auto& e : ev) DMRGX_HIP(hipEventCreate(&e));
  DMRGX_HIP(hipEventRecord(ev[0], s));
  for (int i = 0; i < iters; ++i) r_pass();
  DMRGX_HIP(hipEventRecord(ev[1], s));
  for (int i = 0; i < iters; ++i) l_pass();
  DMRGX_HIP(hipEventRecord(ev[2], s));
  DMRGX_HIP(hipEventSynchronize(ev[2]));
  float a = 0, b = 0;
  DMRGX_HIP(hipEventElapsedTime(&a, ev[0], ev[1]));
  DMRGX_HIP(hipEventElapsedTime(&b, ev[1], ev[2]));
  for (auto& e : ev) (void)hipEventDestroy(e);
  *ms_r = a / iters;
  *ms_l = b / iters;
  // the L-pass kernel alone leaves split-K partials unreduced: restore out
  apply(in, out, s);
}

}  // namespace dmrgx

using namespace dmrgx;

#define CAPI_BEGIN try {
#define CAPI_END                      \
  return 0;                           \
  }                                   \
  catch (const dmrgx::Error& e) {     \
    set_last_error(e.what());         \
    return e.code;                    \
  }                                   \
  catch (const std::exception& e) {   \
    set_last_error(e.what());         \
    return -2;                        \
  }

static int64_t blocks_len(const int64_t* off, const int64_t* size, int n) {
  int64_t e = 0;
  for (int i = 0; i < n; ++i) e = std::max(e, off[i] + size[i]);
  return e;
}

extern "C" {

int dmrgx_task_gemm(int dtype, int prec, const void* A, const int64_t* blk_off, const int64_t* blk_size,
                    int n_blk, const int32_t* map, int n_task, const void* in, void* out, int64_t n_out) {
  CAPI_BEGIN
  DMRGX_CHECK(dtype == DT_R64 || dtype == DT_C64, "task_gemm: dtype must be R64 or C64");
  DMRGX_CHECK(prec >= PREC_FP64 && prec <= PREC_BF16X3, "task_gemm: unknown precision");
  auto t = tasks_from_map(map, n_task, blk_off, blk_size, n_blk);
  int64_t n_in = 0;
  for (auto& g : t) {
    n_in = std::max(n_in, g.b_off + (int64_t)g.N * g.K);
    DMRGX_CHECK(g.c_off + (int64_t)g.M * g.N <= n_out, "task_gemm: output past n_out");
  }
  ensure_init();
  hipStream_t s = rt().stream;
  int64_t alen = blocks_len(blk_off, blk_size, n_blk);
  // outputs without any task are left untouched, as the reference's loop does
  auto plan = make_gemm_plan(t, dtype, prec, 'N', 'T', {});
  const int c = dtype == DT_C64 ? 2 : 1;
  if (prec == PREC_FP64) {
    run_gemm_plan(*plan, A, in, out, s);
  } else {
    DevBuf a32(std::max<int64_t>(alen, 1) * 4 * c), i32(std::max<int64_t>(n_in, 1) * 4 * c),
        o32(std::max<int64_t>(n_out, 1) * 4 * c);
    convert_f64_to_f32((const double*)A, (float*)a32.ptr, alen * c, s);
    convert_f64_to_f32((const double*)in, (float*)i32.ptr, n_in * c, s);
    convert_f64_to_f32((const double*)out, (float*)o32.ptr, n_out * c, s);
    run_gemm_plan(*plan, a32.ptr, i32.ptr, o32.ptr, s);
    convert_f32_to_f64((const float*)o32.ptr, (double*)out, n_out * c, s);
    DMRGX_HIP(hipStreamSynchronize(s));
  }
  DMRGX_HIP(hipStreamSynchronize(s));
  CAPI_END
}

int dmrgx_heff_create(dmrgx_heff** plan, int dtype, int prec, const void* renv, const int64_t* rblk_off,
                      const int64_t* rblk_size, int n_rblk, const int32_t* rmap, int n_rtask,
                      const void* lenv, const int64_t* lblk_off, const int64_t* lblk_size, int n_lblk,
                      const int32_t* lmap, int n_ltask, int64_t n_psi, int64_t n_store) {
  CAPI_BEGIN
  DMRGX_CHECK(plan != nullptr, "heff_create: null plan pointer");
  auto h = new dmrgx_heff;
  try {
    h->p = make_heff(dtype, prec, renv, rblk_off, rblk_size, n_rblk, blocks_len(rblk_off, rblk_size, n_rblk),
                     rmap, n_rtask, lenv, lblk_off, lblk_size, n_lblk, blocks_len(lblk_off, lblk_size, n_lblk),
                     lmap, n_ltask, n_psi, n_store);
  } catch (...) {
    delete h;
    throw;
  }
  *plan = h;
  CAPI_END
}

int dmrgx_heff_apply(dmrgx_heff* plan, const void* in, void* out) {
  CAPI_BEGIN
  DMRGX_CHECK(plan && plan->p, "heff_apply: null plan");
  plan->p->apply((const double*)in, (double*)out, rt().stream);
  DMRGX_HIP(hipStreamSynchronize(rt().stream));
  CAPI_END
}

int dmrgx_heff_sizes(const dmrgx_heff* plan, int64_t* out) {
  CAPI_BEGIN
  DMRGX_CHECK(plan && plan->p && out, "heff_sizes: null arguments");
  out[0] = plan->p->n_psi;
  out[1] = plan->p->n_store;
  out[2] = plan->p->renv_len;
  out[3] = plan->p->lenv_len;
  CAPI_END
}

int dmrgx_heff_adjoint(dmrgx_heff* plan, const void* psi, const void* g_out, void* g_psi, void* g_renv,
                       void* g_lenv) {
  CAPI_BEGIN
  DMRGX_CHECK(plan && plan->p && psi && g_out, "heff_adjoint: null arguments");
  plan->p->adjoint((const double*)psi, (const double*)g_out, (double*)g_psi, (double*)g_renv, (double*)g_lenv,
                   rt().stream);
  DMRGX_HIP(hipStreamSynchronize(rt().stream));  // complete on return
  CAPI_END
}

int dmrgx_heff_time_adjoint(dmrgx_heff* plan, const void* psi, const void* g_out, void* g_psi, void* g_renv,
                            void* g_lenv, int iters, double* ms) {
  CAPI_BEGIN
  DMRGX_CHECK(plan && plan->p && psi && g_out && iters > 0 && ms, "heff_time_adjoint: bad arguments");
  hipStream_t s = rt().stream;
  hipEvent_t e0, e1;
  DMRGX_HIP(hipEventCreate(&e0));
  DMRGX_HIP(hipEventCreate(&e1));
  DMRGX_HIP(hipEventRecord(e0, s));
  for (int i = 0; i < iters; ++i)
    plan->p->adjoint((const double*)psi, (const double*)g_out, (double*)g_psi, (double*)g_renv, (double*)g_lenv, s);
  DMRGX_HIP(hipEventRecord(e1, s));
  DMRGX_HIP(hipStreamSynchronize(s));
  float tot = 0;
  DMRGX_HIP(hipEventElapsedTime(&tot, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  ms[0] = tot / iters;
  CAPI_END
}

int dmrgx_heff_time(dmrgx_heff* plan, const void* in, void* out, int iters, double* ms) {
  CAPI_BEGIN
  DMRGX_CHECK(plan && plan->p && iters > 0, "heff_time: bad arguments");
  hipStream_t s = rt().stream;
  std::vector<hipEvent_t> ev(4 * iters + 2);
  for (auto& e : ev) DMRGX_HIP(hipEventCreate(&e));
  DMRGX_HIP(hipEventRecord(ev[4 * iters], s));
  for (int i = 0; i < iters; ++i) plan->p->apply((const double*)in, (double*)out, s, &ev[4 * i]);
  DMRGX_HIP(hipEventRecord(ev[4 * iters + 1], s));
  DMRGX_HIP(hipStreamSynchronize(s));
  float tot = 0, r = 0, l = 0, x;
  DMRGX_HIP(hipEventElapsedTime(&tot, ev[4 * iters], ev[4 * iters + 1]));
  for (int i = 0; i < iters; ++i) {
    DMRGX_HIP(hipEventElapsedTime(&x, ev[4 * i], ev[4 * i + 1]));
    r += x;
    DMRGX_HIP(hipEventElapsedTime(&x, ev[4 * i + 2], ev[4 * i + 3]));
    l += x;
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  const HeffPlan& h = *plan->p;
  const double cf = h.dtype == DT_C64 ? 8.0 : 2.0;
  double ar = 0, al = 0;
  for (auto& g : h.rtasks) ar += cf * (double)g.M * g.N * g.K;
  for (auto& g : h.ltasks) al += cf * (double)g.M * g.N * g.K;
  ms[0] = tot / iters;
  ms[1] = r / iters;
  ms[2] = l / iters;
  ms[3] = ar;
  ms[4] = al;
  ms[5] = h.rplan->flops;
  ms[6] = h.lplan->flops;
  CAPI_END
}

int dmrgx_heff_time_kernels(dmrgx_heff* plan, const void* in, void* out, int iters, double* ms) {
  CAPI_BEGIN
  DMRGX_CHECK(plan && plan->p && iters > 0 && ms, "heff_time_kernels: bad arguments");
  hipStream_t s = rt().stream;
  plan->p->time_kernels((const double*)in, (double*)out, iters, s, &ms[0], &ms[1]);
  DMRGX_HIP(hipStreamSynchronize(s));
  CAPI_END
}

int dmrgx_set_skip_zero_slices(int on) {
  set_heff_skip_zeros(on != 0);
  return 0;
}

double dmrgx_heff_flops_exec(const dmrgx_heff* plan) { return plan && plan->p ? plan->p->flops_exec : 0.0; }

double dmrgx_heff_flops(const dmrgx_heff* plan) { return plan && plan->p ? plan->p->flops : 0.0; }

int dmrgx_heff_destroy(dmrgx_heff* plan) {
  CAPI_BEGIN
  if (plan) {
    (void)hipStreamSynchronize(rt().stream);
    delete plan;
  }
  CAPI_END
}

namespace {

// FP64 refinement of a reduced-precision solve: Jacobi-Davidson from the pv
// lowest low-precision Ritz vectors of every chain (in `vecs`, pv columns of
// the concatenated length), the correction equations solved with the
// low-precision operator.  The refined lowest vector of every chain is left
// in column 0.
struct Refined {
  std::vector<double> e0, resid;
  std::vector<char> conv;
  int n64 = 0, nlo = 0;
};
Refined refine_lowest(HeffPlan* lo, HeffPlan* f, const std::vector<int64_t>& len, int pv, double* vecs, double tol) {
  hipStream_t s = rt().stream;
  // the inner solves on the coarse hi-plane product of a BF16x3 plan, as the
  // sweep engine does
  const bool coarse = lo->prec == PREC_BF16X3;
  MatVec op = [lo, s, coarse](const double* x, double* y) { lo->apply(x, y, s, nullptr, coarse); };
  MatVec op64 = [f, s](const double* x, double* y) { f->apply(x, y, s); };
  Refined r;
  // the sweep engine's settings: at most 12 outer steps of at most 16 CG steps
  r.e0 = jd_polish_batched(op, op64, len, f->dtype == DT_C64, pv, vecs, tol, 12, 16, &r.n64, &r.nlo, &r.resid,
                           &r.conv);
  return r;
}

void check_refine(const HeffPlan* h, const HeffPlan* f, const char* who) {
  DMRGX_CHECK(f->n_psi == h->n_psi && f->dtype == h->dtype,
              std::string(who) + ": refine plan of another psi length or element type");
  DMRGX_CHECK(f->prec == PREC_FP64, std::string(who) + ": the refine plan must be an FP64 plan");
}

}  // namespace

int dmrgx_eigsh(dmrgx_heff* plan, int nev, int ncv, double tol, int max_restart, uint64_t seed,
                const void* v0, double* evals, void* evec, dmrgx_heff* refine, double refine_tol, int* info,
                double* resid) {
  CAPI_BEGIN
  DMRGX_CHECK(plan && plan->p && evals && evec, "eigsh: null arguments");
  HeffPlan* h = plan->p.get();
  HeffPlan* f = refine && refine->p ? refine->p.get() : nullptr;
  if (f) check_refine(h, f, "eigsh");
  const bool use_ref = f && h->prec != PREC_FP64;
  const bool cx = h->dtype == DT_C64;
  hipStream_t s = rt().stream;
  MatVec op = [h, s](const double* x, double* y) { h->apply(x, y, s); };
  // a refined solve hands its pv lowest Ritz vectors to the FP64 refinement:
  // four whenever the basis holds them, also when only the lowest pair's
  // residual is tested (nev = 1)
  const int pv = use_ref ? (int)std::max<int64_t>(1, std::min<int64_t>({4, std::max(nev, ncv), h->n_psi})) : 1;
  std::unique_ptr<DevBuf> vecs;
  double* ev = (double*)evec;
  if (pv > 1) {
    vecs = std::make_unique<DevBuf>((size_t)pv * h->n_psi * (cx ? 16 : 8));
    ev = (double*)vecs->ptr;
  }
  // a refined low-precision solve keeps its Krylov basis in FP32
  EigshResult r = eigsh(op, h->n_psi, cx, nev, ncv, tol, max_restart, seed, (const double*)v0, ev, pv, false,
                        use_ref && h->prec != PREC_FP64);
  for (int i = 0; i < (int)r.evals.size(); ++i) evals[i] = r.evals[i];
  Refined rf;
  if (use_ref) {
    rf = refine_lowest(h, f, {h->n_psi}, pv, ev, refine_tol > 0 ? refine_tol : 1e-10);
    evals[0] = rf.e0[0];
    if (pv > 1) copy_d2d(evec, ev, (size_t)h->n_psi * (cx ? 16 : 8), s);
  } else if (resid) {
    rf.resid = residual_batched(op, {h->n_psi}, cx, (const double*)evec, {evals[0]});
  }
  if (info) {
    info[0] = r.n_matvec;
    info[1] = r.n_restart;
    info[2] = r.converged ? 1 : 0;
    info[3] = rf.n64;
    info[4] = use_ref ? (rf.conv[0] ? 1 : 0) : info[2];
    info[5] = rf.nlo;
  }
  if (resid) resid[0] = rf.resid[0];
  DMRGX_HIP(hipStreamSynchronize(rt().stream));  // complete on return
  CAPI_END
}

int dmrgx_eigsh_batched(dmrgx_heff* plan, int n_chains, const int64_t* len, int nev, int ncv, double tol,
                        int max_restart, const uint64_t* seeds, double* evals, void* evec, dmrgx_heff* refine,
                        double refine_tol, int* info, double* resid) {
  CAPI_BEGIN
  DMRGX_CHECK(plan && plan->p && len && seeds && evals && evec && n_chains > 0, "eigsh_batched: null arguments");
  HeffPlan* h = plan->p.get();
  std::vector<int64_t> ln(len, len + n_chains);
  int64_t tot = 0, nmin = ln[0];
  for (auto n : ln) {
    DMRGX_CHECK(n > 0, "eigsh_batched: empty chain");
    tot += n;
    nmin = std::min(nmin, n);
  }
  DMRGX_CHECK(tot == h->n_psi, "eigsh_batched: chain lengths do not add up to the plan's psi length");
  HeffPlan* f = refine && refine->p ? refine->p.get() : nullptr;
  if (f) check_refine(h, f, "eigsh_batched");
  const bool use_ref = f && h->prec != PREC_FP64;
  const bool cx = h->dtype == DT_C64;
  hipStream_t s = rt().stream;
  MatVec op = [h, s](const double* x, double* y) { h->apply(x, y, s); };
  const int pv = use_ref ? (int)std::max<int64_t>(1, std::min<int64_t>({4, std::max(nev, ncv), nmin})) : 1;
  std::unique_ptr<DevBuf> vecs;
  double* ev = (double*)evec;
  if (pv > 1) {
    vecs = std::make_unique<DevBuf>((size_t)pv * h->n_psi * (cx ? 16 : 8));
    ev = (double*)vecs->ptr;
  }
  std::vector<uint64_t> sd(seeds, seeds + n_chains);
  auto r = eigsh_batched(op, ln, cx, nev, ncv, tol, max_restart, sd, nullptr, ev, pv);
  for (int c = 0; c < n_chains; ++c)
    for (int i = 0; i < nev; ++i) evals[(size_t)c * nev + i] = i < (int)r[c].evals.size() ? r[c].evals[i] : 0.0;
  Refined rf;
  if (use_ref) {
    rf = refine_lowest(h, f, ln, pv, ev, refine_tol > 0 ? refine_tol : 1e-10);
    for (int c = 0; c < n_chains; ++c) evals[(size_t)c * nev] = rf.e0[c];
    if (pv > 1) copy_d2d(evec, ev, (size_t)h->n_psi * (cx ? 16 : 8), s);
  } else if (resid) {
    std::vector<double> th(n_chains);
    for (int c = 0; c < n_chains; ++c) th[c] = evals[(size_t)c * nev];
    rf.resid = residual_batched(op, ln, cx, (const double*)evec, th);
  }
  if (info)
    for (int c = 0; c < n_chains; ++c) {
      int* q = info + 6 * c;
      q[0] = r[c].n_matvec;
      q[1] = r[c].n_restart;
      q[2] = r[c].converged ? 1 : 0;
      q[3] = rf.n64;
      q[4] = use_ref ? (rf.conv[c] ? 1 : 0) : q[2];
      q[5] = rf.nlo;
    }
  if (resid)
    for (int c = 0; c < n_chains; ++c) resid[c] = rf.resid[c];
  DMRGX_HIP(hipStreamSynchronize(rt().stream));  // complete on return
  CAPI_END
}

int dmrgx_set_autotune(int on) {
  set_heff_autotune(on != 0);
  return 0;
}

int dmrgx_autotune_keys(void) { return heff_autotune_keys(); }

int dmrgx_heff_set_split(dmrgx_heff* plan, int r_target, int l_target) {
  CAPI_BEGIN
  DMRGX_CHECK(plan && plan->p, "heff_set_split: null plan");
  plan->p->replan(r_target, l_target);
  DMRGX_HIP(hipStreamSynchronize(rt().stream));
  CAPI_END
}

int dmrgx_heff_get_split(const dmrgx_heff* plan, int* targets) {
  CAPI_BEGIN
  DMRGX_CHECK(plan && plan->p && targets, "heff_get_split: null argument");
  targets[0] = plan->p->r_split;
  targets[1] = plan->p->l_split;
  targets[2] = plan->p->rplan->n_items;
  targets[3] = plan->p->lplan->n_items;
  CAPI_END
}

int dmrgx_set_lanczos_graphs(int on) {
  set_lanczos_graphs(on != 0);
  return 0;
}

int dmrgx_block_gemm(int dtype, char ta, char tb, const void* A, const void* B, void* C, const int64_t* tasks,
                     int n_task) {
  CAPI_BEGIN
  ensure_init();
  std::vector<GemmTask> t(n_task);
  for (int i = 0; i < n_task; ++i) {
    const int64_t* q = tasks + 7 * i;
    t[i] = GemmTask{q[0], q[1], q[2], (int)q[3], (int)q[4], (int)q[5], 1, (int)q[6]};
  }
  auto plan = make_gemm_plan(t, dtype, PREC_FP64, ta, tb, {});
  run_gemm_plan(*plan, A, B, C, rt().stream);
  DMRGX_HIP(hipStreamSynchronize(rt().stream));
  CAPI_END
}

}  // extern "C"
