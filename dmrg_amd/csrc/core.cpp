// Step functions (src/core.cpp:88-474) and sweep schedules
// (src/functions.cpp:157-299) over device-resident qtensors.  The matvec,
// eigensolver, SVD and environment contractions are the HIP kernels; the
// control flow, sector bookkeeping and step order follow the reference.
#include "core.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <set>

#include "lanczos.hpp"
#include "observables.hpp"
#include "output.hpp"
#include "svd.hpp"

namespace dmrgx {

using clk = std::chrono::steady_clock;
static double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

template <typename T>
Engine<T>::Engine(const mpo<T>& m, const RunConfig& c) : mpo_(m), mps_(m), cfg(c) {
  if (!cfg.output_suffix.empty()) {
    // src/functions.cpp:112-116, 163-167
    f_energy_ = std::make_unique<std::ofstream>("energy" + cfg.output_suffix);
    f_singular_ = std::make_unique<std::ofstream>("singular" + cfg.output_suffix);
    DMRGX_CHECK(f_energy_->good() && f_singular_->good(), "cannot open the output files for " + cfg.output_suffix);
  }
}

template <typename T>
qtensor<T> Engine<T>::make_lenv(int l) {
  // src/core.cpp:188-189
  qtensor<T> t;
  t.contract(mps_(l - 1), 1, mpo_[l], 0);
  return t.exchange(3, 4);
}

template <typename T>
qtensor<T> Engine<T>::make_renv(int r) {
  // src/core.cpp:190-192
  qtensor<T> ope = mpo_[r].exchange(0, 2);
  qtensor<T> t;
  t.contract(mps_(r + 1), 1, ope, 0, true);
  return t.exchange(0, 1);
}

template <typename T>
qtensor<T> Engine<T>::make_ledge(int l) {
  // src/core.cpp:193-196
  qtensor<T> ope = mps_.edge(l - 1).id(mpo_[l]);
  qtensor<T> t;
  t.contract(mps_.edge(l - 1), 1, ope, 0);
  return t.exchange(3, 4);
}

template <typename T>
qtensor<T> Engine<T>::make_redge(int r, int from) {
  // src/core.cpp:193-199 (the identity built from edge(l-1)), :417-420 (from edge(r+1))
  qtensor<T> ope = mps_.edge(from).id(mpo_[r]);
  qtensor<T> t;
  t.contract(mps_.edge(r + 1), 1, ope, 0, true);
  return t.exchange(0, 1);
}

template <typename T>
void Engine<T>::update_ledge(int l, const qtensor<T>& ledge, qtensor<T> U, int num) {
  // src/core.cpp:273-276
  qtensor<T> vec;
  vec.contract(ledge, U, 'N', 'N', num);
  U.conjugate();
  mps_.edge(l).contract(U, vec, 'T', 'N', num);
}

template <typename T>
void Engine<T>::update_redge(int r, const qtensor<T>& redge, qtensor<T> V, int num) {
  // src/core.cpp:277-281
  qtensor<T> vec;
  vec.contract(redge, V, 'N', 'T', num);
  V.conjugate();
  mps_.edge(r).contract(V, vec, 'N', 'N', num);
}

static void flatten_map(const vector<vector<int>>& m, vector<int32_t>& out) {
  out.clear();
  for (auto& t : m) out.insert(out.end(), t.begin(), t.end());
}

// one H_eff plan over the two environments' task lists
template <typename T>
static std::unique_ptr<HeffPlan> plan_of(const qtensor<T>& renv, const vector<vector<int>>& rmap,
                                         const qtensor<T>& lenv, const vector<vector<int>>& lmap, int64_t d,
                                         int64_t d_store, int prec) {
  vector<int32_t> rm, lm;
  flatten_map(rmap, rm);
  flatten_map(lmap, lm);
  vector<int64_t> rsz(renv.nblocks()), lsz(lenv.nblocks());
  for (int b = 0; b < renv.nblocks(); ++b) rsz[b] = renv.bsize(b);
  for (int b = 0; b < lenv.nblocks(); ++b) lsz[b] = lenv.bsize(b);
  return make_heff(dtype_of<T>::value, prec, renv.data(), renv.off_.data(), rsz.data(), renv.nblocks(), renv.total_,
                   rm.data(), (int)rmap.size(), lenv.data(), lenv.off_.data(), lsz.data(), lenv.nblocks(),
                   lenv.total_, lm.data(), (int)lmap.size(), d, d_store);
}

template <typename T>
void Engine<T>::solve(const qtensor<T>& lenv, const qtensor<T>& renv, int lnum, int rnum, int r_num, int l_num,
                      vector<vector<int>>& dim, vector<vector<int>>& sym, double* vals, qtensor<T>& psi,
                      StepLog& lg, bool allow_ed, int ed_which, const qtensor<T>* guess, const qtensor<T>* ledge,
                      const qtensor<T>* redge) {
  DMRGX_CHECK((ledge == nullptr) == (redge == nullptr), "solve: one edge environment without the other");
  generate_dim_sym(dim, sym, lenv, lnum, renv, rnum, cfg.sector);
  const int64_t d = get_dimension(dim, sym);
  lg.d = d;
  for (int i = 0; i < NEV; ++i) vals[i] = 0;
  if (allow_ed && d < NEV) {
    // exact diagonalisation of the tiny superblock (src/core.cpp:106-124, 314-336, 410-433)
    auto pair = [ed_which](const qtensor<T>& le, const qtensor<T>& re) {
      qtensor<T> w;
      if (ed_which == 0) w.contract(le, 1, re, 1);
      else if (ed_which == 1) w.contract(le.combine(3, 4).combine(0, 1), 1, re, 1);
      else w.contract(le, 1, re.combine(3, 4).combine(0, 1), 1);
      return w.simplify();
    };
    qtensor<T> whole = pair(lenv, renv);
    if (ledge) {
      // bulk + edge (src/core.cpp:111-116, 323-331, 419-427)
      qtensor<T> edge = pair(*ledge, *redge), all;
      if (ed_which == 0) all.plus(edge, whole);
      else all.plus(whole, edge);
      whole = all.simplify();
    }
    whole = whole.exchange(2, 3).combine(2, 3).combine(0, 1);
    vector<T> v;
    std::vector<double> w(std::max<int64_t>(d, NEV) + whole.dimension(0), 0.0);
    int n = whole.eig(w.data(), v, cfg.sector);
    for (int i = 0; i < std::min(n, NEV); ++i) vals[i] = w[i];
    lg.n_eval = (int)d;  // src/core.cpp:123
    DMRGX_CHECK((int64_t)v.size() == d, "exact diagonalisation: eigenvector length differs from the psi dimension");
    DevBuf tmp(std::max<int64_t>(d, 1) * sizeof(T));
    DMRGX_HIP(hipMemcpyAsync(tmp.ptr, v.data(), d * sizeof(T), hipMemcpyHostToDevice, rt().stream));
    psi = qtensor<T>((const T*)tmp.ptr, dim, sym);
    DMRGX_HIP(hipStreamSynchronize(rt().stream));
    return;
  }
  auto tp0 = clk::now();
  vector<vector<int>> rmap, lmap, dim_ret, sym_ret, dim2;
  DMRGX_CHECK(renv.get_map(rmap, dim_ret, sym_ret, dim, sym, 'N', 'T', r_num, 0, false, alpha_on()),
              "get_map (rmap) failed");
  const int64_t d_store = get_dimension(dim_ret, sym_ret);
  vector<vector<int>> psym = sym;
  DMRGX_CHECK(lenv.get_map(lmap, dim2, psym, dim_ret, sym_ret, 'N', 'T', l_num, 0, true, alpha_on()),
              "get_map (lmap) failed");
  const int dt = dtype_of<T>::value;
  auto plan = plan_of(renv, rmap, lenv, lmap, d, d_store, cfg.prec);
  std::unique_ptr<HeffPlan> plan64;
  if (cfg.prec != PREC_FP64) plan64 = plan_of(renv, rmap, lenv, lmap, d, d_store, PREC_FP64);
  if (ledge) {
    // the edge term's own task lists over its own store layout
    // (src/core.cpp:241-247); the reference accumulates every edge L task
    // onto the bulk result (beta = 1), the edge plan writes its first task
    // per output block and apply() adds the result
    vector<vector<int>> remap, lemap, edim_ret, esym_ret, edim2;
    DMRGX_CHECK(redge->get_map(remap, edim_ret, esym_ret, dim, sym, 'N', 'T', r_num, 0, false, alpha_on()),
                "get_map (edge rmap) failed");
    const int64_t e_store = get_dimension(edim_ret, esym_ret);
    vector<vector<int>> esym = sym;
    DMRGX_CHECK(ledge->get_map(lemap, edim2, esym, edim_ret, esym_ret, 'N', 'T', l_num, 1, true, alpha_on()),
                "get_map (edge lmap) failed");
    std::set<int> written;
    for (auto& t : lemap) t[5] = written.insert(t[2]).second ? 0 : 1;
    attach_edge(*plan, plan_of(*redge, remap, *ledge, lemap, d, e_store, cfg.prec));
    if (plan64) attach_edge(*plan64, plan_of(*redge, remap, *ledge, lemap, d, e_store, PREC_FP64));
  }
  hipStream_t s = rt().stream;
  HeffPlan* hp = plan.get();
  MatVec op = [hp, s](const double* x, double* y) { hp->apply(x, y, s); };
  const bool cx = dt == DT_C64;
  // start vector: the predicted state in this solve's block order, when it
  // has weight there (a truncated U S V keeps norm close to 1); else random
  qtensor<T> start;
  const double* v0 = nullptr;
  if (guess) {
    int matched = 0;
    start = embed_state(*guess, dim, sym, &matched);
    if (matched > 0 && dev_norm2((const double*)start.data(), d, cx, s) > 0.1) v0 = (const double*)start.data();
  }
  lg.warm = v0 != nullptr;
  // a reduced-precision solve tests the residuals of its nev_lo lowest pairs,
  // sizes its basis for four and hands its pv lowest Ritz vectors to the
  // FP64 polish; a warm start converges the lowest pair only
  const int nev = lg.warm ? 1 : plan64 ? std::min(cfg.nev_lo, cfg.nev) : cfg.nev;
  const int nb = lg.warm ? 1 : plan64 ? std::max(nev, 4) : nev;
  // ARPACK's basis size for nb pairs (src/dsaupd.cpp:76, src/znaupd.cpp:94);
  // a warm start keeps a 20-vector basis for its one pair
  const int ncv = lg.warm ? 20 : (cx ? 6 : 4) * nb;
  const int pv = plan64 ? (int)std::max<int64_t>(1, std::min<int64_t>({4, nb, d})) : 1;
  DevBuf evec(std::max<int64_t>(d, 1) * sizeof(T) * pv);
  const double tol = cfg.prec == PREC_FP64 ? cfg.tol : cfg.tol_lo;
  uint64_t seed = cfg.seed * 0x9E3779B97F4A7C15ull + (uint64_t)log.size() + 1;
  auto tp1 = clk::now();
  // a reduced-precision solve keeps its Krylov basis in FP32 (its 1e-4
  // tolerance is far above the basis' orthogonality; the FP64 refinement
  // follows)
  EigshResult r = eigsh(op, d, cx, nev, ncv, tol, cfg.max_restart, seed, v0, (double*)evec.ptr, pv, lg.warm,
                        plan64 != nullptr);
  auto tp2 = clk::now();
  lg.t_plan = secs(tp0, tp1);
  lg.t_lanczos = secs(tp1, tp2);
  for (int i = 0; i < std::min<int>(NEV, (int)r.evals.size()); ++i) vals[i] = r.evals[i];
  lg.n_matvec = r.n_matvec;
  lg.flops = plan->flops * r.n_matvec;
  if (plan64) {
    HeffPlan* hp64 = plan64.get();
    MatVec op64 = [hp64, s](const double* x, double* y) { hp64->apply(x, y, s); };
    int pm = 0, plo = 0;
    // the inner solves (a few digits deep) on the coarse product of a BF16x3 plan
    MatVec op_in = cfg.prec == PREC_BF16X3 ? MatVec([hp, s](const double* x, double* y) {
      hp->apply(x, y, s, nullptr, true);
    }) : op;
    vals[0] = jd_polish_batched(op_in, op64, {d}, cx, pv, (double*)evec.ptr, cfg.tol, cfg.polish_cycles,
                                2 * std::max(1, cfg.refine_k), &pm, &plo)[0];
    lg.n_polish = pm;
    lg.n_matvec += plo;
    lg.flops += plan64->flops * pm + plan->flops * plo;
    lg.t_polish = secs(tp2, clk::now());
  }
  psi = qtensor<T>((const T*)evec.ptr, dim, sym);
  DMRGX_HIP(hipStreamSynchronize(s));
}

template <typename T>
void Engine<T>::finish(StepLog& lg, const vector<double>& s, int cutoff) {
  // print_singular, src/useful.cpp:75-102
  double ent = 0;
  for (double x : s) ent += -x * x * std::log(x * x + TOL);
  double err = 1;
  int c = std::min<int>((int)s.size(), cutoff);
  for (int i = 0; i < c; ++i) err -= s[i] * s[i];
  lg.entropy = ent;
  lg.trunc = err;
  lg.s.assign(s.begin(), s.begin() + std::min<size_t>(s.size(), NSI));
  if (f_energy_) {
    write_energy_line(*f_energy_, lg.l, lg.r, lg.e, lg.n_eval);
    write_singular_line(*f_singular_, lg.l, lg.r, s, cutoff);
    f_energy_->flush();
    f_singular_->flush();
  }
  log.push_back(lg);
  if (cfg.verbose) {
    // 1: one line per step on stdout; 2: on stderr (stdout stays free for results)
    FILE* f = cfg.verbose == 2 ? stderr : stdout;
    std::fprintf(f,
                 "l=%d r=%d d=%lld E0=%.12f trunc=%.3e matvec=%d polish=%d t=%.3fs (plan %.3f lanczos %.3f polish "
                 "%.3f svd %.3f env %.3f)\n",
                 lg.l, lg.r, (long long)lg.d, lg.e[0], lg.trunc, lg.n_matvec, lg.n_polish, lg.t_total, lg.t_plan,
                 lg.t_lanczos, lg.t_polish, lg.t_svd, lg.t_env);
    std::fflush(f);
  }
  if (on_step) on_step(lg);
}

// src/core.cpp:88-172
template <typename T>
double Engine<T>::two_sites(int l, int r, int cutoff) {
  auto t0 = clk::now();
  StepLog lg;
  lg.l = l;
  lg.r = r;
  qtensor<T> lenv = mpo_[l], renv = mpo_[r];
  const bool pbc = mps_.pbc();
  qtensor<T> ledge, redge;
  if (pbc) {
    // src/core.cpp:96-99
    ledge = mps_.edge(l);
    redge = mps_.edge(r);
  }
  vector<vector<int>> dim, sym;
  qtensor<T> psi;
  auto t1 = clk::now();
  solve(lenv, renv, 1, 1, 1, 2, dim, sym, lg.e, psi, lg, true, 0, nullptr, pbc ? &ledge : nullptr,
        pbc ? &redge : nullptr);
  auto t2 = clk::now();
  qtensor<T> U, V;
  qtensor<double> S;
  vector<double> s = psi.svd(U, S, V, 1, cutoff);
  auto t3 = clk::now();
  mps_[l] = U;
  mps_.center(l + 1) = S;
  mps_[r] = V;
  const qtensor<T> U0 = U, V0 = V;
  qtensor<T> vec;
  vec.contract(lenv, U, 'N', 'N');
  U.conjugate();
  mps_(l).contract(U, vec, 'T', 'N');
  vec.contract(renv, V, 'N', 'T');
  V.conjugate();
  mps_(r).contract(V, vec, 'N', 'N');
  if (pbc) {
    // src/core.cpp:161-170
    update_ledge(l, ledge, U0, 1);
    update_redge(r, redge, V0, 1);
  }
  auto t4 = clk::now();
  lg.t_env = secs(t0, t1) + secs(t3, t4);
  lg.t_eig = secs(t1, t2);
  lg.t_svd = secs(t2, t3);
  lg.svd_sweeps = block_svd_last_sweeps();
  lg.t_block_svd = block_svd_last_seconds();
  lg.t_total = secs(t0, t4);
  finish(lg, s, cutoff);
  return lg.e[0];
}

// White's prediction (S. R. White, PRL 77, 3633 (1996)): after the step on
// (l-1, l) the state is A_{l-1} S V_l B_{l+1}..., so the two-site state on
// (l, l+1) is (S V_l) B_{l+1}; after (r, r+1) it is A_l (U_r S).  The
// environments of the new step are built from the same A / B tensors, so
// the prediction is already in the new step's basis.
template <typename T>
bool Engine<T>::predict(int l, int r, qtensor<T>& guess) {
  if (!cfg.warm_start || log.empty() || mps_.cut_ <= 0) return false;
  const StepLog& p = log.back();
  const qtensor<T> S = centre_as<T>(mps_.singular_);
  qtensor<T> half;
  if (p.l == l - 1 && p.r == r - 1 && mps_.cut_ == l) {
    half.contract(S, mps_[l]);
    guess.contract(half, mps_[r]);
  } else if (p.l == l + 1 && p.r == r + 1 && mps_.cut_ == r + 1) {
    half.contract(mps_[r], S);
    guess.contract(mps_[l], half);
  } else {
    return false;
  }
  return true;
}

// src/core.cpp:182-283
template <typename T>
double Engine<T>::update_two(int l, int r, int cutoff) {
  auto t0 = clk::now();
  StepLog lg;
  lg.l = l;
  lg.r = r;
  qtensor<T> lenv = make_lenv(l), renv = make_renv(r);
  const bool pbc = mps_.pbc();
  qtensor<T> ledge, redge;
  if (pbc) {
    ledge = make_ledge(l);
    redge = make_redge(r, l - 1);
  }
  vector<vector<int>> dim, sym;
  qtensor<T> psi, guess;
  const bool warm = predict(l, r, guess);
  auto t1 = clk::now();
  solve(lenv, renv, 2, 2, 2, 3, dim, sym, lg.e, psi, lg, false, 0, warm ? &guess : nullptr,
        pbc ? &ledge : nullptr, pbc ? &redge : nullptr);
  auto t2 = clk::now();
  qtensor<T> U, V;
  qtensor<double> S;
  vector<double> s = psi.svd(U, S, V, 2, cutoff);
  auto t3 = clk::now();
  mps_[l] = U;
  mps_.center(l + 1) = S;
  mps_[r] = V;
  const qtensor<T> U0 = U, V0 = V;
  qtensor<T> vec;
  vec.contract(lenv, U, 'N', 'N', 2);
  U.conjugate();
  mps_(l).contract(U, vec, 'T', 'N', 2);
  vec.contract(renv, V, 'N', 'T', 2);
  V.conjugate();
  mps_(r).contract(V, vec, 'N', 'N', 2);
  if (pbc) {
    update_ledge(l, ledge, U0, 2);
    update_redge(r, redge, V0, 2);
  }
  auto t4 = clk::now();
  lg.t_env = secs(t0, t1) + secs(t3, t4);
  lg.t_eig = secs(t1, t2);
  lg.t_svd = secs(t2, t3);
  lg.svd_sweeps = block_svd_last_sweeps();
  lg.t_block_svd = block_svd_last_seconds();
  lg.t_total = secs(t0, t4);
  finish(lg, s, cutoff);
  return lg.e[0];
}

// src/core.cpp:293-378
template <typename T>
double Engine<T>::move2right(int l, int r, int cutoff) {
  auto t0 = clk::now();
  StepLog lg;
  lg.l = l;
  lg.r = r;
  qtensor<T> lenv = make_lenv(l), renv = mps_(r);
  const bool pbc = mps_.pbc();
  qtensor<T> ledge, redge;
  if (pbc) {
    // src/core.cpp:302-308
    ledge = make_ledge(l);
    redge = mps_.edge(r);
  }
  vector<vector<int>> dim, sym;
  qtensor<T> psi;
  auto t1 = clk::now();
  solve(lenv, renv, 2, 1, 1, 3, dim, sym, lg.e, psi, lg, true, 1, nullptr, pbc ? &ledge : nullptr,
        pbc ? &redge : nullptr);
  auto t2 = clk::now();
  qtensor<T> U, V;
  qtensor<double> S;
  vector<double> s = psi.svd(U, S, V, 2, cutoff);
  auto t3 = clk::now();
  mps_[l] = U;
  mps_.center(l + 1) = S;
  const qtensor<T> U0 = U;
  qtensor<T> vec;
  vec.contract(lenv, U, 'N', 'N', 2);
  U.conjugate();
  mps_(l).contract(U, vec, 'T', 'N', 2);
  if (pbc) update_ledge(l, ledge, U0, 2);  // src/core.cpp:371-376
  auto t4 = clk::now();
  lg.t_env = secs(t0, t1) + secs(t3, t4);
  lg.t_eig = secs(t1, t2);
  lg.t_svd = secs(t2, t3);
  lg.svd_sweeps = block_svd_last_sweeps();
  lg.t_block_svd = block_svd_last_seconds();
  lg.t_total = secs(t0, t4);
  finish(lg, s, cutoff);
  return lg.e[0];
}

// src/core.cpp:388-474
template <typename T>
double Engine<T>::move2left(int l, int r, int cutoff) {
  auto t0 = clk::now();
  StepLog lg;
  lg.l = l;
  lg.r = r;
  qtensor<T> lenv = mps_(l), renv = make_renv(r);
  const bool pbc = mps_.pbc();
  qtensor<T> ledge, redge;
  if (pbc) {
    // src/core.cpp:398-404
    ledge = mps_.edge(l);
    redge = make_redge(r, r + 1);
  }
  vector<vector<int>> dim, sym;
  qtensor<T> psi;
  auto t1 = clk::now();
  solve(lenv, renv, 1, 2, 2, 2, dim, sym, lg.e, psi, lg, true, 2, nullptr, pbc ? &ledge : nullptr,
        pbc ? &redge : nullptr);
  auto t2 = clk::now();
  qtensor<T> U, V;
  qtensor<double> S;
  vector<double> s = psi.svd(U, S, V, 1, cutoff);
  auto t3 = clk::now();
  mps_.center(l + 1) = S;
  mps_[r] = V;
  const qtensor<T> V0 = V;
  qtensor<T> vec;
  vec.contract(renv, V, 'N', 'T', 2);
  V.conjugate();
  mps_(r).contract(V, vec, 'N', 'N', 2);
  if (pbc) update_redge(r, redge, V0, 2);  // src/core.cpp:467-472
  auto t4 = clk::now();
  lg.t_env = secs(t0, t1) + secs(t3, t4);
  lg.t_eig = secs(t1, t2);
  lg.t_svd = secs(t2, t3);
  lg.svd_sweeps = block_svd_last_sweeps();
  lg.t_block_svd = block_svd_last_seconds();
  lg.t_total = secs(t0, t4);
  finish(lg, s, cutoff);
  return lg.e[0];
}

// single-site schedule, src/functions.cpp:157-227
template <typename T>
double Engine<T>::dmrg(int cutoff, int sweep) {
  const int L = mps_.size();
  int pre_cutoff = sweep == -1 ? cutoff : 10;
  int pre_sweep = cutoff / 100;
  double e = two_sites(0, L - 1, cutoff * 10);
  for (int i = 1; i < L / 2; ++i) e = update_two(i, L - i - 1, pre_cutoff);
  if (sweep != -1) {
    for (int i = L / 2; i < L - 1; ++i) e = move2right(i, i + 1, pre_cutoff);
    for (int s = 0; s < pre_sweep - 1; ++s) {
      for (int i = L - 2; i > 0; --i) e = move2left(i - 1, i, 100 * (s + 1));
      for (int i = 1; i < L - 1; ++i) e = move2right(i, i + 1, 100 * (s + 1));
    }
  }
  for (int s = 0; s < sweep; ++s) {
    for (int i = L - 2; i > 0; --i) e = move2left(i - 1, i, cutoff);
    for (int i = 1; i < L - 1; ++i) e = move2right(i, i + 1, cutoff);
  }
  if (sweep != -1)
    for (int i = L - 2; i >= L / 2; --i) e = move2left(i - 1, i, cutoff);
  return e;
}

// two-site schedule, src/functions.cpp:234-299
double schedule_dmrg2(int L, int cutoff, int sweep, bool ladder, const std::function<double()>& first,
                      const std::function<double(int, int, int)>& step) {
  int pre_cutoff = sweep == -1 ? cutoff : 10;
  int pre_sweep = ladder ? cutoff / 100 : 0;
  double e = first();
  for (int i = 1; i < L / 2; ++i) e = step(i, L - i - 1, pre_cutoff);
  if (sweep != -1) {
    for (int i = L / 2; i < L - 2; ++i) e = step(i, i + 1, pre_cutoff);
    for (int s = 0; s < pre_sweep - 1; ++s) {
      for (int i = L - 3; i > 1; --i) e = step(i - 1, i, 100 * (s + 1));
      for (int i = 2; i < L - 2; ++i) e = step(i, i + 1, 100 * (s + 1));
    }
  }
  for (int s = 0; s < sweep; ++s) {
    for (int i = L - 3; i > 1; --i) e = step(i - 1, i, cutoff);
    for (int i = 2; i < L - 2; ++i) e = step(i, i + 1, cutoff);
  }
  if (sweep != -1)
    for (int i = L - 3; i >= L / 2; --i) e = step(i - 1, i, cutoff);
  return e;
}

template <typename T>
double Engine<T>::dmrg2(int cutoff, int sweep, bool ladder) {
  const int L = mps_.size();
  return schedule_dmrg2(
      L, cutoff, sweep, ladder, [&] { return two_sites(0, L - 1, 0); },
      [&](int l, int r, int c) { return update_two(l, r, c); });
}

template class Engine<double>;
template class Engine<cplx>;

}  // namespace dmrgx
