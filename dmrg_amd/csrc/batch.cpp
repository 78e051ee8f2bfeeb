// Lockstep sweeps of many independent chains (see batch.hpp).
#include "batch.hpp"

#include <chrono>
#include <limits>

#include "heff.hpp"
#include "lanczos.hpp"
#include "svd.hpp"

namespace dmrgx {

using clk = std::chrono::steady_clock;
static double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

template <typename T>
BatchEngine<T>::BatchEngine(const std::vector<mpo<T>>& m, const RunConfig& c) : cfg(c) {
  DMRGX_CHECK(!m.empty(), "batch: no chains");
  for (auto& x : m) {
    DMRGX_CHECK(x.size() == m[0].size(), "batch: chains of different lengths");
    chains.push_back(std::make_unique<Engine<T>>(x, c));
  }
}

static int32_t narrow(int64_t v, const char* what) {
  DMRGX_CHECK(v >= 0 && v <= std::numeric_limits<int32_t>::max(), std::string("batch: ") + what +
                                                                       " offset beyond the int32 task format");
  return (int32_t)v;
}

// The chains' two-site problems as one: psi / store of chain c at element
// offsets P_c / S_c of the concatenated vectors, environment blocks in two
// arenas, get_map tuples shifted accordingly (block index, input and output
// offsets).  Then eigsh_batched (+ the FP64 polish of a reduced-precision
// solve) as in Engine::solve.
template <typename T>
void BatchEngine<T>::solve(std::vector<qtensor<T>>& lenv, std::vector<qtensor<T>>& renv,
                           std::vector<qtensor<T>>& psi, std::vector<StepLog>& lg, BatchStepLog& bl) {
  const int n = (int)chains.size();
  const bool alpha = chains[0]->alpha_on();
  const int dt = dtype_of<T>::value;
  const bool cx = dt == DT_C64;
  const double cf = cx ? 8.0 : 2.0;
  hipStream_t s = rt().stream;
  auto tp0 = clk::now();
  vector<vector<vector<int>>> dim(n), sym(n);
  vector<int64_t> d(n), poff(n);
  vector<int32_t> rm, lm;
  vector<int64_t> roff, rsz, loff, lsz;
  vector<double> cflops(n, 0.0);
  vector<int> rt_end(n), lt_end(n);  // each chain's tasks end here in the merged lists
  int64_t P = 0, SX = 0, RT = 0, LT = 0;
  for (int c = 0; c < n; ++c) {
    const qtensor<T>& L = lenv[c];
    const qtensor<T>& R = renv[c];
    generate_dim_sym(dim[c], sym[c], L, 2, R, 2, cfg.sector);
    d[c] = get_dimension(dim[c], sym[c]);
    DMRGX_CHECK(d[c] > 0, "batch: empty two-site space");
    lg[c].d = d[c];
    vector<vector<int>> rmap, lmap, dim_ret, sym_ret, dim2;
    DMRGX_CHECK(R.get_map(rmap, dim_ret, sym_ret, dim[c], sym[c], 'N', 'T', 2, 0, false, alpha),
                "get_map (rmap) failed");
    const int64_t ds = get_dimension(dim_ret, sym_ret);
    vector<vector<int>> psym = sym[c];
    DMRGX_CHECK(L.get_map(lmap, dim2, psym, dim_ret, sym_ret, 'N', 'T', 3, 0, true, alpha),
                "get_map (lmap) failed");
    const int rb = (int)roff.size(), lb = (int)loff.size();
    for (int b = 0; b < R.nblocks(); ++b) {
      roff.push_back(R.off_[b] + RT);
      rsz.push_back(R.bsize(b));
    }
    for (int b = 0; b < L.nblocks(); ++b) {
      loff.push_back(L.off_[b] + LT);
      lsz.push_back(L.bsize(b));
    }
    for (auto& t : rmap) {
      rm.insert(rm.end(), {narrow((int64_t)t[0] + rb, "block"), narrow(t[1] + P, "psi"), narrow(t[2] + SX, "store"),
                           t[3], t[4], t[5], t[6]});
      cflops[c] += cf * (double)(R.bsize(t[0]) / t[4]) * t[3] * t[4];
    }
    for (auto& t : lmap) {
      lm.insert(lm.end(), {narrow((int64_t)t[0] + lb, "block"), narrow(t[1] + SX, "store"), narrow(t[2] + P, "psi"),
                           t[3], t[4], t[5], t[6]});
      cflops[c] += cf * (double)(L.bsize(t[0]) / t[4]) * t[3] * t[4];
    }
    rt_end[c] = (int)(rm.size() / 7);
    lt_end[c] = (int)(lm.size() / 7);
    poff[c] = P;
    P += d[c];
    SX += ds;
    RT += R.total_;
    LT += L.total_;
  }
  // the environments of every chain in two arenas
  DevBuf Ra(std::max<int64_t>(RT, 1) * sizeof(T)), La(std::max<int64_t>(LT, 1) * sizeof(T));
  {
    int64_t ro = 0, lo = 0;
    for (int c = 0; c < n; ++c) {
      if (renv[c].total_) copy_d2d((T*)Ra.ptr + ro, renv[c].data(), renv[c].total_ * sizeof(T), s);
      if (lenv[c].total_) copy_d2d((T*)La.ptr + lo, lenv[c].data(), lenv[c].total_ * sizeof(T), s);
      ro += renv[c].total_;
      lo += lenv[c].total_;
    }
  }
  auto plan = make_heff(dt, cfg.prec, Ra.ptr, roff.data(), rsz.data(), (int)roff.size(), RT, rm.data(),
                        (int)(rm.size() / 7), La.ptr, loff.data(), lsz.data(), (int)loff.size(), LT, lm.data(),
                        (int)(lm.size() / 7), P, SX);
  std::unique_ptr<HeffPlan> plan64;
  if (cfg.prec != PREC_FP64)
    plan64 = make_heff(dt, PREC_FP64, Ra.ptr, roff.data(), rsz.data(), (int)roff.size(), RT, rm.data(),
                       (int)(rm.size() / 7), La.ptr, loff.data(), lsz.data(), (int)loff.size(), LT, lm.data(),
                       (int)(lm.size() / 7), P, SX);
  bl.t_plan = secs(tp0, clk::now());
  bl.d = P;
  HeffPlan* hp = plan.get();
  MatVec op = [hp, s](const double* x, double* y) { hp->apply(x, y, s); };
  const double tol = cfg.prec == PREC_FP64 ? cfg.tol : cfg.tol_lo;
  vector<uint64_t> seeds(n);
  for (int c = 0; c < n; ++c)
    seeds[c] = chains[c]->cfg.seed * 0x9E3779B97F4A7C15ull + (uint64_t)chains[c]->log.size() + 1;
  // a reduced-precision solve hands its pv lowest Ritz vectors to the FP64 polish
  int64_t dmin = d[0];
  for (auto x : d) dmin = std::min(dmin, x);
  // (it tests the residuals of its nev_lo lowest pairs and sizes its basis for four)
  const int nev = plan64 ? std::min(cfg.nev_lo, cfg.nev) : cfg.nev;
  const int nb = plan64 ? std::max(nev, 4) : nev;
  // ARPACK's basis size for nb pairs (src/dsaupd.cpp:76, src/znaupd.cpp:94)
  const int ncv = (cx ? 6 : 4) * nb;
  const int pv = plan64 ? (int)std::max<int64_t>(1, std::min<int64_t>({4, nb, dmin})) : 1;
  DevBuf evec(P * sizeof(T) * pv);
  // converged chains leave the operator at the end of their restart cycle
  RetireFn retire = [&](const std::vector<char>& active) {
    std::vector<char> kr(rt_end[n - 1]), kl(lt_end[n - 1]);
    for (int c = 0; c < n; ++c) {
      for (int i = c ? rt_end[c - 1] : 0; i < rt_end[c]; ++i) kr[i] = active[c];
      for (int i = c ? lt_end[c - 1] : 0; i < lt_end[c]; ++i) kl[i] = active[c];
    }
    hp->restrict_tasks(kr, kl);
  };
  vector<EigshResult> r = eigsh_batched(op, d, cx, nev, ncv, tol, cfg.max_restart, seeds, nullptr,
                                        (double*)evec.ptr, pv, retire);
  hp->restrict_tasks({}, {});
  int nm = 0;
  for (int c = 0; c < n; ++c) {
    for (int i = 0; i < NEV; ++i) lg[c].e[i] = i < (int)r[c].evals.size() ? r[c].evals[i] : 0.0;
    lg[c].n_matvec = r[c].n_matvec;
    lg[c].flops = cflops[c] * r[c].n_matvec;
    nm = std::max(nm, r[c].n_matvec);
  }
  bl.n_matvec = nm;
  for (int c = 0; c < n; ++c) bl.flops += lg[c].flops;  // retired chains stop adding work
  if (plan64) {
    HeffPlan* hp64 = plan64.get();
    MatVec op64 = [hp64, s](const double* x, double* y) { hp64->apply(x, y, s); };
    int pm = 0, plo = 0;
    // the inner solves (a few digits deep) on the coarse product of a BF16x3 plan
    MatVec op_in = cfg.prec == PREC_BF16X3 ? MatVec([hp, s](const double* x, double* y) {
      hp->apply(x, y, s, nullptr, true);
    }) : op;
    vector<double> e0 = jd_polish_batched(op_in, op64, d, cx, pv, (double*)evec.ptr, cfg.tol, cfg.polish_cycles,
                                          2 * std::max(1, cfg.refine_k), &pm, &plo);
    for (int c = 0; c < n; ++c) {
      lg[c].e[0] = e0[c];
      lg[c].n_polish = pm;
      lg[c].n_matvec += plo;
      lg[c].flops += cflops[c] * (pm + plo);
    }
    bl.n_polish = pm;
    bl.n_matvec += plo;
    bl.flops += plan64->flops * pm + plan->flops * plo;
  }
  for (int c = 0; c < n; ++c) psi[c] = qtensor<T>((const T*)evec.ptr + poff[c], dim[c], sym[c]);
  DMRGX_HIP(hipStreamSynchronize(s));
}

// src/core.cpp:182-283 on every chain
template <typename T>
double BatchEngine<T>::update_two(int l, int r, int cutoff) {
  const int n = (int)chains.size();
  auto t0 = clk::now();
  BatchStepLog bl;
  bl.l = l;
  bl.r = r;
  vector<StepLog> lg(n);
  vector<qtensor<T>> lenv(n), renv(n), psi(n);
  for (int c = 0; c < n; ++c) {
    lg[c].l = l;
    lg[c].r = r;
    lenv[c] = chains[c]->make_lenv(l);
    renv[c] = chains[c]->make_renv(r);
  }
  auto t1 = clk::now();
  solve(lenv, renv, psi, lg, bl);
  auto t2 = clk::now();
  vector<const qtensor<T>*> pp(n);
  for (int c = 0; c < n; ++c) pp[c] = &psi[c];
  vector<qtensor<T>> U, V;
  vector<qtensor<double>> S;
  vector<vector<double>> sv = svd_batched(pp, U, S, V, 2, cutoff);
  const int sweeps = block_svd_last_sweeps();
  const double tbs = block_svd_last_seconds();
  auto t3 = clk::now();
  for (int c = 0; c < n; ++c) {
    Engine<T>& e = *chains[c];
    e.mps_[l] = U[c];
    e.mps_.center(l + 1) = S[c];
    e.mps_[r] = V[c];
    qtensor<T> vec;
    vec.contract(lenv[c], U[c], 'N', 'N', 2);
    U[c].conjugate();
    e.mps_(l).contract(U[c], vec, 'T', 'N', 2);
    vec.contract(renv[c], V[c], 'N', 'T', 2);
    V[c].conjugate();
    e.mps_(r).contract(V[c], vec, 'N', 'N', 2);
  }
  auto t4 = clk::now();
  bl.t_env = secs(t0, t1) + secs(t3, t4);
  bl.t_eig = secs(t1, t2);
  bl.t_svd = secs(t2, t3);
  bl.t_total = secs(t0, t4);
  for (int c = 0; c < n; ++c) {
    // the step's wall time is shared by all chains; each chain logs its share
    lg[c].t_env = bl.t_env / n;
    lg[c].t_eig = bl.t_eig / n;
    lg[c].t_svd = bl.t_svd / n;
    lg[c].t_total = bl.t_total / n;
    lg[c].svd_sweeps = sweeps;
    lg[c].t_block_svd = tbs / n;
    chains[c]->finish(lg[c], sv[c], cutoff);
  }
  log.push_back(bl);
  return lg[0].e[0];
}

template <typename T>
double BatchEngine<T>::dmrg2(int cutoff, int sweep, bool ladder) {
  const int L = chains[0]->mps_.size();
  auto first = [&] {
    // the infinite-size start on the two edge sites: tiny, per chain
    auto t0 = clk::now();
    double e = 0;
    for (int c = (int)chains.size() - 1; c >= 0; --c) e = chains[c]->two_sites(0, L - 1, 0);
    BatchStepLog bl;
    bl.l = 0;
    bl.r = L - 1;
    bl.t_total = secs(t0, clk::now());
    for (auto& ch : chains) {
      bl.d += ch->log.back().d;
      bl.n_matvec = std::max(bl.n_matvec, ch->log.back().n_matvec);
      bl.flops += ch->log.back().flops;
    }
    log.push_back(bl);
    return e;
  };
  return schedule_dmrg2(L, cutoff, sweep, ladder, first, [&](int l, int r, int c) { return update_two(l, r, c); });
}

template class BatchEngine<double>;
template class BatchEngine<cplx>;

}  // namespace dmrgx
