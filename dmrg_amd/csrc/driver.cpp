// C entry points of the host mirror: full DMRG runs, synthetic operators for
// benchmarking, and structure-only get_map / generate_dim_sym.
#include <chrono>
#include <cstring>
#include <type_traits>

#include "dmrgx_driver.h"
#include "batch.hpp"
#include "core.hpp"
#include "heff.hpp"
#include "lanczos.hpp"
#include "model_4e.hpp"
#include "observables.hpp"
#include "output.hpp"
#include "qtensor.hpp"

using namespace dmrgx;

// a finished run's MPS (centre absorbed), one of r / c set
struct dmrgx_state {
  int model = 0, L = 0, sector = 0;
  double p[DMRGX_NPARAM] = {};  // the model parameters of the run
  bool signs = false;  // graded signs on: fermionic model, or the reference's sign rule
  bool pbc = false;    // the run carried periodic edge terms
  std::unique_ptr<mps<double>> r, raw_r;  // prepared (centre absorbed) and as left by the run
  std::unique_ptr<mps<cplx>> c, raw_c;
};

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

namespace {

bool model_complex(int model) {
  return model == DMRGX_MODEL_XYZ || model == DMRGX_MODEL_POTTS3 || model == DMRGX_MODEL_4E;
}
bool model_fermion(int model) { return model == DMRGX_MODEL_KITAEV || model == DMRGX_MODEL_4E; }
int model_symmetry(int model) {
  return model == DMRGX_MODEL_POTTS3 ? 3 : model == DMRGX_MODEL_4E ? SYM_S4E : 2;
}
// MPO tensors of a chain of L physical sites: the 4e model puts two sites in
// each (4e.cpp:46-47, which asks for L % 4 == 0)
int model_tensors(int model, int L) { return model == DMRGX_MODEL_4E ? L / 2 : L; }

// the model MPOs of src/operators.cpp with their edges (include/mpo.h:45-59);
// sets the charge group of the model
mpo<cplx> model_mpo_c(int model, const double* p, int L) {
  set_symmetry(model_symmetry(model));
  if (model == DMRGX_MODEL_4E) return mpo_4e(model_tensors(model, L), p);
  qtensor<cplx> H;
  if (model == DMRGX_MODEL_XYZ) H = H_XYZ(p[0], p[1], p[2]);
  else if (model == DMRGX_MODEL_POTTS3) H = H_Potts3(p[0], p[1], p[2], p[3]);
  else DMRGX_CHECK(false, "unknown complex model");
  mpo<cplx> m(L, H);
  m.edge();
  return m;
}

mpo<double> model_mpo_r(int model, const double* p, int L) {
  set_symmetry(model_symmetry(model));
  if (model == DMRGX_MODEL_KITAEV) {
    qtensor<double> H = H_Kitaev(p[0], p[1], p[2]);
    mpo<double> m(L, H);
    m.edge(H.left(), H.right());
    return m;
  }
  DMRGX_CHECK(model == DMRGX_MODEL_POTTS2, "unknown real model");
  qtensor<double> H = H_Potts2(p[0], p[1]);
  mpo<double> m(L, H);
  m.edge();
  return m;
}

// the periodic edge terms the reference's mains add (xyz.cpp:39-43,
// kitaev.cpp:43-46, potts.cpp:45-48): the two halves of the boundary term
template <typename T>
void add_model_edge(mps<T>& s, int model, const double* p);
template <>
void add_model_edge(mps<cplx>& s, int model, const double* p) {
  if (model == DMRGX_MODEL_4E) {
    qtensor<cplx> a, b;
    edges_4e(s.size(), p, a, b);
    s.add_edge(a, b);
  } else if (model == DMRGX_MODEL_XYZ) {
    qtensor<cplx> H = H_XYZ(p[0], p[1], p[2]);
    s.add_edge(H.left(), H.right());
  } else {
    DMRGX_CHECK(model == DMRGX_MODEL_POTTS3, "pbc: no edge terms for this model");
    qtensor<cplx> H = H_Potts3(0.0, p[1], p[2], p[3]);
    s.add_edge(H.left(), H.right());
  }
}
template <>
void add_model_edge(mps<double>& s, int model, const double* p) {
  DMRGX_CHECK(model == DMRGX_MODEL_KITAEV, "pbc: the reference has no periodic edge terms for this model");
  qtensor<double> H = H_Kitaev(p[0], -p[1], 0.0);
  s.add_edge(H.left(), H.right());
}

template <typename T>
mps<T> without_stores(const mps<T>& m) {
  mps<T> r;
  r.cut_ = m.cut_;
  r.state_ = m.state_;
  r.singular_ = m.singular_;
  r.store_.assign(m.state_.size(), qtensor<T>());
  return r;
}

// the MPO of H(p + dp) - H(p) (the builders are linear in the couplings)
template <typename T>
mpo<T> mpo_difference(mpo<T>& a, mpo<T>& b, int L) {
  qtensor<T> d;
  d.plus(a.operators_[0], b.operators_[0], T(1), T(-1));
  mpo<T> m(L, d);
  m.edge();
  return m;
}

// one row of DMRGX_STEP_LOG_WIDTH doubles per step, at most cap rows
void write_rows(const std::vector<StepLog>& lg, double* log, int cap) {
  int k = 0;
  for (auto& g : lg) {
    if (!log || k >= cap) break;
    double* row = log + (size_t)k * DMRGX_STEP_LOG_WIDTH;
    row[0] = g.l;
    row[1] = g.r;
    for (int i = 0; i < NEV; ++i) row[2 + i] = g.e[i];
    row[8] = g.trunc;
    row[9] = g.entropy;
    row[10] = g.n_matvec;
    row[11] = g.n_polish;
    row[12] = g.t_total;
    row[13] = g.t_eig;
    row[14] = g.t_svd;
    row[15] = g.t_env;
    row[16] = (double)g.d;
    row[17] = g.svd_sweeps;
    row[18] = g.t_block_svd;
    ++k;
  }
}

template <typename T>
void run_engine(Engine<T>& eng, const dmrgx_run_config* c, dmrgx_run_stats* st, double* log, int cap, int* n) {
  using clk = std::chrono::steady_clock;
  auto t0 = clk::now();
  // sweep boundaries: remember the step index at which each full sweep starts
  const int L = eng.mps_.size();
  double e = c->schedule == DMRGX_SCHEDULE_DMRG ? eng.dmrg(c->m, c->sweeps) : eng.dmrg2(c->m, c->sweeps, c->ladder != 0);
  auto t1 = clk::now();
  std::memset(st, 0, sizeof(*st));
  st->energy = e;
  st->seconds = std::chrono::duration<double>(t1 - t0).count();
  st->steps = (int)eng.log.size();
  for (auto& g : eng.log) {
    st->matvecs += g.n_matvec;
    st->polish_matvecs += g.n_polish;
    st->matvec_flops += g.flops;
    st->t_eig += g.t_eig;
    st->t_svd += g.t_svd;
    st->t_env += g.t_env;
  }
  if (!eng.log.empty()) {
    st->trunc_err = eng.log.back().trunc;
    st->entropy = eng.log.back().entropy;
  }
  // full sweeps are the last `sweeps` blocks of steps before the final half sweep
  if (c->sweeps > 0) {
    const int per = c->schedule == DMRGX_SCHEDULE_DMRG ? 2 * (L - 2) : 2 * (L - 4);
    const int fin = c->schedule == DMRGX_SCHEDULE_DMRG ? (L - 2) - (L / 2) + 1 : (L - 3) - (L / 2) + 1;
    int end = (int)eng.log.size() - std::max(fin, 0);
    int beg = end - per * c->sweeps;
    if (beg >= 0) {
      double tot = 0, last = 0;
      for (int i = beg; i < end; ++i) {
        tot += eng.log[i].t_total;
        if (i >= end - per) last += eng.log[i].t_total;
      }
      st->sweep_seconds = tot / c->sweeps;
      st->last_sweep_seconds = last;
    }
  }
  write_rows(eng.log, log, cap);
  if (n) *n = (int)eng.log.size();
}

template <typename T>
void run_batch(BatchEngine<T>& b, const dmrgx_run_config* c, double* energy, double* trunc, dmrgx_run_stats* st,
               double* log, int cap, int* n, double* chain_log) {
  using clk = std::chrono::steady_clock;
  auto t0 = clk::now();
  double e = b.dmrg2(c->m, c->sweeps, c->ladder != 0);
  auto t1 = clk::now();
  std::memset(st, 0, sizeof(*st));
  st->energy = e;
  st->seconds = std::chrono::duration<double>(t1 - t0).count();
  st->steps = (int)b.log.size();
  for (auto& g : b.log) {
    st->matvecs += g.n_matvec;
    st->polish_matvecs += g.n_polish;
    st->matvec_flops += g.flops;
    st->t_eig += g.t_eig;
    st->t_svd += g.t_svd;
    st->t_env += g.t_env;
  }
  const int nc = (int)b.chains.size();
  for (int i = 0; i < nc; ++i) {
    const StepLog& g = b.chains[i]->log.back();
    if (energy) energy[i] = g.e[0];
    if (trunc) trunc[i] = g.trunc;
  }
  if (c->sweeps > 0) {
    const int L = c->L, per = 2 * (L - 4), fin = (L - 3) - (L / 2) + 1;
    int end = (int)b.log.size() - std::max(fin, 0), beg = end - per * c->sweeps;
    if (beg >= 0) {
      double tot = 0, last = 0;
      for (int i = beg; i < end; ++i) {
        tot += b.log[i].t_total;
        if (i >= end - per) last += b.log[i].t_total;
      }
      st->sweep_seconds = tot / c->sweeps;
      st->last_sweep_seconds = last;
    }
  }
  int k = 0;
  for (auto& g : b.log) {
    if (!log || k >= cap) break;
    double* row = log + (size_t)k * DMRGX_BATCH_LOG_WIDTH;
    row[0] = g.l;
    row[1] = g.r;
    row[2] = (double)g.d;
    row[3] = g.n_matvec;
    row[4] = g.n_polish;
    row[5] = g.flops;
    row[6] = g.t_total;
    row[7] = g.t_eig;
    row[8] = g.t_svd;
    row[9] = g.t_env;
    row[10] = g.t_plan;
    ++k;
  }
  if (n) *n = (int)b.log.size();
  if (chain_log)
    for (int i = 0; i < nc; ++i)
      write_rows(b.chains[i]->log, chain_log + (size_t)i * cap * DMRGX_STEP_LOG_WIDTH, cap);
}

// Host-side checks of a run configuration, made before the device is touched.
// The single-site schedule under the reference's sign rule is refused for
// bosonic models: there get_map's alpha (src/qtensor.cpp:1078-1085) makes the
// move2right operator non-Hermitian (||H - H^H|| / ||H|| up to 0.6 on an open
// XYZ chain), which the reference hands to complex Arnoldi (znaupd, which="SR",
// src/znaupd.cpp:78-197) and this library's Hermitian Lanczos cannot solve.
void validate_cfg(const dmrgx_run_config* c) {
  DMRGX_CHECK(c != nullptr, "run: null configuration");
  DMRGX_CHECK(c->schedule == DMRGX_SCHEDULE_DMRG || c->schedule == DMRGX_SCHEDULE_DMRG2, "run: unknown schedule");
  DMRGX_CHECK(c->sign_rule == DMRGX_SIGNS_PHYSICAL || c->sign_rule == DMRGX_SIGNS_REFERENCE, "run: unknown sign rule");
  DMRGX_CHECK(c->prec >= PREC_FP64 && c->prec <= PREC_BF16X3, "run: unknown precision");
  DMRGX_CHECK(c->model >= DMRGX_MODEL_XYZ && c->model <= DMRGX_MODEL_4E, "run: unknown model");
  DMRGX_CHECK(c->pbc == 0 || c->model != DMRGX_MODEL_POTTS2, "run: the reference has no periodic edge terms for potts2");
  DMRGX_CHECK(c->model != DMRGX_MODEL_4E || (c->L % 4 == 0 && c->L >= 8),
              "run: the 4e chain needs L a multiple of 4 and at least 8 (two sites per MPO tensor)");
  DMRGX_CHECK(!(c->schedule == DMRGX_SCHEDULE_DMRG && c->sign_rule == DMRGX_SIGNS_REFERENCE && !model_fermion(c->model)),
              "run: the single-site schedule (dmrg) with the reference sign rule makes the operator of a bosonic "
              "model non-Hermitian; use schedule dmrg2 or the physical sign rule");
}

RunConfig make_cfg(const dmrgx_run_config* c, bool fermion) {
  RunConfig r;
  r.prec = c->prec;
  r.physical_signs = c->sign_rule == DMRGX_SIGNS_PHYSICAL;
  r.fermion = fermion;
  r.sector = c->sector;
  r.seed = c->seed;
  r.tol = c->tol > 0 ? c->tol : 1e-10;
  r.tol_lo = c->tol_lo > 0 ? c->tol_lo : 1e-6;
  r.refine_k = c->refine_k > 0 ? c->refine_k : 8;
  r.nev_lo = c->nev_lo > 0 ? c->nev_lo : 1;
  if (c->output) r.output_suffix = c->output;
  r.warm_start = c->warm_start != 0;
  r.verbose = c->verbose == 2 ? 2 : (c->verbose != 0 ? 1 : 0);
  return r;
}

// desc = [nleg, nsec[nleg], dims..., nblk, sym..., (dir...)]
struct Desc {
  vector<vector<int>> dim, sym;
  vector<int> dir;
};
Desc read_desc(const int32_t* d, bool with_dir) {
  Desc r;
  int p = 0;
  int nleg = d[p++];
  DMRGX_CHECK(nleg >= 0 && nleg <= 8, "desc: 0..8 legs");
  vector<int> ns(nleg);
  for (int i = 0; i < nleg; ++i) ns[i] = d[p++];
  r.dim.resize(nleg);
  for (int i = 0; i < nleg; ++i)
    for (int k = 0; k < ns[i]; ++k) r.dim[i].push_back(d[p++]);
  int nb = d[p++];
  r.sym.assign(nb, vector<int>(nleg));
  for (int b = 0; b < nb; ++b)
    for (int i = 0; i < nleg; ++i) r.sym[b][i] = d[p++];
  if (with_dir) {
    r.dir.resize(nleg);
    for (int i = 0; i < nleg; ++i) r.dir[i] = d[p++];
  }
  return r;
}
int write_desc(const vector<vector<int>>& dim, const vector<vector<int>>& sym, int32_t* out, int cap) {
  vector<int32_t> v;
  v.push_back((int)dim.size());
  for (auto& x : dim) v.push_back((int)x.size());
  for (auto& x : dim) v.insert(v.end(), x.begin(), x.end());
  v.push_back((int)sym.size());
  for (auto& s : sym) v.insert(v.end(), s.begin(), s.end());
  DMRGX_CHECK((int)v.size() <= cap, "description buffer too small");
  std::memcpy(out, v.data(), v.size() * 4);
  return (int)v.size();
}
qtensor<double> structure_only(const Desc& d) {
  qtensor<double> t;
  t.dim_ = d.dim;
  t.sym_ = d.sym;
  t.dir_ = d.dir.empty() ? vector<int>(d.dim.size(), 0) : d.dir;
  t.layout();
  return t;
}

}  // namespace

extern "C" {

int dmrgx_run_dmrg(const dmrgx_run_config* c, dmrgx_run_stats* st, double* step_log, int log_cap, int* n_steps) {
  return dmrgx_run_dmrg_state(c, st, step_log, log_cap, n_steps, nullptr);
}

int dmrgx_run_dmrg_state(const dmrgx_run_config* c, dmrgx_run_stats* st, double* step_log, int log_cap,
                         int* n_steps, dmrgx_state** state) {
  CAPI_BEGIN
  DMRGX_CHECK(c && st, "run_dmrg: null arguments");
  validate_cfg(c);
  DMRGX_CHECK(c->L >= 4, "run_dmrg: L must be at least 4");
  ensure_init();
  if (state) *state = nullptr;
  std::unique_ptr<dmrgx_state> out;
  if (state) {
    out = std::make_unique<dmrgx_state>();
    out->model = c->model;
    out->L = c->L;
    for (int i = 0; i < DMRGX_NPARAM; ++i) out->p[i] = c->p[i];
    out->pbc = c->pbc != 0;
    out->signs = model_fermion(c->model) || c->sign_rule == DMRGX_SIGNS_REFERENCE;
    out->sector = c->sector;
  }
  if (model_complex(c->model)) {
    mpo<cplx> m = model_mpo_c(c->model, c->p, c->L);
    Engine<cplx> eng(m, make_cfg(c, model_fermion(c->model)));
    if (c->pbc) add_model_edge(eng.mps_, c->model, c->p);
    run_engine(eng, c, st, step_log, log_cap, n_steps);
    if (out) {
      out->c = std::make_unique<mps<cplx>>(prepared(eng.mps_));
      out->raw_c = std::make_unique<mps<cplx>>(without_stores(eng.mps_));
    }
  } else {
    mpo<double> m = model_mpo_r(c->model, c->p, c->L);
    Engine<double> eng(m, make_cfg(c, model_fermion(c->model)));
    if (c->pbc) add_model_edge(eng.mps_, c->model, c->p);
    run_engine(eng, c, st, step_log, log_cap, n_steps);
    if (out) {
      out->r = std::make_unique<mps<double>>(prepared(eng.mps_));
      out->raw_r = std::make_unique<mps<double>>(without_stores(eng.mps_));
    }
  }
  if (state) *state = out.release();
  CAPI_END
}

int dmrgx_parse_args(int argc, const char* const* argv, int n_para, const char* const* names, int* L, int* m,
                     int* sweeps, int* sector, double* params, char* suffix, int cap) {
  CAPI_BEGIN
  DMRGX_CHECK(argc >= 0 && n_para >= 0 && (argc == 0 || argv) && (n_para == 0 || names), "parse_args: null arguments");
  std::vector<std::string> nm(names, names + n_para);
  Params p = set_para_val(argc, argv, nm);
  if (L) *L = p.sites;
  if (m) *m = p.cutoff;
  if (sweeps) *sweeps = p.sweep;
  if (sector) *sector = p.sector;
  if (params)
    for (int i = 0; i < n_para; ++i) params[i] = p.values[i];
  if (suffix) {
    DMRGX_CHECK((int)p.suffix.size() < cap, "parse_args: suffix buffer too small");
    std::memcpy(suffix, p.suffix.c_str(), p.suffix.size() + 1);
  }
  CAPI_END
}

int dmrgx_run_batch(const dmrgx_run_config* c, int n_chains, const double* p, double* energy, double* trunc,
                    dmrgx_run_stats* st, double* step_log, int log_cap, int* n_steps, double* chain_log) {
  CAPI_BEGIN
  DMRGX_CHECK(c && st && p && n_chains > 0, "run_batch: null arguments or no chains");
  validate_cfg(c);
  DMRGX_CHECK(c->L >= 6, "run_batch: L must be at least 6");
  DMRGX_CHECK(c->schedule == DMRGX_SCHEDULE_DMRG2, "run_batch: only the two-site schedule is batched");
  DMRGX_CHECK(c->warm_start == 0, "run_batch: warm starts are not batched (warm_start must be 0)");
  DMRGX_CHECK(c->pbc == 0, "run_batch: periodic chains are not batched (pbc must be 0)");
  ensure_init();
  const RunConfig cfg = make_cfg(c, model_fermion(c->model));
  if (model_complex(c->model)) {
    vector<mpo<cplx>> m;
    for (int i = 0; i < n_chains; ++i) m.push_back(model_mpo_c(c->model, p + 4 * i, c->L));
    BatchEngine<cplx> b(m, cfg);
    run_batch(b, c, energy, trunc, st, step_log, log_cap, n_steps, chain_log);
  } else {
    vector<mpo<double>> m;
    for (int i = 0; i < n_chains; ++i) m.push_back(model_mpo_r(c->model, p + 4 * i, c->L));
    BatchEngine<double> b(m, cfg);
    run_batch(b, c, energy, trunc, st, step_log, log_cap, n_steps, chain_log);
  }
  CAPI_END
}

int dmrgx_state_norm(const dmrgx_state* a, double* out) {
  CAPI_BEGIN
  DMRGX_CHECK(a && out, "state_norm: null arguments");
  set_symmetry(model_symmetry(a->model));
  if (a->c) {
    cplx v = calc(*a->c, a->signs);
    out[0] = v.real();
    out[1] = v.imag();
  } else {
    out[0] = calc(*a->r, a->signs);
    out[1] = 0;
  }
  CAPI_END
}

int dmrgx_state_overlap(const dmrgx_state* a, const dmrgx_state* b, double* out) {
  CAPI_BEGIN
  DMRGX_CHECK(a && b && out, "state_overlap: null arguments");
  DMRGX_CHECK(a->model == b->model && a->L == b->L, "state_overlap: states of different models or lengths");
  set_symmetry(model_symmetry(a->model));
  if (a->c) {
    cplx v = calc(*a->c, *b->c, a->signs);
    out[0] = v.real();
    out[1] = v.imag();
  } else {
    out[0] = calc(*a->r, *b->r, a->signs);
    out[1] = 0;
  }
  CAPI_END
}

int dmrgx_state_expect(const dmrgx_state* a, int model, const double* p, double* out) {
  CAPI_BEGIN
  DMRGX_CHECK(a && p && out, "state_expect: null arguments");
  DMRGX_CHECK(model_complex(model) == (a->c != nullptr) && model_symmetry(model) == model_symmetry(a->model),
              "state_expect: operator and state differ in element type or symmetry");
  if (a->c) {
    mpo<cplx> m = model_mpo_c(model, p, a->L);
    cplx v = calc(*a->c, m, a->signs);
    out[0] = v.real();
    out[1] = v.imag();
  } else {
    mpo<double> m = model_mpo_r(model, p, a->L);
    out[0] = calc(*a->r, m, a->signs);
    out[1] = 0;
  }
  CAPI_END
}

int dmrgx_state_correlations_4e(const dmrgx_state* a, double* rows, int cap, int* n) {
  CAPI_BEGIN
  DMRGX_CHECK(a && n && (rows || cap == 0), "state_correlations_4e: null arguments");
  DMRGX_CHECK(a->model == DMRGX_MODEL_4E && a->c, "state_correlations_4e: not a state of the 4e model");
  set_symmetry(model_symmetry(a->model));
  vector<Corr4e> c = correlations_4e(*a->c, true);
  *n = 0;
  for (auto& r : c) {
    if (*n >= cap) break;
    rows[3 * *n] = r.x;
    rows[3 * *n + 1] = r.value.real();
    rows[3 * *n + 2] = r.value.imag();
    ++*n;
  }
  CAPI_END
}

int dmrgx_state_position(const dmrgx_state* a, int* cut) {
  CAPI_BEGIN
  DMRGX_CHECK(a && cut, "state_position: null arguments");
  DMRGX_CHECK(a->raw_c || a->raw_r, "state_position: the state holds no MPS");
  *cut = a->raw_c ? a->raw_c->position() : a->raw_r->position();
  CAPI_END
}

int dmrgx_state_move_cut(dmrgx_state* a, int steps) {
  CAPI_BEGIN
  DMRGX_CHECK(a, "state_move_cut: null state");
  DMRGX_CHECK(a->raw_c || a->raw_r, "state_move_cut: the state holds no MPS");
  set_symmetry(model_symmetry(a->model));
  auto move = [&](auto& raw, auto& prep) {
    const int n = raw->size(), c = raw->position();
    DMRGX_CHECK(c + steps >= 1 && c + steps <= n - 1,
                "state_move_cut: the centre would leave bonds 1.." + std::to_string(n - 1));
    for (int i = 0; i < steps; ++i) raw->move_cut_right();
    for (int i = 0; i < -steps; ++i) raw->move_cut_left();
    prep = std::make_unique<std::remove_reference_t<decltype(*raw)>>(prepared(*raw));
  };
  if (a->raw_c) move(a->raw_c, a->c);
  else move(a->raw_r, a->r);
  CAPI_END
}

int dmrgx_state_inverse(const dmrgx_state* a, dmrgx_state** out) {
  CAPI_BEGIN
  DMRGX_CHECK(a && out, "state_inverse: null arguments");
  DMRGX_CHECK(a->raw_c || a->raw_r, "state_inverse: the state holds no MPS");
  set_symmetry(model_symmetry(a->model));
  auto r = std::make_unique<dmrgx_state>();
  r->model = a->model;
  r->L = a->L;
  r->sector = a->sector;
  r->signs = a->signs;
  r->pbc = a->pbc;
  for (int i = 0; i < DMRGX_NPARAM; ++i) r->p[i] = a->p[i];
  if (a->raw_c) {
    r->raw_c = std::make_unique<mps<cplx>>(a->raw_c->inverse());
    r->c = std::make_unique<mps<cplx>>(prepared(*r->raw_c));
  } else {
    r->raw_r = std::make_unique<mps<double>>(a->raw_r->inverse());
    r->r = std::make_unique<mps<double>>(prepared(*r->raw_r));
  }
  *out = r.release();
  CAPI_END
}

int dmrgx_state_derivative(const dmrgx_state* a, const double* p, const double* dp, int prec, double* out) {
  CAPI_BEGIN
  DMRGX_CHECK(a && p && dp && out, "state_derivative: null arguments");
  DMRGX_CHECK(prec == DMRGX_FP64 || prec == DMRGX_FP32, "state_derivative: prec must be FP64 or FP32");
  DMRGX_CHECK(!a->pbc, "state_derivative: the environments of a periodic run carry edge terms this call omits");
  DMRGX_CHECK(a->model != DMRGX_MODEL_4E, "state_derivative: the 4e MPO is not one tensor repeated");
  double q[DMRGX_NPARAM];
  for (int i = 0; i < DMRGX_NPARAM; ++i) q[i] = p[i] + dp[i];
  double e = 0;
  if (a->raw_c) {
    mpo<cplx> W = model_mpo_c(a->model, p, a->L), Wq = model_mpo_c(a->model, q, a->L);
    mpo<cplx> dW = mpo_difference(Wq, W, a->L);
    out[0] = centre_derivative(*a->raw_c, W, dW, a->sector, a->signs, prec, &e);
  } else {
    mpo<double> W = model_mpo_r(a->model, p, a->L), Wq = model_mpo_r(a->model, q, a->L);
    mpo<double> dW = mpo_difference(Wq, W, a->L);
    out[0] = centre_derivative(*a->raw_r, W, dW, a->sector, a->signs, prec, &e);
  }
  out[1] = e;
  CAPI_END
}

int dmrgx_state_heff_create(const dmrgx_state* a, int prec, dmrgx_heff** plan, int64_t* n_psi, const void** psi) {
  CAPI_BEGIN
  DMRGX_CHECK(a && plan && n_psi, "state_heff_create: null arguments");
  DMRGX_CHECK(prec >= PREC_FP64 && prec <= PREC_BF16X3, "state_heff_create: unknown precision");
  DMRGX_CHECK(a->raw_c || a->raw_r, "state_heff_create: the state holds no MPS");
  DMRGX_CHECK(!a->pbc, "state_heff_create: the environments of a periodic run carry edge terms this call omits");
  ensure_init();
  auto h = std::make_unique<dmrgx_heff>();
  auto take = [&](auto& c) {
    *n_psi = get_dimension(c.psi.dim_, c.psi.sym_);
    h->p = std::move(c.plan);
    h->keep.push_back(c.lenv.buf_);
    h->keep.push_back(c.renv.buf_);
    h->keep.push_back(c.psi.buf_);
    if (psi) *psi = c.psi.data();
  };
  if (a->raw_c) {
    mpo<cplx> W = model_mpo_c(a->model, a->p, a->L);
    auto c = centre_operator(*a->raw_c, W, a->sector, a->signs, prec);
    take(c);
  } else {
    mpo<double> W = model_mpo_r(a->model, a->p, a->L);
    auto c = centre_operator(*a->raw_r, W, a->sector, a->signs, prec);
    take(c);
  }
  DMRGX_HIP(hipStreamSynchronize(rt().stream));
  *plan = h.release();
  CAPI_END
}

int dmrgx_state_destroy(dmrgx_state* a) {
  CAPI_BEGIN
  delete a;
  CAPI_END
}

int dmrgx_synth_heff_create(int S, int D, const int* bond, int n_sec, int dtype, int prec, uint64_t seed, int fermion,
                            dmrgx_heff** plan, int64_t* n_psi) {
  CAPI_BEGIN
  ensure_init();
  DMRGX_CHECK(S >= 1 && n_sec == S && D >= 1, "synth_heff: need one bond dimension per sector");
  set_symmetry(S);
  auto build = [&](bool right, auto tag) {
    using T = decltype(tag);
    qtensor<T> t;
    vector<int> phys(S, 1), bd(bond, bond + n_sec);
    if (right) t.dim_ = {phys, bd, {D}, phys, bd};
    else t.dim_ = {bd, phys, {D}, bd, phys};
    t.dir_ = right ? vector<int>{0, -1, 0, 0, 1} : vector<int>{-1, 0, 0, 1, 0};
    for (int a = 0; a < S; ++a)
      for (int b = 0; b < S; ++b)
        for (int c = 0; c < S; ++c)
          for (int e = 0; e < S; ++e) t.sym_.push_back({a, b, 0, c, e});
    t.layout();
    t.allocate();
    dev_random((double*)t.data(), t.total_, dtype == DT_C64, seed + (right ? 17 : 29), rt().stream);
    return t;
  };
  auto make = [&](auto tag) {
    using T = decltype(tag);
    qtensor<T> R = build(true, tag), Lq = build(false, tag);
    vector<vector<int>> dim, sym, rmap, lmap, dim_ret, sym_ret, dim2;
    generate_dim_sym(dim, sym, Lq, 2, R, 2, 0);
    int64_t d = get_dimension(dim, sym);
    R.get_map(rmap, dim_ret, sym_ret, dim, sym, 'N', 'T', 2, 0, false, fermion != 0);
    int64_t ds = get_dimension(dim_ret, sym_ret);
    Lq.get_map(lmap, dim2, sym, dim_ret, sym_ret, 'N', 'T', 3, 0, true, fermion != 0);
    vector<int32_t> rm, lm;
    for (auto& x : rmap) rm.insert(rm.end(), x.begin(), x.end());
    for (auto& x : lmap) lm.insert(lm.end(), x.begin(), x.end());
    vector<int64_t> rs(R.nblocks()), ls(Lq.nblocks());
    for (int b = 0; b < R.nblocks(); ++b) rs[b] = R.bsize(b);
    for (int b = 0; b < Lq.nblocks(); ++b) ls[b] = Lq.bsize(b);
    auto h = new dmrgx_heff;
    h->p = make_heff(dtype, prec, R.data(), R.off_.data(), rs.data(), R.nblocks(), R.total_, rm.data(),
                     (int)rmap.size(), Lq.data(), Lq.off_.data(), ls.data(), Lq.nblocks(), Lq.total_, lm.data(),
                     (int)lmap.size(), d, ds);
    h->keep.push_back(R.buf_);
    h->keep.push_back(Lq.buf_);
    *plan = h;
    *n_psi = d;
  };
  if (dtype == DT_C64) make(cplx());
  else make(double());
  CAPI_END
}

int dmrgx_get_map(int sym_n, const int32_t* tdesc, const int32_t* adesc, char ta, char tb, int num, int beta, int back,
                  int fermion, const int32_t* symret_in, int32_t* map_out, int map_cap, int* n_map, int32_t* ret_desc,
                  int ret_cap) {
  CAPI_BEGIN
  set_symmetry(sym_n);
  Desc t = read_desc(tdesc, true), a = read_desc(adesc, false);
  qtensor<double> q = structure_only(t);
  vector<vector<int>> map, dim_ret, sym_ret;
  if (back) {
    int nb = symret_in[0];
    int nleg = (int)(t.dim.size() + a.dim.size()) - 2 * num;
    sym_ret.assign(nb, vector<int>(nleg));
    for (int b = 0; b < nb; ++b)
      for (int i = 0; i < nleg; ++i) sym_ret[b][i] = symret_in[1 + b * nleg + i];
  }
  DMRGX_CHECK(q.get_map(map, dim_ret, sym_ret, a.dim, a.sym, ta, tb, num, beta, back != 0, fermion != 0),
              "get_map: dimensions differ");
  DMRGX_CHECK((int)map.size() <= map_cap, "get_map: map buffer too small");
  for (size_t i = 0; i < map.size(); ++i)
    for (int k = 0; k < 7; ++k) map_out[i * 7 + k] = map[i][k];
  *n_map = (int)map.size();
  write_desc(dim_ret, sym_ret, ret_desc, ret_cap);
  CAPI_END
}

int dmrgx_generate_dim_sym(int sym_n, const int32_t* ldesc, int lnum, const int32_t* rdesc, int rnum, int sector,
                           int32_t* out_desc, int cap) {
  CAPI_BEGIN
  set_symmetry(sym_n);
  qtensor<double> l = structure_only(read_desc(ldesc, true)), r = structure_only(read_desc(rdesc, true));
  vector<vector<int>> dim, sym;
  generate_dim_sym(dim, sym, l, lnum, r, rnum, sector);
  write_desc(dim, sym, out_desc, cap);
  CAPI_END
}

}  // extern "C"
