/*
 * dmrgx driver API: the C++ host mirror of the reference's sweep drivers
 * (dmrg / dmrg2, src/functions.cpp:157-299) and step functions
 * (src/core.cpp:88-474) running every numerical kernel through dmrgx.h, plus
 * the integer bookkeeping entry points used to check it bit for bit.
 * Same conventions as dmrgx.h (0 = success, dmrgx_last_error()).
 */
#ifndef DMRGX_DRIVER_H
#define DMRGX_DRIVER_H

#include <stdint.h>

#include "dmrgx.h"

#ifdef __cplusplus
extern "C" {
#endif

enum { DMRGX_MODEL_XYZ = 0, DMRGX_MODEL_KITAEV = 1, DMRGX_MODEL_POTTS3 = 2, DMRGX_MODEL_POTTS2 = 3, DMRGX_MODEL_4E = 4 };
/* model parameters carried by a configuration and by a finished state */
#define DMRGX_NPARAM 8
enum { DMRGX_SCHEDULE_DMRG = 0, DMRGX_SCHEDULE_DMRG2 = 1 };
enum { DMRGX_SIGNS_PHYSICAL = 0, DMRGX_SIGNS_REFERENCE = 1 };

typedef struct dmrgx_run_config {
  int model;        /* DMRGX_MODEL_* */
  double p[DMRGX_NPARAM];  /* xyz: jx jy jz; kitaev: t p u; potts3: f j p t; potts2: h j;
                              4e: t1 t3 f0 f1 frac (4e.cpp:39); unused: 0 */
  int L, m, sweeps, sector;   /* the reference's l= m= n= s= (src/functions.cpp:46-66); for
                                 4e, L physical sites (a multiple of 4) in L/2 MPO tensors
                                 (4e.cpp:46-47) and a Z2 x Z2 x Z2 sector 0..7 */
  int schedule;     /* DMRGX_SCHEDULE_DMRG (the mains' dmrg) or _DMRG2 (two-site) */
  int prec;         /* matvec arithmetic: DMRGX_FP64 / DMRGX_FP32 / DMRGX_BF16X3 */
  int sign_rule;    /* DMRGX_SIGNS_PHYSICAL (alpha=1 for bosonic models) or _REFERENCE */
  int ladder;       /* 1: pre-sweep cutoff ladder 100,200,.. as the reference; 0: skip */
  uint64_t seed;    /* Lanczos start vectors */
  double tol;       /* FP64 Lanczos tolerance (reference: 1e-10) */
  double tol_lo;    /* reduced-precision Lanczos tolerance */
  int refine_k;     /* FP64 Krylov polish steps after a reduced-precision solve */
  int verbose;      /* 0: quiet; 1: one line per step on stdout; 2: the same on stderr */
  int nev_lo;       /* eigenpairs whose residual a reduced-precision solve tests (0: 1; its basis is sized
                       for four and its four lowest Ritz vectors start the FP64 refinement); FP64 solves use 6 */
  const char* output;  /* NULL: no files; else every step appends to "energy" + output and
                          "singular" + output in the working directory in the reference's
                          formats (print_energy / print_singular, src/useful.cpp:54-102) */
  int warm_start;   /* 0: every two-site eigensolve starts from a seeded random vector, as the
                       reference's ARPACK calls (info=0, src/dsaupd.cpp:97) and converges the
                       reference's six pairs; 1: a sweep step starts from the state the previous
                       step left (White's prediction) and converges the lowest pair only
                       (E1..E5 read 0).  dmrgx_run_batch accepts 0 only. */
  int pbc;          /* 1: periodic chain, the reference's PBC build: the mains' add_edge terms
                       (xyz.cpp:39-43, kitaev.cpp:43-46, potts.cpp:45-48) carried by edge
                       environments, H_eff = bulk + L_edge W_id W_id R_edge (src/core.cpp:27-30);
                       not for potts2, dmrgx_run_batch or the derivative / centre-operator calls */
} dmrgx_run_config;

typedef struct dmrgx_run_stats {
  double energy;          /* value returned by the last step */
  double seconds;         /* whole run */
  double sweep_seconds;   /* mean wall seconds of the full sweeps (0 if none) */
  double last_sweep_seconds;
  int steps;
  long long matvecs, polish_matvecs;
  double matvec_flops;    /* algorithmic flops of all matvecs */
  double trunc_err, entropy;   /* of the last step */
  double t_eig, t_svd, t_env;  /* summed step phases, seconds */
} dmrgx_run_stats;

/* One row per step: l, r, E0..E5, trunc, entropy, matvecs, polish matvecs,
 * t_total, t_eig, t_svd, t_env, d, SVD Jacobi sweeps, t_block_svd (19
 * doubles; t_block_svd is the blockwise SVD alone, inside t_svd). */
#define DMRGX_STEP_LOG_WIDTH 19
int dmrgx_run_dmrg(const dmrgx_run_config* cfg, dmrgx_run_stats* stats, double* step_log, int log_cap,
                   int* n_steps);

/* set_para_val (src/functions.cpp:46-80): reads "key=value" words (l, m, n, s
 * and the model's n_para parameter names; defaults l=4, m=10, n=-1, s=0, the
 * first parameter 1 and the others 0) and writes the output file suffix
 * "_L_<l>_m_<m>_n_<n>_s_<s>_<name>_<value>..._.txt" (values in the default
 * stream format) into suffix (cap bytes).  Host only: no device is touched. */
int dmrgx_parse_args(int argc, const char* const* argv, int n_para, const char* const* names, int* L, int* m,
                     int* sweeps, int* sector, double* params, char* suffix, int cap);

/* ---- many chains in lockstep ----------------------------------------------------
 * dmrgx_run_batch sweeps n_chains independent chains of cfg->model (couplings
 * p[4 i .. 4 i + 3] for chain i; cfg->p is ignored) with the two-site schedule
 * (cfg->schedule must be DMRGX_SCHEDULE_DMRG2), every step of every chain in
 * lockstep: one H_eff plan over all chains' get_map tasks (one grouped launch
 * per pass), a segmented batched Lanczos over the concatenated two-site
 * states, one block SVD over all chains' blocks.  The reference has no
 * counterpart (one chain per process, src/functions.cpp:234-299 per chain);
 * each chain's result is that of dmrgx_run_dmrg with its couplings, up to
 * rounding and to the extra Lanczos steps a chain takes while others converge.
 * energy[i], trunc[i]: chain i's last E0 and truncation error.  stats sum the
 * batch (matvecs = batched matvec calls, matvec_flops = all chains).  One row
 * per batch step: l, r, summed dimension, matvec calls, polish calls, flops,
 * t_total, t_eig, t_svd, t_env, t_plan (plan construction, inside t_eig).
 * chain_log (may be NULL): n_chains x log_cap rows of DMRGX_STEP_LOG_WIDTH,
 * each chain's own step log as dmrgx_run_dmrg writes it (its times are its
 * 1/n_chains share of the batch step). */
#define DMRGX_BATCH_LOG_WIDTH 11
int dmrgx_run_batch(const dmrgx_run_config* cfg, int n_chains, const double* p, double* energy, double* trunc,
                    dmrgx_run_stats* stats, double* step_log, int log_cap, int* n_steps, double* chain_log);

/* ---- finished states and observables ------------------------------------------
 * dmrgx_run_dmrg_state runs like dmrgx_run_dmrg and also returns the final MPS
 * (device-resident, the centre absorbed as mps::prep_calc does,
 * src/mps.cpp:33-54).  The observables are the reference's calc functions
 * (src/functions.cpp:394-511), out = {re, im}:
 *   dmrgx_state_norm:    calc(mps)        = <a|a>
 *   dmrgx_state_overlap: calc(mps, mps)   = <b|a>  (same model, chain length)
 *   dmrgx_state_expect:  calc(mps, mpo)   = <a|H(model, p)|a>, H built by the
 *                        model builder with couplings p[DMRGX_NPARAM] (same
 *                        symmetry as a; open-chain MPO, no edge terms);
 *                        linear builders give Hellmann-Feynman derivatives,
 *                        e.g. XYZ with p = {0, 0, 1} is dH/dJz. */
typedef struct dmrgx_state dmrgx_state;
int dmrgx_run_dmrg_state(const dmrgx_run_config* cfg, dmrgx_run_stats* stats, double* step_log, int log_cap,
                         int* n_steps, dmrgx_state** state);
int dmrgx_state_norm(const dmrgx_state* a, double* out);
int dmrgx_state_overlap(const dmrgx_state* a, const dmrgx_state* b, double* out);
int dmrgx_state_expect(const dmrgx_state* a, int model, const double* p, double* out);
/* dmrgx_state_derivative: d E0 / dt along the coupling direction dp at the
 * couplings p (the state's own), from the adjoint H_eff kernels at the
 * state's centre: the environments and their tangents along the MPO of
 * H(p + dp) - H(p) are recomputed from the MPS site by site, and
 * dE0/dt = Re(<g_L, dL> + <g_R, dR>) with g_L, g_R the environment
 * cotangents for the output cotangent psi (Hellmann-Feynman at fixed MPS
 * basis).  out = {dE0/dt, E0 of the centre}; prec DMRGX_FP64 or DMRGX_FP32.
 * p and dp hold DMRGX_NPARAM values.  States of periodic runs are refused. */
int dmrgx_state_derivative(const dmrgx_state* a, const double* p, const double* dp, int prec, double* out);
/* The two-site H_eff at the centre bond of a finished run's state (the
 * operator its last step solved, rebuilt): both environments recomputed from
 * the MPS site by site (src/core.cpp:188-192, 266-271), the get_map task lists
 * (src/qtensor.cpp:997-1119) and a plan of arithmetic `prec` that owns the
 * environments.  *psi (may be NULL) receives the device address of the
 * state's two-site vector in the plan's psi layout (n_psi elements, owned by
 * the plan).  No counterpart in the reference beyond av() itself. */
int dmrgx_state_heff_create(const dmrgx_state* a, int prec, dmrgx_heff** plan, int64_t* n_psi, const void** psi);
/* mps::position, mps::move_cut_left / move_cut_right (repeated |steps| times;
 * steps > 0 moves right) and mps::inverse (src/mps.cpp:56-128, include/mps.h:55-74)
 * on the state as the run left it; the observables above then use the moved
 * or reversed state.  The centre stays on bonds 1..L-1 (an error otherwise;
 * the reference prints one and returns).  dmrgx_state_inverse returns a new
 * state, released by dmrgx_state_destroy. */
int dmrgx_state_position(const dmrgx_state* a, int* cut);
int dmrgx_state_move_cut(dmrgx_state* a, int steps);
int dmrgx_state_inverse(const dmrgx_state* a, dmrgx_state** out);
/* The correlation functions the 4e main writes to its data file
 * (4e.cpp:71-196), for a state of the 4e model: the end-to-end pair, c_up and
 * c_down correlations (L1 R1 .. L2 R2 for each), the normalization calc(mps),
 * then the Cooper-pair expectation on every interior tensor (positions 2q+1,
 * 2q+2) and the pair-pair ones (2q+1.5, 2q+2.5).  Row i of rows = {x, re, im}
 * (x: the number of tensors for the first 13 lines, else the position), as
 * the main prints them; *n rows are written (at most cap). */
int dmrgx_state_correlations_4e(const dmrgx_state* a, double* rows, int cap, int* n);
int dmrgx_state_destroy(dmrgx_state* a);

/* Synthetic two-site operator with the block structure of a Z_S model at bond
 * sectors `bond` (n_sec entries), MPO bond D, physical sectors of dimension 1,
 * every sector combination present, uniform(-1,1) entries from `seed`.  The
 * plan owns its environments. */
int dmrgx_synth_heff_create(int S, int D, const int* bond, int n_sec, int dtype, int prec, uint64_t seed,
                            int fermion, dmrgx_heff** plan, int64_t* n_psi);

/* `symmetry` below: N >= 1 for the Z_N group of the reference's SYMMETRY N
 * build, or DMRGX_SYM_S4E for the Z2 x Z2 x Z2 group of its S4E build
 * (eight sectors, fusion by XOR and the fermion-parity table of
 * src/useful.cpp:4-27). */
#define DMRGX_SYM_S4E (-8)

/* qtensor<T>::get_map on structure only (src/qtensor.cpp:997-1119).
 * Tensor / array descriptions: desc = [nleg, nsec[nleg], dims (per leg, per
 * sector), nblk, sym (nblk x nleg), dir (nleg, tensor only)].  `symret_in`
 * ([nblk, sym...]) is read when back != 0.  Writes n_map tuples of 7 int32
 * and the result description (dims + syms, no dirs) into ret_desc. */
int dmrgx_get_map(int symmetry, const int32_t* tensor_desc, const int32_t* array_desc, char ta, char tb, int num,
                  int beta, int back, int fermion, const int32_t* symret_in, int32_t* map_out, int map_cap,
                  int* n_map, int32_t* ret_desc, int ret_cap);

/* generate_dim_sym (src/useful.cpp:151-194) on structure only; lenv / renv
 * descriptions as above (dirs included); out_desc as get_map's ret_desc. */
int dmrgx_generate_dim_sym(int symmetry, const int32_t* lenv_desc, int lnum, const int32_t* renv_desc, int rnum,
                           int sector, int32_t* out_desc, int cap);

#ifdef __cplusplus
}
#endif

#endif /* DMRGX_DRIVER_H */
