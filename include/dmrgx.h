/*
 * dmrgx -- MI355X (gfx950) implementation of the two-site DMRG inner loop of
 * clover199/DMRG: the effective-Hamiltonian matvec over get_map task lists,
 * its eigensolver, the blockwise SVD truncation and the environment updates.
 *
 * C ABI.  Conventions (as in the reference's numerics wrappers, include/basic.h):
 *   - element types: DMRGX_R64 = double, DMRGX_C64 = complex double stored as
 *     interleaved (re, im) pairs, i.e. the layout of std::complex<double>;
 *   - all matrices are row-major; block data are contiguous in block order;
 *   - "device" pointers come from dmrgx_malloc (or any hipMalloc'd memory of
 *     the device passed to dmrgx_init);
 *   - every function returns 0 on success and a negative code on failure;
 *     dmrgx_last_error() describes the failure (per host thread);
 *   - one device per process, one in-order stream; calls are not re-entrant
 *     on one plan;
 *   - a call that writes device memory returns with that work complete, so
 *     results may be read at once from any stream (dmrgx_stream() gives the
 *     library's stream to callers that order their own work after it).
 */
#ifndef DMRGX_H
#define DMRGX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { DMRGX_R64 = 0, DMRGX_C64 = 1 };
enum { DMRGX_FP64 = 0, DMRGX_FP32 = 1, DMRGX_BF16X3 = 2 };

/* ---- runtime ------------------------------------------------------------ */
int dmrgx_init(int device);
int dmrgx_finalize(void);
const char* dmrgx_last_error(void);
int dmrgx_device_info(int* num_cu, char* arch_name, int name_len);
int dmrgx_malloc(size_t bytes, void** ptr);
int dmrgx_free(void* ptr);
int dmrgx_memcpy_h2d(void* dst_dev, const void* src_host, size_t bytes);
int dmrgx_memcpy_d2h(void* dst_host, const void* src_dev, size_t bytes);
int dmrgx_memset_zero(void* dst_dev, size_t bytes);
int dmrgx_synchronize(void);
/* the library's HIP stream (hipStream_t), for callers that order their own work */
void* dmrgx_stream(void);

/* ---- task-list GEMM ------------------------------------------------------
 * Replaces qtensor<T>::contract(T* out, T* in, map, 'N', 'T', num)
 * (src/qtensor.cpp:986-994 -> src/tensor.cpp:698-718 -> src/dgemm.cpp:29-47,
 * src/zgemm.cpp:32-51).  `map` holds n_task get_map tuples of 7 int32
 * (block, in_beg, out_beg, row, col, beta, alpha) (src/qtensor.cpp:1046).
 * For task t: out[out_beg..] (m x row) = alpha * A_blk (m x col) *
 * in[in_beg..] (row x col)^T + beta * out, m = blk_size[blk] / col.
 * blk_off / blk_size (host) give each block of A in elements.  A, in, out are
 * device pointers of element type `dtype`; `prec` selects the arithmetic
 * (FP64 operands are converted on the fly for FP32 / BF16X3). */
int dmrgx_task_gemm(int dtype, int prec, const void* A, const int64_t* blk_off,
                    const int64_t* blk_size, int n_blk, const int32_t* map, int n_task,
                    const void* in, void* out, int64_t n_out);

/* dmrgx_task_gemm for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  FP64 arithmetic; same arguments and
 * results as dmrgx_task_gemm with prec = DMRGX_FP64 (outputs without a task
 * are left untouched).  It allocates nothing: the work decomposition and the
 * split-K partial tiles go to a fixed 64 MiB device array per device, set up
 * by dmrgx_init (call it first: initialisation allocates).  The split-K
 * decomposition is the default one when it fits, else one with fewer K cuts;
 * a task list whose tables alone exceed 64 MiB is an error.  Every call on a
 * device shares that array, so calls that overlap on one device (several host
 * threads) race; calls return with their work complete. */
int dmrgx_task_gemm_noalloc(int dtype, const void* A, const int64_t* blk_off,
                            const int64_t* blk_size, int n_blk, const int32_t* map, int n_task,
                            const void* in, void* out, int64_t n_out);

/* ---- effective Hamiltonian ------------------------------------------------
 * Replaces av(in, out, lanczos<T>&) (src/core.cpp:22-31): store = R * in over
 * rmap (r_num contracted legs), out = L * store over lmap.  renv / lenv are
 * device pointers (FP64 storage, kept by reference: they must outlive the
 * plan).  n_psi, n_store in elements.  Output blocks of `out` that no lmap
 * task writes read as zero. */
typedef struct dmrgx_heff dmrgx_heff;
int dmrgx_heff_create(dmrgx_heff** plan, int dtype, int prec,
                      const void* renv, const int64_t* rblk_off, const int64_t* rblk_size,
                      int n_rblk, const int32_t* rmap, int n_rtask,
                      const void* lenv, const int64_t* lblk_off, const int64_t* lblk_size,
                      int n_lblk, const int32_t* lmap, int n_ltask,
                      int64_t n_psi, int64_t n_store);
int dmrgx_heff_apply(dmrgx_heff* plan, const void* in, void* out);
/* Adjoint of out = H_eff(psi) with respect to psi and the two environments
 * (no counterpart in the reference, which has no derivatives): for a
 * cotangent g_out of out, writes g_psi (n_psi), g_renv and g_lenv (the
 * environments' lengths, same block layout) such that
 * Re<g_out, d out> = Re<g_psi, d psi> + Re<g_renv, d renv> + Re<g_lenv, d lenv>
 * (<a, b> = sum conj(a) b).  Any of g_psi / g_renv / g_lenv may be NULL.
 * FP64 for FP64 plans, FP32 MFMA for FP32 / BF16X3 plans; all buffers FP64
 * (complex interleaved), device pointers. */
int dmrgx_heff_adjoint(dmrgx_heff* plan, const void* psi, const void* g_out, void* g_psi, void* g_renv,
                       void* g_lenv);
/* {n_psi, n_store, renv length, lenv length} in elements */
int dmrgx_heff_sizes(const dmrgx_heff* plan, int64_t* out);
/* algorithmic flops of one apply: sum over tasks of c*M*N*K (c = 2 real, 8 complex),
 * every task in full as the reference's BLAS calls multiply it */
double dmrgx_heff_flops(const dmrgx_heff* plan);
/* flops the plan executes per apply: the algorithmic count without the
 * environments' structurally zero slices (see dmrgx_set_skip_zero_slices) */
double dmrgx_heff_flops_exec(const dmrgx_heff* plan);
/* Whether plans created from now on skip the environments' structurally zero
 * slices (on = 1, the default).  No counterpart in the reference, which
 * multiplies them: the MPO bond index w of an environment block carries only
 * the operators its symmetry sectors allow, so most (block, w) slices are
 * exact zeros (about 70 % of the flops of an XYZ step, 83 % of a Z3 Potts
 * step).  A plan scans its environments once at creation: R-pass rows (s, b,
 * w) that are zero in every task of an output block are not computed (their
 * `store` elements are exact zeros), and each L-pass task contracts only the
 * K columns (w, a', s') that are nonzero in the environment and computed in
 * `store`.  Results are the full sums with zero terms left out. */
int dmrgx_set_skip_zero_slices(int on);
int dmrgx_heff_destroy(dmrgx_heff* plan);
/* Run `iters` applies on the library stream with HIP events around each GEMM
 * launch: ms[0] = mean ms per apply, ms[1] / ms[2] = mean ms of the R / L pass
 * GEMM kernel, ms[3] / ms[4] = algorithmic flops of the R / L pass, ms[5] /
 * ms[6] = flops the R / L pass executes (7 doubles). */
int dmrgx_heff_time(dmrgx_heff* plan, const void* in, void* out, int iters, double* ms);
/* One apply of `in`, then `iters` back-to-back launches of the R-pass GEMM
 * kernel alone and of the L-pass GEMM kernel alone between HIP events (no
 * zero fills or split-K reductions): ms[0] / ms[1] = ms per launch of the R /
 * L kernel (2 doubles); `out` holds H_eff in on return.  A measurement of
 * this build (the reference has no GPU kernels). */
int dmrgx_heff_time_kernels(dmrgx_heff* plan, const void* in, void* out, int iters, double* ms);
/* dmrgx_heff_adjoint run `iters` times between two HIP events on the library's
 * stream: ms[0] = event time per call (no counterpart in the reference, which
 * has no derivatives) */
int dmrgx_heff_time_adjoint(dmrgx_heff* plan, const void* psi, const void* g_out, void* g_psi, void* g_renv,
                            void* g_lenv, int iters, double* ms);

/* ---- eigensolver -------------------------------------------------------------
 * Replaces znaupd(n, nev, Evals, Evecs, lanczos<T>&) (src/dsaupd.cpp:65-155,
 * src/znaupd.cpp:78-197): the nev lowest eigenvalues (ascending) of the
 * Hermitian operator of `plan` and the eigenvector of the lowest.  FP64
 * thick-restart Lanczos, full reorthogonalisation, ARPACK convergence test
 * |beta*y_last| <= tol*max(|theta|, eps^(2/3)).  evec (device, n_psi elements)
 * receives the vector; v0 (device, may be NULL) is the start vector, else a
 * vector seeded by `seed`.
 * Mixed precision: when `plan` is an FP32 / BF16X3 plan and `refine` is an FP64
 * plan of the same operator (same psi length and element type, else an error),
 * the Lanczos solve runs on `plan` until the residuals of its nev lowest pairs
 * pass `tol` (e.g. nev 1, ncv 24, tol 3e-5, as the sweep engine does) and its
 * lowest min(4, ncv) Ritz vectors start an FP64 Jacobi-Davidson refinement: the
 * Rayleigh-Ritz basis grows by one vector per step (one FP64 matvec), the
 * correction from at most 16 conjugate-gradient steps on the projected
 * correction equation with the low-precision operator (for a BF16X3 plan its
 * coarse product of the hi planes alone), at most 12 steps, until
 * the FP64 residual passes refine_tol*max(|theta|, eps^(2/3)) (refine_tol <= 0:
 * 1e-10).  Starting from several Ritz vectors keeps near-degenerate lowest
 * pairs apart.  evals[0] and evec then come from the refinement.
 * info (6 ints, may be NULL): {Lanczos matvecs, restarts, Lanczos converged
 * (0/1), FP64 refinement matvecs, refined pair converged (0/1; the Lanczos flag
 * when not refined), low-precision matvecs of the refinement}.
 * resid (may be NULL): |H u - theta u| of the returned pair, computed from
 * FP64 matvecs of the refinement; without refinement from one extra matvec of
 * `plan`. */
int dmrgx_eigsh(dmrgx_heff* plan, int nev, int ncv, double tol, int max_restart,
                uint64_t seed, const void* v0, double* evals, void* evec,
                dmrgx_heff* refine, double refine_tol, int* info, double* resid);

/* dmrgx_eigsh on many independent operators at once: `plan` maps the
 * concatenation of n_chains vectors (chain c: len[c] elements, in order) to
 * itself without coupling them (a plan built from all chains' get_map tasks
 * with shifted offsets).  Each chain starts from its own seeded vector
 * (seeds[c]) and follows dmrgx_eigsh's algorithm and convergence test; the
 * chains' Gram-Schmidt reductions and vector updates are segmented launches
 * and every matvec is one call of `plan` for all chains; all chains take the
 * same number of steps until the last one has converged.  Chains of at most
 * 64 elements are diagonalised densely.  evals: n_chains x nev; evec: the
 * concatenated lowest vectors; refine / refine_tol as in dmrgx_eigsh (one
 * batched refinement); info: n_chains x 6 as in dmrgx_eigsh (the refinement
 * counts are for the whole batch); resid: n_chains.  No counterpart in the
 * reference (one ARPACK call per chain, src/dsaupd.cpp:65-155,
 * src/znaupd.cpp:78-197). */
int dmrgx_eigsh_batched(dmrgx_heff* plan, int n_chains, const int64_t* len, int nev, int ncv, double tol,
                        int max_restart, const uint64_t* seeds, double* evals, void* evec, dmrgx_heff* refine,
                        double refine_tol, int* info, double* resid);

/* Work-decomposition autotuner of the H_eff plans (on by default).  Each
 * pass plan (R and L) cuts its output tiles and the concatenated K of every
 * output block into work items; the number of items per launch decides
 * split-K depth and load balance.  The first plan with a given sector-shape
 * histogram (task shapes with M, N, K rounded to powers of two, counts too,
 * per pass, with precision and element type) of at least 1e9 flop times 1, 2,
 * 4 and 8 items per CU, and for BF16x3 a decomposition balanced over the
 * resident work-groups (whole output blocks cut into two K halves, items
 * ordered per XCD), with HIP events and keeps the fastest for every later plan
 * with that histogram in the process (the default of 2 per CU is kept unless
 * another is 3 % faster).  Choices change the summation order only.
 * dmrgx_autotune_keys: histograms timed so far.  No counterpart in the
 * reference. */
int dmrgx_set_autotune(int on);
int dmrgx_autotune_keys(void);
/* Pin the work decomposition of a plan's two passes (no counterpart in the
 * reference): target > 0 cuts K so that the pass has about that many work
 * items, 0 is the default (two per CU), < 0 balances the pass over -target
 * resident work-groups (whole output blocks cut into two K halves, items
 * ordered per XCD).  dmrgx_heff_get_split writes {R target, L target, R work
 * items, L work items}.  Results change in summation order only. */
int dmrgx_heff_set_split(dmrgx_heff* plan, int r_target, int l_target);
int dmrgx_heff_get_split(const dmrgx_heff* plan, int* targets);

/* dmrgx_eigsh_batched (and the batched sweeps) run each Lanczos step -- the
 * matvec and both Gram-Schmidt passes, then the normalisation of the next
 * vector -- as hipGraphs captured at their first use and replayed in every
 * restart cycle (on = 1); off (0, the default) launches the same kernels one
 * by one.  Results are identical.  Off is the default because a DMRG step's
 * plan lives for a few restart cycles only: on the 64-chain L=32 m=128 sweep
 * the capture and instantiation cost more than the replays save (27.9 s
 * against 25.5 s).  No counterpart in the reference. */
int dmrgx_set_lanczos_graphs(int on);

/* ---- blockwise SVD -----------------------------------------------------------
 * Replaces the per-sector tensor<T>::svd -> ?gesvd loop of qtensor<T>::svd
 * (src/qtensor.cpp:1167-1176, src/tensor.cpp:736-785, src/dgesvd.cpp:32-97,
 * src/zgesvd.cpp:36-104).  For each block b: A_b (rows x cols, row-major at
 * a_off[b]) = U_b diag(S_b) V_b with U_b full (rows x rows, at u_off[b]),
 * S_b descending (min(rows,cols), at s_off[b] in S), V_b full (cols x cols,
 * at v_off[b]).  All-zero blocks (every |x| <= 1e-10) give identity U, V and
 * S = 0; 1x1 blocks give U = 1, S = |a|, V = a/|a| (src/tensor.cpp:764-782).
 * FP64: Householder QR of each block, then batched blocked one-sided Jacobi on
 * R^H; A is not modified.  All pointers device. */
int dmrgx_block_svd(int dtype, int n_blk, const int32_t* rows, const int32_t* cols,
                    const void* A, const int64_t* a_off, void* U, const int64_t* u_off,
                    double* S, const int64_t* s_off, void* V, const int64_t* v_off);

/* Jacobi sweeps taken by the last dmrgx_block_svd call on this host thread
 * (diagnostic; no counterpart in the reference, whose ?gesvd reports none). */
int dmrgx_block_svd_sweeps(void);

/* ---- block contraction kernels used by the environment updates --------------
 * dmrgx_block_gemm: the batched form of qtensor<T>::contract(A, B, ta, tb, num)
 * (src/qtensor.cpp:824-899 -> src/tensor.cpp:603-649): one task per block
 * pair, tasks[t] = {a_off, b_off, c_off, M, N, K, beta} (7 int64); C(MxN)
 * = op(A) op(B) (+ C if beta); op(A) = A (MxK) for ta 'N', A^T (A is KxM) for
 * 'T'; op(B) = B (KxN) for tb 'N', B^T (B is NxK) for 'T'.  All FP64. */
int dmrgx_block_gemm(int dtype, char ta, char tb, const void* A, const void* B, void* C,
                     const int64_t* tasks, int n_task);

/* dmrgx_leg_contract: the block kernel of qtensor<T>::contract(A, a, B, b,
 * change, fermion) (src/qtensor.cpp:902-983 -> src/tensor.cpp:652-695):
 * out[i1,j1,j2,i2] (+)= sign * sum_k A[i1,k,i2] * B[j1,k,j2] for each task
 * tasks[t] = {a_off, b_off, c_off, al, mid, ar, bl, br, sign, accumulate}
 * (10 int64). */
int dmrgx_leg_contract(int dtype, const void* A, const void* B, void* C,
                       const int64_t* tasks, int n_task);

/* dmrgx_permute: out block = transpose of an in block with up to 8 legs:
 * tasks[t] = {in_off, out_off, nleg, dims[8] (input order), perm[8]} (19 int64);
 * output leg q is input leg perm[q] (tensor<T>::exchange, src/tensor.cpp:335-366). */
int dmrgx_permute(int dtype, const void* in, void* out, const int64_t* tasks, int n_task);

/* dmrgx_copy_boxes: strided 3-D box copies used by combine/split/resize
 * (src/qtensor.cpp:466-656, src/tensor.cpp:482-513): tasks[t] = {src_off,
 * src_s0, src_s1, dst_off, dst_s0, dst_s1, n0, n1, n2, scale_sign} (10 int64):
 * dst[dst_off + i*dst_s0 + j*dst_s1 + k] = src[src_off + i*src_s0 + j*src_s1 + k]
 * (times -1 when scale_sign < 0; conjugated when scale_sign == 2). */
int dmrgx_copy_boxes(int dtype, const void* src, void* dst, const int64_t* tasks, int n_task);

/* ---- measurement ---------------------------------------------------------
 * No counterpart in the reference (which has no GPU path): calibration for
 * the roofline figures.  dmrgx_calibrate_mfma times `rounds` x 4 independent
 * instructions of one matrix-core kind per wavefront, register operands only,
 * on waves_per_simd wavefronts per SIMD of every CU, and returns the rate in
 * TFLOP/s: kind 0 v_mfma_f32_32x32x16_bf16, 1 v_mfma_f32_16x16x32_bf16,
 * 2 v_mfma_f64_16x16x4f64, 3 v_mfma_f32_16x16x4f32. */
int dmrgx_calibrate_mfma(int kind, int waves_per_simd, int rounds, double* tflops);

#ifdef __cplusplus
}
#endif

#endif /* DMRGX_H */
