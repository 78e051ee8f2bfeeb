// Batched FP64 one-sided (Hestenes) Jacobi SVD, blocked.
//
// Each block works on W = A (rows >= cols) or W = A^H (rows < cols), stored
// with one contiguous row per column of W, and accumulates the rotations in
// Q (same layout).  The columns are grouped into column blocks of BJ; one
// round of the round-robin ordering pairs every column block with one other,
// and for each pair (I, J) the 2*BJ columns are orthogonalised together:
//   1. partial Gram matrices G = [W_I W_J]^H [W_I W_J] over row chunks,
//   2. G = Z diag Z^H by cyclic two-sided Jacobi in LDS (the rotations are
//      the ones one-sided Jacobi would apply to the columns themselves),
//   3. [W_I W_J] <- [W_I W_J] Z and [Q_I Q_J] <- [Q_I Q_J] Z over row chunks.
// Two launches per round cover every pair of every block, so all sectors of
// a qtensor::svd run together.  A sweep ends with one read-back of the largest
// |w_p^H w_q| / (|w_p||w_q|) seen, the convergence test.  Replaces LAPACK
// ?gesvd in the reference's per-sector loop (src/qtensor.cpp:1167-1176,
// src/tensor.cpp:736-785).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>
#include <cstring>

#include "dmrgx.h"
#include "common.hpp"
#include "svd.hpp"

namespace dmrgx {
namespace {

constexpr double TOLZERO = 1e-10;  // global.h:19, all-zero block test (src/tensor.cpp:764)
// off-diagonal convergence threshold: |w_p^H w_q| <= tol |w_p| |w_q| with
// tol = max(rows, cols) * eps (the one-sided Jacobi test of Demmel and Veselic)
constexpr double EPS = 2.220446049250313e-16;
constexpr int MAX_SWEEPS = 40;
// Columns of norm below NEGL sqrt(rows cols) max|a| (between 1e-14 and ~2e-11
// of sigma_max) take no part in the rotations or the convergence test: their
// directions are rounding noise, and they are replaced by the Householder
// complement of the others at the end.
constexpr double NEGL = 1e-14;

constexpr int BJ = 16;        // columns per column block
constexpr int PJ = 2 * BJ;    // columns per block pair
constexpr int RCG = 64;       // rows per Gram chunk
constexpr int RCA = 64;       // rows per update chunk
constexpr int MAXIN = 1;      // inner Jacobi sweeps on one Gram matrix per visit
constexpr int NB = 8;         // Householder panel width
constexpr int CT = 16;        // columns per work-group of the block-reflector update

struct JBlk {
  int m, n;         // W has n columns of length m
  int transposed;   // W = A^H
  int rows, cols;   // original A shape
  int special;      // 0 normal, 1 all-zero, 2 one-by-one
  int nbk, nplay;   // column blocks; players in the round robin (nbk rounded up to even)
  int nchg, ncha;   // Gram row chunks; update row chunks (W rows then Q rows)
  int npan, nbc;    // QR panels; columns of the back-transformed factor B
  double tol;
  double negl2;     // columns with |x|^2 <= negl2 are negligible (below ~1e-14 of the block norm)
  int64_t a_off, w_off, q_off;
  int64_t gram_base;   // first global work-group of the Gram launch
  int64_t apply_base;  // first global work-group of the update launch
  int64_t gp_off;      // partial Gram matrices (pairs x nchg x PJ x PJ elements)
  int64_t pair_base;   // first global pair (work-group of the pair eigensolver)
  int64_t elem_base;   // first flat element of the element-wise kernels
  int64_t c_off;       // per-column reflector data (beta, tau)
  int64_t t_off;       // triangular factors T of the panels (npan x NB x NB)
  int64_t b_off;       // B = Q_h [V; 0], m x nbc
};

enum { BY_ELEM = 0, BY_GRAM = 1, BY_APPLY = 2, BY_PAIR = 3 };

__device__ __forceinline__ int find_blk(const JBlk* b, int nb, int64_t x, int by) {
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    int64_t s = by == BY_GRAM ? b[mid].gram_base
                : (by == BY_APPLY ? b[mid].apply_base : (by == BY_PAIR ? b[mid].pair_base : b[mid].elem_base));
    if (s <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// round-robin pairing of N (even) players: pair p of round r
__device__ __forceinline__ void rr_pair(int N, int r, int p, int& i, int& j) {
  if (p == 0) { i = N - 1; j = r; }
  else { i = (r + p) % (N - 1); j = (r - p + N - 1) % (N - 1); }
  if (i > j) { int t = i; i = j; j = t; }
}

// column of W behind slot c of the pair (I, J); -1 if none
__device__ __forceinline__ int pair_col(const JBlk& k, int I, int J, int c) {
  const int blk = c < BJ ? I : J;
  const int col = blk * BJ + (c & (BJ - 1));
  return (blk < k.nbk && col < k.n) ? col : -1;
}

// W (column j contiguous, length m) from A or A^H
template <bool CPLX>
__global__ __launch_bounds__(256) void k_jinit(const double* __restrict__ A, double* __restrict__ W,
                                               const JBlk* __restrict__ blks, int nb, int64_t total) {
  int64_t x = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (x >= total) return;
  const int b = find_blk(blks, nb, x, BY_ELEM);
  const JBlk k = blks[b];
  const int64_t e = x - k.elem_base;
  if (e >= (int64_t)k.n * k.m) return;
  const int j = (int)(e / k.m), i = (int)(e % k.m);  // W column j, row i
  int64_t src = k.transposed ? k.a_off + (int64_t)j * k.cols + i : k.a_off + (int64_t)i * k.cols + j;
  if (CPLX) {
    W[2 * (k.w_off + e)] = A[2 * src];
    W[2 * (k.w_off + e) + 1] = k.transposed ? -A[2 * src + 1] : A[2 * src + 1];
  } else {
    W[k.w_off + e] = A[src];
  }
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// ---------------------------------------------------------------------------
// Householder QR of W (the preconditioning step of Drmac and Veselic: one-sided
// Jacobi on R^H needs a handful of sweeps where Jacobi on W itself needs
// dozens for exponentially decaying spectra).  Reflectors are Hermitian,
// H = I - tau v v^H with v_0 = x_0 + e^{i arg x_0}|x| and tau = 2 / |v|^2, so
// H x = beta e_1 with beta = -e^{i arg x_0}|x|.  Panels of NB columns; the
// product of a panel's reflectors is I - V T V^H (T upper triangular).

// Squared norms of every column of W and the identity permutation: one
// wavefront per column.
template <bool CPLX>
__global__ __launch_bounds__(256) void k_qr_colnorm(const double* __restrict__ W, const JBlk* __restrict__ blks,
                                                    const int64_t* __restrict__ colbase, int nb, int64_t total,
                                                    double* __restrict__ CN, int* __restrict__ PIV) {
  constexpr int C = CPLX ? 2 : 1;
  const int64_t g = blockIdx.x * (int64_t)4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (g >= total) return;
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (colbase[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  const JBlk k = blks[lo];
  const int j = (int)(g - colbase[lo]);
  const double* w = W + C * (k.w_off + (int64_t)j * k.m);
  double a = 0;
  for (int x = lane; x < k.m; x += 64) {
    a += w[C * x] * w[C * x];
    if (CPLX) a += w[C * x + 1] * w[C * x + 1];
  }
  a = wave_sum(a);
  if (lane == 0) {
    CN[k.c_off + j] = a;
    PIV[k.c_off + j] = j;
  }
}

// One work-group per block factors panel kp; the panel rows [k0, m) live in
// registers, RPT rows per thread.  Column j of the panel keeps v_j in rows
// >= j (v_0 at row j) and R entries above; beta_j, tau_j and T go to their
// own arrays.  Column pivoting by panel: the panel takes, in order, the
// trailing columns of largest norm over rows >= k0 (the norms are exact at
// the start of a panel; they are not downdated inside it), swapping whole
// columns of W and recording the permutation in PIV.
template <bool CPLX, int RPT>
__global__ __launch_bounds__(512) void k_qr_panel(double* __restrict__ W, const JBlk* __restrict__ blks, int kp,
                                                  double* __restrict__ BETA, double* __restrict__ TAU,
                                                  double* __restrict__ TW, double* __restrict__ CN,
                                                  int* __restrict__ PIV) {
  constexpr int C = CPLX ? 2 : 1;
  __shared__ double sred[8][2 * NB];
  __shared__ double smaxv[8];
  __shared__ int smaxi[8];
  __shared__ double sx0[2];
  __shared__ double sS[NB][NB][2];
  __shared__ double sTau[NB];
  __shared__ double sT[NB][NB][2];
  const JBlk k = blks[blockIdx.x];
  if (k.special) return;
  const int k0 = kp * NB;
  if (k0 >= k.n) return;
  const int ncl = min(NB, k.n - k0);
  const int NT = blockDim.x, t = threadIdx.x, lane = t & 63, wave = t >> 6, nw = NT >> 6;
  // pivot selection and column swaps
  for (int c = 0; c < ncl; ++c) {
    const int j = k0 + c;
    double bv = -1;
    int bi = k.n;
    for (int x = j + t; x < k.n; x += NT) {
      const double v = CN[k.c_off + x];
      if (v > bv) { bv = v; bi = x; }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(bv, off, 64);
      const int oi = __shfl_xor(bi, off, 64);
      if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
    }
    if (lane == 0) { smaxv[wave] = bv; smaxi[wave] = bi; }
    __syncthreads();
    int p = smaxi[0];
    double pv = smaxv[0];
    for (int w = 1; w < nw; ++w)
      if (smaxv[w] > pv || (smaxv[w] == pv && smaxi[w] < p)) { pv = smaxv[w]; p = smaxi[w]; }
    if (p != j && p < k.n) {
      for (int r = t; r < k.m; r += NT) {
        const int64_t ea = k.w_off + (int64_t)j * k.m + r, eb = k.w_off + (int64_t)p * k.m + r;
        const double ar = W[C * ea], br = W[C * eb];
        W[C * ea] = br;
        W[C * eb] = ar;
        if (CPLX) {
          const double ai = W[C * ea + 1], bi2 = W[C * eb + 1];
          W[C * ea + 1] = bi2;
          W[C * eb + 1] = ai;
        }
      }
      if (t == 0) {
        const double cv = CN[k.c_off + j];
        CN[k.c_off + j] = CN[k.c_off + p];
        CN[k.c_off + p] = cv;
        const int pv2 = PIV[k.c_off + j];
        PIV[k.c_off + j] = PIV[k.c_off + p];
        PIV[k.c_off + p] = pv2;
      }
    }
    __syncthreads();
  }
  double pr[RPT][NB], pim[RPT][NB];
#pragma unroll
  for (int i = 0; i < RPT; ++i) {
    const int r = k0 + t + NT * i;
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      const bool ok = r < k.m && c < ncl;
      const int64_t e = k.w_off + (int64_t)(k0 + c) * k.m + r;
      pr[i][c] = ok ? W[C * e] : 0.0;
      pim[i][c] = (CPLX && ok) ? W[C * e + 1] : 0.0;
    }
  }
#pragma unroll
  for (int jj = 0; jj < NB; ++jj) {
    if (jj >= ncl) break;  // uniform across the work-group
    const int j = k0 + jj;
    // |x|^2 over rows >= j and x_0 = row j
    double s2 = 0;
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int r = k0 + t + NT * i;
      if (r >= j && r < k.m) s2 += pr[i][jj] * pr[i][jj] + pim[i][jj] * pim[i][jj];
      if (r == j) { sx0[0] = pr[i][jj]; sx0[1] = pim[i][jj]; }
    }
    s2 = wave_sum(s2);
    if (lane == 0) sred[wave][0] = s2;
    __syncthreads();
    double nrm2 = 0;
    for (int w = 0; w < nw; ++w) nrm2 += sred[w][0];
    const double x0r = sx0[0], x0i = sx0[1];
    const double nrm = sqrt(nrm2), ax0 = CPLX ? hypot(x0r, x0i) : fabs(x0r);
    const double er = ax0 > 0 ? x0r / ax0 : 1.0, ei = ax0 > 0 ? x0i / ax0 : 0.0;
    const double vn2 = 2 * nrm * (nrm + ax0);
    const double tau = vn2 > 0 ? 2 / vn2 : 0.0;
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      if (k0 + t + NT * i == j) { pr[i][jj] = x0r + er * nrm; pim[i][jj] = x0i + ei * nrm; }
    }
    // y_c = v^H P[:, c] over rows >= j, every c != jj (c < jj gives conj(v_c^H v))
    double yr[NB], yi[NB];
#pragma unroll
    for (int c = 0; c < NB; ++c) { yr[c] = 0; yi[c] = 0; }
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int r = k0 + t + NT * i;
      if (r >= j && r < k.m) {
        const double vr = pr[i][jj], vi = pim[i][jj];
#pragma unroll
        for (int c = 0; c < NB; ++c) {
          yr[c] += vr * pr[i][c] + vi * pim[i][c];
          if (CPLX) yi[c] += vr * pim[i][c] - vi * pr[i][c];
        }
      }
    }
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      yr[c] = wave_sum(yr[c]);
      if (CPLX) yi[c] = wave_sum(yi[c]);
    }
    __syncthreads();  // sred of the norm is read by everyone
    if (lane == 0) {
#pragma unroll
      for (int c = 0; c < NB; ++c) { sred[wave][2 * c] = yr[c]; sred[wave][2 * c + 1] = yi[c]; }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      double a = 0, bq = 0;
      for (int w = 0; w < nw; ++w) { a += sred[w][2 * c]; bq += sred[w][2 * c + 1]; }
      yr[c] = a;
      yi[c] = bq;
    }
    // P[:, c] -= tau v y_c for c > jj
#pragma unroll
    for (int i = 0; i < RPT; ++i) {
      const int r = k0 + t + NT * i;
      if (r >= j && r < k.m) {
        const double vr = pr[i][jj], vi = pim[i][jj];
#pragma unroll
        for (int c = 0; c < NB; ++c) {
          if (c > jj) {
            pr[i][c] -= tau * (vr * yr[c] - vi * yi[c]);
            if (CPLX) pim[i][c] -= tau * (vr * yi[c] + vi * yr[c]);
          }
        }
      }
    }
    if (t == 0) {
      for (int c = 0; c < jj; ++c) { sS[c][jj][0] = yr[c]; sS[c][jj][1] = -yi[c]; }
      sTau[jj] = tau;
      BETA[C * (k.c_off + j)] = -er * nrm;
      if (CPLX) BETA[C * (k.c_off + j) + 1] = -ei * nrm;
      TAU[k.c_off + j] = tau;
    }
    __syncthreads();
  }
  // T: T_ii = tau_i, T[0:i, i] = -tau_i T[0:i, 0:i] (V[:, 0:i]^H v_i)
  if (t == 0) {
    for (int a = 0; a < NB; ++a)
      for (int c = 0; c < NB; ++c) { sT[a][c][0] = 0; sT[a][c][1] = 0; }
    for (int i = 0; i < ncl; ++i) {
      sT[i][i][0] = sTau[i];
      for (int a = 0; a < i; ++a) {
        double zr = 0, zi = 0;
        for (int l = a; l < i; ++l) {
          zr += sT[a][l][0] * sS[l][i][0] - sT[a][l][1] * sS[l][i][1];
          zi += sT[a][l][0] * sS[l][i][1] + sT[a][l][1] * sS[l][i][0];
        }
        sT[a][i][0] = -sTau[i] * zr;
        sT[a][i][1] = -sTau[i] * zi;
      }
    }
  }
  __syncthreads();
  if (t < NB * NB) {
    double* T = TW + C * (k.t_off + (int64_t)kp * NB * NB);
    T[C * t] = sT[t / NB][t % NB][0];
    if (CPLX) T[C * t + 1] = sT[t / NB][t % NB][1];
  }
#pragma unroll
  for (int i = 0; i < RPT; ++i) {
    const int r = k0 + t + NT * i;
    if (r >= k.m) continue;
#pragma unroll
    for (int c = 0; c < NB; ++c) {
      if (c >= ncl) continue;
      const int64_t e = k.w_off + (int64_t)(k0 + c) * k.m + r;
      W[C * e] = pr[i][c];
      if (CPLX) W[C * e + 1] = pim[i][c];
    }
  }
}

// Block reflector of panel kp applied to CT columns of a target with columns
// of length m: X <- (I - V op(T) V^H) X with op(T) = T^H (trailing update of
// the QR, target W, columns [k0 + NB, n)) or T (back-transformation, target
// B, columns [0, nbc)).  Work-group = (block, tile of CT columns).
template <bool CPLX>
__global__ __launch_bounds__(256) void k_qr_apply(const double* __restrict__ W, double* __restrict__ X,
                                                  const JBlk* __restrict__ blks, int tiles, int kp,
                                                  const double* __restrict__ TW, int trailing,
                                                  double* __restrict__ CN) {
  constexpr int C = CPLX ? 2 : 1;
  constexpr int RC = 64;
  __shared__ double sV[RC][NB][2];
  __shared__ double sC[RC][CT][2];
  __shared__ double sY[2][NB][CT][2];
  __shared__ double sT[NB][NB][2];
  const int b = blockIdx.x / tiles, tile = blockIdx.x % tiles;
  const JBlk k = blks[b];
  if (k.special) return;
  const int k0 = kp * NB;
  if (k0 >= k.n) return;
  const int ncl = min(NB, k.n - k0);
  const int c_lo = trailing ? k0 + NB : 0, c_hi = trailing ? k.n : k.nbc;
  const int cb = c_lo + tile * CT;
  if (cb >= c_hi) return;
  const int64_t xbase = trailing ? k.w_off : k.b_off;
  const int t = threadIdx.x;
  if (t < NB * NB) {
    const double* T = TW + C * (k.t_off + (int64_t)kp * NB * NB);
    sT[t / NB][t % NB][0] = T[C * t];
    sT[t / NB][t % NB][1] = CPLX ? T[C * t + 1] : 0.0;
  }
  auto vval = [&](int r, int kk, double& vr, double& vi) {  // V[r][kk], zero above the diagonal
    vr = vi = 0;
    if (kk < ncl && r >= k0 + kk && r < k.m) {
      const int64_t e = k.w_off + (int64_t)(k0 + kk) * k.m + r;
      vr = W[C * e];
      if (CPLX) vi = W[C * e + 1];
    }
  };
  // Y = V^H X: thread (kk, c, half)
  const int kk = (t >> 4) & (NB - 1), c = t & (CT - 1), half = t >> 7;
  double ar = 0, ai = 0;
  for (int r0 = k0; r0 < k.m; r0 += RC) {
    for (int e = t; e < RC * NB; e += 256) {
      const int rr = e % RC, q = e / RC;
      vval(r0 + rr, q, sV[rr][q][0], sV[rr][q][1]);
    }
    for (int e = t; e < RC * CT; e += 256) {
      const int rr = e % RC, q = e / RC;
      const int r = r0 + rr, col = cb + q;
      const bool ok = r < k.m && col < c_hi;
      const int64_t x = xbase + (int64_t)col * k.m + r;
      sC[rr][q][0] = ok ? X[C * x] : 0.0;
      sC[rr][q][1] = (CPLX && ok) ? X[C * x + 1] : 0.0;
    }
    __syncthreads();
    for (int rr = half * (RC / 2); rr < (half + 1) * (RC / 2); ++rr) {
      const double vr = sV[rr][kk][0], vi = sV[rr][kk][1], xr = sC[rr][c][0], xi = sC[rr][c][1];
      ar += vr * xr + vi * xi;  // conj(v) x
      if (CPLX) ai += vr * xi - vi * xr;
    }
    __syncthreads();
  }
  sY[half][kk][c][0] = ar;
  sY[half][kk][c][1] = ai;
  __syncthreads();
  // Y2 = op(T) Y
  double y2r = 0, y2i = 0;
  if (t < NB * CT) {
    for (int l = 0; l < NB; ++l) {
      const double yr = sY[0][l][c][0] + sY[1][l][c][0], yi = sY[0][l][c][1] + sY[1][l][c][1];
      double tr, ti;
      if (trailing) { tr = sT[l][kk][0]; ti = -sT[l][kk][1]; }  // (T^H)[kk][l]
      else { tr = sT[kk][l][0]; ti = sT[kk][l][1]; }
      y2r += tr * yr - ti * yi;
      y2i += tr * yi + ti * yr;
    }
  }
  __syncthreads();
  if (t < NB * CT) { sY[0][kk][c][0] = y2r; sY[0][kk][c][1] = y2i; }
  __syncthreads();
  // X -= V Y2: thread (row, group of 4 columns); the trailing update also
  // sums the squared norms of its columns over rows >= k0 + NB (the next
  // panel's pivot norms)
  const int row = t & (RC - 1), cq = t >> 6;
  double nrm[CT / 4] = {};
  for (int r0 = k0; r0 < k.m; r0 += RC) {
    const int r = r0 + row;
    if (r >= k.m) continue;
    double vr[NB], vi[NB];
#pragma unroll
    for (int q = 0; q < NB; ++q) vval(r, q, vr[q], vi[q]);
#pragma unroll
    for (int q4 = 0; q4 < CT / 4; ++q4) {
      const int cc = cq * (CT / 4) + q4, col = cb + cc;
      if (col >= c_hi) continue;
      double dr = 0, di = 0;
#pragma unroll
      for (int q = 0; q < NB; ++q) {
        dr += vr[q] * sY[0][q][cc][0] - vi[q] * sY[0][q][cc][1];
        di += vr[q] * sY[0][q][cc][1] + vi[q] * sY[0][q][cc][0];
      }
      const int64_t x = xbase + (int64_t)col * k.m + r;
      const double nr = X[C * x] - dr, ni = CPLX ? X[C * x + 1] - di : 0.0;
      X[C * x] = nr;
      if (CPLX) X[C * x + 1] = ni;
      if (r >= k0 + NB) nrm[q4] += nr * nr + ni * ni;
    }
  }
  if (trailing) {
#pragma unroll
    for (int q4 = 0; q4 < CT / 4; ++q4) {
      const double v = wave_sum(nrm[q4]);
      const int col = cb + cq * (CT / 4) + q4;
      if (row == 0 && col < c_hi) CN[k.c_off + col] = v;
    }
  }
}

// X = R^H (n x n, column j contiguous) and Q = I; jb holds the Jacobi view of
// each block (m = n, w_off = X offset), qb the QR view (W offsets, beta).
template <bool CPLX>
__global__ __launch_bounds__(256) void k_qr_rh(const double* __restrict__ W, const double* __restrict__ BETA,
                                               double* __restrict__ X, double* __restrict__ Q,
                                               const JBlk* __restrict__ qb, const JBlk* __restrict__ jb, int nb,
                                               int64_t total) {
  constexpr int C = CPLX ? 2 : 1;
  int64_t x = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (x >= total) return;
  const int b = find_blk(jb, nb, x, BY_ELEM);
  const JBlk k = jb[b], q = qb[b];
  int64_t e = x - k.elem_base;
  const int64_t nn = (int64_t)k.n * k.n;
  const bool isX = e < nn;
  if (!isX) e -= nn;
  if (e >= nn) return;
  const int j = (int)(e / k.n), i = (int)(e % k.n);  // column j, row i
  double vr = 0, vi = 0;
  if (isX) {
    if (q.special) {  // no QR: zero block, or the 1x1 block |a| e^{i arg a}
      if (q.special == 2 && i == 0 && j == 0) {
        vr = W[C * q.w_off];
        vi = CPLX ? -W[C * q.w_off + 1] : 0.0;
      }
    } else if (i == j) {
      vr = BETA[C * (q.c_off + j)];
      vi = CPLX ? -BETA[C * (q.c_off + j) + 1] : 0.0;
    } else if (i > j) {  // conj(R[j][i]), R[j][i] = W column i, row j
      const int64_t s = q.w_off + (int64_t)i * q.m + j;
      vr = W[C * s];
      vi = CPLX ? -W[C * s + 1] : 0.0;
    }
    X[C * (k.w_off + e)] = vr;
    if (CPLX) X[C * (k.w_off + e) + 1] = vi;
  } else {
    Q[C * (k.q_off + e)] = i == j ? 1.0 : 0.0;
    if (CPLX) Q[C * (k.q_off + e) + 1] = 0.0;
  }
}

// B = [Q[:, perm]; 0 | e_n ...] (m x nbc, column c contiguous)
template <bool CPLX>
__global__ __launch_bounds__(256) void k_binit(const double* __restrict__ Q, double* __restrict__ B,
                                               const JBlk* __restrict__ qb, const JBlk* __restrict__ jb,
                                               const int64_t* __restrict__ bbase, const int64_t* __restrict__ col0,
                                               const int* __restrict__ perm, int nb, int64_t total) {
  constexpr int C = CPLX ? 2 : 1;
  int64_t x = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (x >= total) return;
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (bbase[mid] <= x) lo = mid;
    else hi = mid - 1;
  }
  const JBlk k = qb[lo], jv = jb[lo];
  const int64_t e = x - bbase[lo];
  const int c = (int)(e / k.m), r = (int)(e % k.m);
  double vr = 0, vi = 0;
  if (c < k.n) {
    if (r < k.n) {
      const int src = perm[col0[lo] + c];
      const int64_t s = jv.q_off + (int64_t)src * k.n + r;
      vr = Q[C * s];
      vi = CPLX ? Q[C * s + 1] : 0.0;
    }
  } else {
    vr = r == c ? 1.0 : 0.0;
  }
  B[C * (k.b_off + e)] = vr;
  if (CPLX) B[C * (k.b_off + e) + 1] = vi;
}

// Partial Gram matrices of one round: work-group = (pair, row chunk).  256
// threads: thread t owns the 4x4 tile (t & 63) of the 32x32 result over the
// chunk rows [16 * (t >> 6), +16); the four row quarters are summed by the
// first wavefront through one LDS tile that reuses the column stage, so a
// work-group holds 33 KB of LDS (complex) and four fit on a CU.
template <bool CPLX>
__global__ __launch_bounds__(256) void k_bgram(const double* __restrict__ W, const JBlk* __restrict__ blks, int nb,
                                               int64_t total, int round, double* __restrict__ GP) {
  constexpr int C = CPLX ? 2 : 1;
  constexpr int XS = (RCG + 1) * C;  // LDS row stride of one column (padded)
  __shared__ double sX[PJ * XS];
  static_assert(PJ * PJ * C <= PJ * XS, "the reduction tile must fit in the column stage");
  const int64_t g = blockIdx.x;
  if (g >= total) return;
  const int b = find_blk(blks, nb, g, BY_GRAM);
  const JBlk k = blks[b];
  if (k.special || round >= k.nplay - 1) return;
  const int64_t li = g - k.gram_base;
  const int pair = (int)(li / k.nchg), chunk = (int)(li % k.nchg);
  int I, J;
  rr_pair(k.nplay, round, pair, I, J);
  const int t = threadIdx.x;
  const int r0 = chunk * RCG;
  for (int e = t; e < PJ * RCG; e += 256) {
    const int c = e / RCG, x = e % RCG;
    const int col = pair_col(k, I, J, c);
    const bool ok = col >= 0 && r0 + x < k.m;
    const int64_t src = k.w_off + (int64_t)col * k.m + r0 + x;
    sX[c * XS + x * C] = ok ? W[C * src] : 0.0;
    if (CPLX) sX[c * XS + x * C + 1] = ok ? W[C * src + 1] : 0.0;
  }
  __syncthreads();
  const int tile = t & 63, ph = t >> 6;
  const int pi0 = (tile >> 3) * 4, qj0 = (tile & 7) * 4;
  double ar[4][4] = {}, ai[4][4] = {};
  for (int x = ph * (RCG / 4); x < (ph + 1) * (RCG / 4); ++x) {
    double pr_[4], pim[4], qr_[4], qim[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      pr_[i] = sX[(pi0 + i) * XS + x * C];
      qr_[i] = sX[(qj0 + i) * XS + x * C];
      pim[i] = CPLX ? sX[(pi0 + i) * XS + x * C + 1] : 0.0;
      qim[i] = CPLX ? sX[(qj0 + i) * XS + x * C + 1] : 0.0;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        ar[i][j] += pr_[i] * qr_[j] + pim[i] * qim[j];  // conj(p) * q
        if (CPLX) ai[i][j] += pr_[i] * qim[j] - pim[i] * qr_[j];
      }
  }
  // quarters 1, 2, 3 in turn through sX (now free): quarter 0's sum is
  // ((q0 + q1) + q2) + q3
  double* sR = sX;
  for (int q = 1; q < 4; ++q) {
    __syncthreads();
    if (ph == q) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int e = (pi0 + i) * PJ + qj0 + j;
          sR[e * C] = ar[i][j];
          if (CPLX) sR[e * C + 1] = ai[i][j];
        }
    }
    __syncthreads();
    if (ph == 0) {
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int e = (pi0 + i) * PJ + qj0 + j;
          ar[i][j] += sR[e * C];
          if (CPLX) ai[i][j] += sR[e * C + 1];
        }
    }
  }
  if (ph == 0) {
    double* out = GP + C * (k.gp_off + ((int64_t)pair * k.nchg + chunk) * PJ * PJ);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int e = (pi0 + i) * PJ + qj0 + j;
        out[e * C] = ar[i][j];
        if (CPLX) out[e * C + 1] = ai[i][j];
      }
  }
}

// The rotations of one pair: work-group = pair.  Sums the partial Gram
// matrices, records the largest off-diagonal ratio for the convergence test
// and, if some ratio exceeds the tolerance, runs cyclic Jacobi on G and
// writes Z (PJ x PJ, the accumulated rotations) with flag 1; flag 0 means
// Z = identity.
template <bool CPLX>
__global__ __launch_bounds__(256) void k_beig(const JBlk* __restrict__ blks, int nb, int64_t total, int round,
                                              const double* __restrict__ GP, double* __restrict__ ZW,
                                              int* __restrict__ ZF, unsigned long long* __restrict__ flag) {
  constexpr int C = CPLX ? 2 : 1;
  constexpr int GS = (PJ + 1) * C;  // LDS row stride of G and Z
  __shared__ double sG[PJ * GS], sZ[PJ * GS], sG2[PJ * GS], sZ2[PJ * GS];
  __shared__ double sPar[PJ / 2][4];
  __shared__ int sPQ[PJ / 2][2];
  __shared__ double sRed[4];
  __shared__ int sChanged;
  const int64_t g = blockIdx.x;
  if (g >= total) return;
  const int b = find_blk(blks, nb, g, BY_PAIR);
  const JBlk k = blks[b];
  if (k.special || round >= k.nplay - 1) return;
  const int pair = (int)(g - k.pair_base);
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;

  // 1. G = sum of the partial Gram matrices; Z = identity
  const double* gp = GP + C * (k.gp_off + (int64_t)pair * k.nchg * PJ * PJ);
  for (int e = t; e < PJ * PJ; e += 256) {
    const int r = e / PJ, c = e % PJ;
    double vr = 0, vi = 0;
    for (int ch = 0; ch < k.nchg; ++ch) {
      vr += gp[C * ((int64_t)ch * PJ * PJ + e)];
      if (CPLX) vi += gp[C * ((int64_t)ch * PJ * PJ + e) + 1];
    }
    sG[r * GS + c * C] = vr;
    if (CPLX) sG[r * GS + c * C + 1] = vi;
    sZ[r * GS + c * C] = r == c ? 1.0 : 0.0;
    if (CPLX) sZ[r * GS + c * C + 1] = 0.0;
  }
  __syncthreads();
  // 2. largest |G_pq| / sqrt(G_pp G_qq), p < q (as its square, one root at the end)
  double mx = 0;
  for (int e = t; e < PJ * PJ; e += 256) {
    const int r = e / PJ, c = e % PJ;
    if (r < c) {
      const double a = sG[r * GS + r * C], d = sG[c * GS + c * C];
      if (a > k.negl2 && d > k.negl2) {
        const double gr = sG[r * GS + c * C], gi = CPLX ? sG[r * GS + c * C + 1] : 0.0;
        mx = fmax(mx, (gr * gr + gi * gi) / (a * d));
      }
    }
  }
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
  if (lane == 0) sRed[wave] = mx;
  __syncthreads();
  mx = sqrt(fmax(fmax(sRed[0], sRed[1]), fmax(sRed[2], sRed[3])));
  if (t == 0) {
    if (mx > 0) atomicMax(&flag[b], (unsigned long long)__double_as_longlong(mx));
    ZF[k.pair_base + pair] = mx > k.tol ? 1 : 0;
  }
  if (!(mx > k.tol)) return;  // uniform across the work-group

  // 3. cyclic Jacobi on G, rotations accumulated in Z: G <- U^H G U, Z <- Z U
  //    with U = [[c, s e^{i phi}], [-s e^{-i phi}, c]] on columns (p, q).  The
  //    16 rotations of a round are disjoint, so G splits into 16 x 16 blocks
  //    rows {p_k, q_k} x cols {p_l, q_l}, each updated as U_k^H B U_l by one
  //    thread from the current buffer into the next one.
  double* Gc = sG;
  double* Gn = sG2;
  double* Zc = sZ;
  double* Zn = sZ2;
  for (int sw = 0; sw < MAXIN; ++sw) {
    if (t == 0) sChanged = 0;
    for (int rd = 0; rd < PJ - 1; ++rd) {
      if (t < PJ / 2) {
        int p, q;
        rr_pair(PJ, rd, t, p, q);
        const double a = Gc[p * GS + p * C], d = Gc[q * GS + q * C];
        const double gr = Gc[p * GS + q * C], gi = CPLX ? Gc[p * GS + q * C + 1] : 0.0;
        const double ag2 = gr * gr + gi * gi;
        double c = 1, sn = 0, er = 1, ei = 0;
        if (a > k.negl2 && d > k.negl2 && ag2 > k.tol * k.tol * (a * d)) {
          // t = tan(theta) = sign(d - a) 2|g| / (|d - a| + sqrt((d - a)^2 + 4|g|^2)),
          // the smaller root of t^2 + 2 zeta t - 1 = 0 with zeta = (d - a) / (2|g|);
          // the two square roots are independent, so the chain is sqrt, div,
          // sqrt, div
          const double ag = sqrt(ag2), dd = d - a;
          const double tt = (dd >= 0 ? 2.0 : -2.0) * ag / (fabs(dd) + sqrt(dd * dd + 4 * ag2));
          c = 1 / sqrt(1 + tt * tt);
          sn = c * tt;
          const double rag = 1 / ag;
          er = CPLX ? gr * rag : (gr >= 0 ? 1.0 : -1.0);
          ei = CPLX ? gi * rag : 0.0;
          sChanged = 1;
        }
        sPQ[t][0] = p;
        sPQ[t][1] = q;
        sPar[t][0] = c;
        sPar[t][1] = sn;
        sPar[t][2] = er;
        sPar[t][3] = ei;
      }
      __syncthreads();
      {  // block (kb, lb) of G
        const int kb = t >> 4, lb = t & 15;
        const int pk = sPQ[kb][0], qk = sPQ[kb][1], pl = sPQ[lb][0], ql = sPQ[lb][1];
        const double ck = sPar[kb][0], sk = sPar[kb][1], ekr = sPar[kb][2], eki = sPar[kb][3];
        const double cl = sPar[lb][0], sl = sPar[lb][1], elr = sPar[lb][2], eli = sPar[lb][3];
        const int rk[2] = {pk, qk};
        double br[2][2], bi[2][2];
#pragma unroll
        for (int x = 0; x < 2; ++x) {
          br[x][0] = Gc[rk[x] * GS + pl * C];
          br[x][1] = Gc[rk[x] * GS + ql * C];
          bi[x][0] = CPLX ? Gc[rk[x] * GS + pl * C + 1] : 0.0;
          bi[x][1] = CPLX ? Gc[rk[x] * GS + ql * C + 1] : 0.0;
        }
        // T = U_k^H B: T_p = c B_p - s e B_q, T_q = s conj(e) B_p + c B_q
        double tr[2][2], ti[2][2];
#pragma unroll
        for (int y = 0; y < 2; ++y) {
          tr[0][y] = ck * br[0][y] - sk * (ekr * br[1][y] - eki * bi[1][y]);
          ti[0][y] = ck * bi[0][y] - sk * (ekr * bi[1][y] + eki * br[1][y]);
          tr[1][y] = sk * (ekr * br[0][y] + eki * bi[0][y]) + ck * br[1][y];
          ti[1][y] = sk * (ekr * bi[0][y] - eki * br[0][y]) + ck * bi[1][y];
        }
        // B' = T U_l: B'_p = c T_p - s conj(e) T_q, B'_q = s e T_p + c T_q
#pragma unroll
        for (int x = 0; x < 2; ++x) {
          Gn[rk[x] * GS + pl * C] = cl * tr[x][0] - sl * (elr * tr[x][1] + eli * ti[x][1]);
          Gn[rk[x] * GS + ql * C] = sl * (elr * tr[x][0] - eli * ti[x][0]) + cl * tr[x][1];
          if (CPLX) {
            Gn[rk[x] * GS + pl * C + 1] = cl * ti[x][0] - sl * (elr * ti[x][1] - eli * tr[x][1]);
            Gn[rk[x] * GS + ql * C + 1] = sl * (elr * ti[x][0] + eli * tr[x][0]) + cl * ti[x][1];
          }
        }
      }
#pragma unroll
      for (int h = 0; h < 2; ++h) {  // Z rows x pairs: (row, pair) tasks t and t + 256
        const int task = t + 256 * h;
        const int lb = task & 15, row = task >> 4;
        const int pl = sPQ[lb][0], ql = sPQ[lb][1];
        const double cl = sPar[lb][0], sl = sPar[lb][1], elr = sPar[lb][2], eli = sPar[lb][3];
        const double xpr = Zc[row * GS + pl * C], xpi = CPLX ? Zc[row * GS + pl * C + 1] : 0.0;
        const double xqr = Zc[row * GS + ql * C], xqi = CPLX ? Zc[row * GS + ql * C + 1] : 0.0;
        Zn[row * GS + pl * C] = cl * xpr - sl * (elr * xqr + eli * xqi);
        Zn[row * GS + ql * C] = sl * (elr * xpr - eli * xpi) + cl * xqr;
        if (CPLX) {
          Zn[row * GS + pl * C + 1] = cl * xpi - sl * (elr * xqi - eli * xqr);
          Zn[row * GS + ql * C + 1] = sl * (elr * xpi + eli * xpr) + cl * xqi;
        }
      }
      __syncthreads();
      double* x = Gc; Gc = Gn; Gn = x;
      x = Zc; Zc = Zn; Zn = x;
    }
    const bool changed = sChanged != 0;
    __syncthreads();
    if (!changed) break;
  }
  double* z = ZW + C * ((k.pair_base + pair) * PJ * PJ);
  for (int e = t; e < PJ * PJ; e += 256) {
    const int r = e / PJ, c = e % PJ;
    z[C * e] = Zc[r * GS + c * C];
    if (CPLX) z[C * e + 1] = Zc[r * GS + c * C + 1];
  }
}

// One round's update: work-group = (pair, row chunk of W or Q): rows
// [r0, r0 + RCA) of the pair's 32 columns times the pair's Z.
template <bool CPLX>
__global__ __launch_bounds__(256) void k_bapply(double* __restrict__ W, double* __restrict__ Q,
                                                const JBlk* __restrict__ blks, int nb, int64_t total, int round,
                                                const double* __restrict__ ZW, const int* __restrict__ ZF) {
  constexpr int C = CPLX ? 2 : 1;
  __shared__ double sZ[PJ * PJ * C];
  __shared__ double sX[PJ * RCA * C];
  const int64_t g = blockIdx.x;
  if (g >= total) return;
  const int b = find_blk(blks, nb, g, BY_APPLY);
  const JBlk k = blks[b];
  if (k.special || round >= k.nplay - 1) return;
  const int64_t li = g - k.apply_base;
  const int pair = (int)(li / k.ncha), chunk = (int)(li % k.ncha);
  if (!ZF[k.pair_base + pair]) return;
  int I, J;
  rr_pair(k.nplay, round, pair, I, J);
  const int t = threadIdx.x;
  const double* z = ZW + C * ((k.pair_base + pair) * PJ * PJ);
  for (int e = t; e < PJ * PJ * C; e += 256) sZ[e] = z[e];
  const int chw = (k.m + RCA - 1) / RCA;
  const bool inW = chunk < chw;
  double* X = inW ? W : Q;
  const int len = inW ? k.m : k.n;
  const int64_t base = inW ? k.w_off : k.q_off;
  const int r0 = (inW ? chunk : chunk - chw) * RCA;
  for (int e = t; e < PJ * RCA; e += 256) {
    const int c = e / RCA, x = e % RCA;
    const int col = pair_col(k, I, J, c);
    const bool ok = col >= 0 && r0 + x < len;
    const int64_t src = base + (int64_t)col * len + r0 + x;
    sX[(c * RCA + x) * C] = ok ? X[C * src] : 0.0;
    if (CPLX) sX[(c * RCA + x) * C + 1] = ok ? X[C * src + 1] : 0.0;
  }
  __syncthreads();
  const int row = t & (RCA - 1), cg = t >> 6;  // 64 rows x 4 groups of 8 columns
  double yr[8] = {}, yi[8] = {};
#pragma unroll 2
  for (int i = 0; i < PJ; ++i) {
    const double xr = sX[(i * RCA + row) * C], xi = CPLX ? sX[(i * RCA + row) * C + 1] : 0.0;
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
      const int j = cg * 8 + jj;
      const double zr = sZ[(i * PJ + j) * C], zi = CPLX ? sZ[(i * PJ + j) * C + 1] : 0.0;
      yr[jj] += xr * zr - xi * zi;
      if (CPLX) yi[jj] += xr * zi + xi * zr;
    }
  }
  if (r0 + row < len) {
#pragma unroll
    for (int jj = 0; jj < 8; ++jj) {
      const int col = pair_col(k, I, J, cg * 8 + jj);
      if (col < 0) continue;
      const int64_t dst = base + (int64_t)col * len + r0 + row;
      X[C * dst] = yr[jj];
      if (CPLX) X[C * dst + 1] = yi[jj];
    }
  }
}

template <bool CPLX>
__global__ __launch_bounds__(256) void k_jnorms(const double* __restrict__ W, const JBlk* __restrict__ blks, int nb,
                                                int64_t total_cols, const int64_t* __restrict__ col_base,
                                                double* __restrict__ sig) {
  const int lane = threadIdx.x & 63;
  const int64_t g = blockIdx.x * (int64_t)4 + (threadIdx.x >> 6);
  if (g >= total_cols) return;
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (col_base[mid] <= g) lo = mid;
    else hi = mid - 1;
  }
  const JBlk k = blks[lo];
  const int j = (int)(g - col_base[lo]);
  const double* w = W + (CPLX ? 2 : 1) * (k.w_off + (int64_t)j * k.m);
  double a = 0;
  for (int x = lane; x < k.m; x += 64) {
    if (CPLX) a += w[2 * x] * w[2 * x] + w[2 * x + 1] * w[2 * x + 1];
    else a += w[x] * w[x];
  }
  a = wave_sum(a);
  if (lane == 0) sig[g] = sqrt(a);
}

struct FinBlk {
  int rows, cols, m, n, transposed, special, nbc, pad;
  int64_t x_off, b_off, u_off, v_off, a_off;
  int64_t c_off;       // offset into the inverse column permutation of the QR
  int64_t col0;        // offset into perm / sig arrays
  int64_t elem_base;   // flat element base (rows*rows + cols*cols)
};

// U, V in the reference layout (A = U diag(S) V), sorted by perm.  With
// W = Q_h R and R^H Qj = X (X columns orthogonal, norms sig):
// W P = (Q_h Qj) diag(sig) (X / sig)^H with the QR column permutation P, and
// B = Q_h [Qj[:, perm]; 0].
template <bool CPLX>
__global__ __launch_bounds__(256) void k_jfinal(const double* __restrict__ X, const double* __restrict__ B,
                                                const double* __restrict__ A, const FinBlk* __restrict__ fb, int nb,
                                                int64_t total, const int* __restrict__ perm,
                                                const double* __restrict__ sig, const int* __restrict__ pinv,
                                                double* __restrict__ U, double* __restrict__ V) {
  constexpr int C = CPLX ? 2 : 1;
  int64_t x = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (x >= total) return;
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (fb[mid].elem_base <= x) lo = mid;
    else hi = mid - 1;
  }
  const FinBlk f = fb[lo];
  int64_t e = x - f.elem_base;
  const int64_t nu = (int64_t)f.rows * f.rows;
  double vr = 0, vi = 0;
  const bool isU = e < nu;
  if (!isU) e -= nu;
  if (!isU && e >= (int64_t)f.cols * f.cols) return;
  const int dim = isU ? f.rows : f.cols;
  const int r = (int)(e / dim), c = (int)(e % dim);  // U[r][c] or V[r][c]
  // vector k (a column of U or a row of V) from X / sig, element i; W P =
  // Q_h R puts element i of the original order at row pinv[i] of X
  auto from_x = [&](int kv, int i) {
    if (kv >= f.n) return;
    const int src = perm[f.col0 + kv];
    const double sg = sig[f.col0 + src];
    if (sg > 0) {
      const int64_t s = f.x_off + (int64_t)src * f.n + pinv[f.c_off + i];
      vr = X[C * s] / sg;
      vi = CPLX ? X[C * s + 1] / sg : 0.0;
    }
  };
  auto from_b = [&](int kv, int i) {
    if (kv >= f.nbc) return;
    const int64_t s = f.b_off + (int64_t)kv * f.m + i;
    vr = B[C * s];
    vi = CPLX ? B[C * s + 1] : 0.0;
  };
  if (f.special == 1) {
    vr = r == c ? 1.0 : 0.0;
  } else if (f.special == 2) {
    if (isU) vr = 1.0;
    else {
      const double ar = CPLX ? A[2 * f.a_off] : A[f.a_off], ai = CPLX ? A[2 * f.a_off + 1] : 0.0;
      const double z = CPLX ? hypot(ar, ai) : fabs(ar);
      vr = ar / z;
      vi = ai / z;
    }
  } else if (isU) {
    if (f.transposed) from_x(c, r);  // U = X / sig
    else from_b(c, r);               // U = B
  } else {
    if (f.transposed) from_b(r, c);  // V = B^H
    else from_x(r, c);               // V = (X / sig)^H
    vi = -vi;
  }
  double* D = isU ? U : V;
  const int64_t out = (isU ? f.u_off : f.v_off) + e;
  D[C * out] = vr;
  if (CPLX) D[C * out + 1] = vi;
}

// Re-orthonormalisation and completion of vectors [good, k1) of X (length L
// each; element (k, i) at off + k*ks + i*is): Householder QR of the first k0
// vectors (X = Q R); vector j in [good, k0) becomes Q e_j times the phase of
// R_jj (the Gram-Schmidt orthonormalisation of the original direction, same
// phase), and vector j in [k0, k1) becomes the complement vector Q e_j.  The
// reflectors live in a workspace of L*k0 elements at w_off, their squared
// norms and the phases of R_jj in 3*k0 doubles at h_off.  The QR runs in
// panels of CPB columns: k_cqr_panel (one workgroup of 1024 threads per task)
// forms the panel's reflectors, k_cqr_trail applies them to the later columns
// (one wavefront per column over the grid); k_complete_apply then builds the
// target vectors, one wavefront each.
struct CompTask {
  int64_t off, ks, is, w_off, h_off;
  int L, k0, k1, good;  // QR of vectors [0,k0); rewrite vectors [good,k1)
};
// four columns or target vectors (one per wavefront) of one task
struct CompItem {
  int task, j0;
};

constexpr int CNT = 1024;
constexpr int CPB = CNT / 64;  // panel width: one wavefront per panel column

// W = the task's kept vectors [0, k0), column-major L x k0
template <bool CPLX>
__global__ __launch_bounds__(256) void k_cqr_copy(const double* __restrict__ X, double* __restrict__ Wk,
                                                   const CompTask* __restrict__ tasks) {
  const CompTask t = tasks[blockIdx.x];
  constexpr int C = CPLX ? 2 : 1;
  double* W = Wk + C * t.w_off;
  for (int64_t e = blockIdx.y * 256 + threadIdx.x; e < (int64_t)t.L * t.k0; e += (int64_t)gridDim.y * 256) {
    const int j = (int)(e / t.L), i = (int)(e % t.L);
    const double* x = X + C * (t.off + (int64_t)j * t.ks + (int64_t)i * t.is);
    W[C * e] = x[0];
    if (CPLX) W[C * e + 1] = x[1];
  }
}

// y <- H_i y for the reflector v = W[i:, i] with |v|^2 = vn2, by one wavefront
template <bool CPLX>
__device__ __forceinline__ void reflect(const double* __restrict__ col, double* __restrict__ y, int i, int L,
                                        double vn2, int lane) {
  double dr = 0, di = 0;  // v^H y
  for (int r = i + lane; r < L; r += 64) {
    if (CPLX) {
      dr += col[2 * r] * y[2 * r] + col[2 * r + 1] * y[2 * r + 1];
      di += col[2 * r] * y[2 * r + 1] - col[2 * r + 1] * y[2 * r];
    } else {
      dr += col[r] * y[r];
    }
  }
  for (int o = 32; o > 0; o >>= 1) { dr += __shfl_xor(dr, o, 64); if (CPLX) di += __shfl_xor(di, o, 64); }
  const double fr = 2 * dr / vn2, fi = 2 * di / vn2;
  for (int r = i + lane; r < L; r += 64) {
    if (CPLX) {
      y[2 * r] -= col[2 * r] * fr - col[2 * r + 1] * fi;
      y[2 * r + 1] -= col[2 * r] * fi + col[2 * r + 1] * fr;
    } else {
      y[r] -= col[r] * fr;
    }
  }
}

// reflectors p0 .. min(p0 + CPB, k0) - 1, each applied at once to the panel's
// later columns
template <bool CPLX>
__global__ __launch_bounds__(CNT) void k_cqr_panel(double* __restrict__ Wk, double* __restrict__ Hk,
                                                    const CompTask* __restrict__ tasks, int p0) {
  const CompTask t = tasks[blockIdx.x];
  const int p1 = min(p0 + CPB, t.k0);
  if (p0 >= p1) return;  // uniform over the block
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = CNT / 64;
  constexpr int C = CPLX ? 2 : 1;
  __shared__ double red[CNT / 64];
  __shared__ double s_vn2;
  double* W = Wk + C * t.w_off;
  double* H = Hk + t.h_off;
  for (int i = p0; i < p1; ++i) {
    double* col = W + C * (int64_t)i * t.L;
    // |x|^2 over rows i..L-1
    double a = 0;
    for (int r = i + tid; r < t.L; r += CNT) a += CPLX ? col[2 * r] * col[2 * r] + col[2 * r + 1] * col[2 * r + 1] : col[r] * col[r];
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    if (lane == 0) red[wave] = a;
    __syncthreads();
    if (tid == 0) {
      double nrm2 = 0;
      for (int q = 0; q < nw; ++q) nrm2 += red[q];
      double nrm = sqrt(nrm2);
      double x0r = col[C * i], x0i = CPLX ? col[C * i + 1] : 0;
      double ax0 = CPLX ? hypot(x0r, x0i) : fabs(x0r);
      double pr = ax0 > 0 ? x0r / ax0 : 1, pi = ax0 > 0 ? x0i / ax0 : 0;  // e^{i arg x0}
      // beta = -e^{i arg x0} |x|;  v = x - beta e_1  =>  v_0 = x0 + e^{i arg x0}|x|
      col[C * i] = x0r + pr * nrm;
      if (CPLX) col[C * i + 1] = x0i + pi * nrm;
      // |v|^2 = 2|x|(|x| + |x0|); R_ii = beta = -e^{i arg x0}|x|
      const double vn2 = 2 * nrm * (nrm + ax0);
      H[i] = vn2;
      H[t.k0 + 2 * i] = -pr;
      H[t.k0 + 2 * i + 1] = -pi;
      s_vn2 = vn2;
    }
    __syncthreads();
    const double vn2 = s_vn2;
    const int c = i + 1 + wave;  // at most CPB - 1 later panel columns, one wavefront each
    if (vn2 > 0 && c < p1) reflect<CPLX>(col, W + C * (int64_t)c * t.L, i, t.L, vn2, lane);
    __syncthreads();
  }
}

// the panel's reflectors applied to column c = j0 + wave of item blockIdx.x,
// for the columns past the panel
template <bool CPLX>
__global__ __launch_bounds__(256) void k_cqr_trail(double* __restrict__ Wk, const double* __restrict__ Hk,
                                                    const CompTask* __restrict__ tasks,
                                                    const CompItem* __restrict__ items, int p0) {
  const CompItem it = items[blockIdx.x];
  const CompTask t = tasks[it.task];
  const int lane = threadIdx.x & 63, c = it.j0 + (threadIdx.x >> 6);
  const int p1 = min(p0 + CPB, t.k0);
  if (c < p1 || c >= t.k0) return;
  constexpr int C = CPLX ? 2 : 1;
  double* W = Wk + C * t.w_off;
  const double* H = Hk + t.h_off;
  double* y = W + C * (int64_t)c * t.L;
  for (int i = p0; i < p1; ++i) {
    const double vn2 = H[i];
    if (vn2 <= 0) continue;
    reflect<CPLX>(W + C * (int64_t)i * t.L, y, i, t.L, vn2, lane);
  }
}

// Q e_j = H_0 H_1 ... H_{k0-1} e_j for j = j0 + wave of item blockIdx.x
template <bool CPLX>
__global__ __launch_bounds__(256) void k_complete_apply(double* __restrict__ X, const double* __restrict__ Wk,
                                                        const double* __restrict__ Hk,
                                                        const CompTask* __restrict__ tasks,
                                                        const CompItem* __restrict__ items) {
  const CompItem it = items[blockIdx.x];
  const CompTask t = tasks[it.task];
  const int lane = threadIdx.x & 63, j = it.j0 + (threadIdx.x >> 6);
  if (j >= t.k1) return;
  constexpr int C = CPLX ? 2 : 1;
  const double* W = Wk + C * t.w_off;
  const double* H = Hk + t.h_off;
  auto xel = [&](int k, int i) -> double* { return X + C * (t.off + (int64_t)k * t.ks + (int64_t)i * t.is); };
  for (int r = lane; r < t.L; r += 64) {
    double* x = xel(j, r);
    x[0] = r == j ? 1.0 : 0.0;
    if (CPLX) x[1] = 0.0;
  }
  for (int i = t.k0 - 1; i >= 0; --i) {
    const double vn2 = H[i];
    if (vn2 <= 0) continue;
    const double* col = W + C * (int64_t)i * t.L;
    double dr = 0, di = 0;
    for (int r = i + lane; r < t.L; r += 64) {
      const double* x = xel(j, r);
      if (CPLX) {
        dr += col[2 * r] * x[0] + col[2 * r + 1] * x[1];
        di += col[2 * r] * x[1] - col[2 * r + 1] * x[0];
      } else {
        dr += col[r] * x[0];
      }
    }
    for (int o = 32; o > 0; o >>= 1) { dr += __shfl_xor(dr, o, 64); if (CPLX) di += __shfl_xor(di, o, 64); }
    const double fr = 2 * dr / vn2, fi = 2 * di / vn2;
    for (int r = i + lane; r < t.L; r += 64) {
      double* x = xel(j, r);
      if (CPLX) {
        x[0] -= col[2 * r] * fr - col[2 * r + 1] * fi;
        x[1] -= col[2 * r] * fi + col[2 * r + 1] * fr;
      } else {
        x[0] -= col[r] * fr;
      }
    }
  }
  if (j < t.k0) {  // keep the phase of the original direction: multiply by R_jj/|R_jj|
    const double qr = H[t.k0 + 2 * j], qi = H[t.k0 + 2 * j + 1];
    for (int r = lane; r < t.L; r += 64) {
      double* x = xel(j, r);
      if (CPLX) {
        double a = x[0], b = x[1];
        x[0] = a * qr - b * qi;
        x[1] = a * qi + b * qr;
      } else {
        x[0] *= qr;
      }
    }
  }
}

template <bool CPLX>
__global__ __launch_bounds__(256) void k_blk_maxabs(const double* __restrict__ A, const int64_t* __restrict__ off,
                                                    const int64_t* __restrict__ len, double* __restrict__ out) {
  const int64_t o = off[blockIdx.x], n = len[blockIdx.x];
  double m = 0;
  for (int64_t e = threadIdx.x; e < n; e += 256)
    m = fmax(m, CPLX ? hypot(A[2 * (o + e)], A[2 * (o + e) + 1]) : fabs(A[o + e]));
  __shared__ double red[256];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

std::vector<double> block_maxabs(bool cplx, const void* A, const std::vector<SvdBlock>& blocks, hipStream_t s) {
  const int nb = (int)blocks.size();
  std::vector<int64_t> off(nb), len(nb);
  for (int b = 0; b < nb; ++b) { off[b] = blocks[b].a_off; len[b] = (int64_t)blocks[b].rows * blocks[b].cols; }
  DevBuf doff(nb * 8), dlen(nb * 8), dout(nb * 8);
  upload(doff.ptr, off.data(), nb * 8, s);
  upload(dlen.ptr, len.data(), nb * 8, s);
  if (cplx) hipLaunchKernelGGL(k_blk_maxabs<true>, dim3(nb), dim3(256), 0, s, (const double*)A,
                               (const int64_t*)doff.ptr, (const int64_t*)dlen.ptr, (double*)dout.ptr);
  else hipLaunchKernelGGL(k_blk_maxabs<false>, dim3(nb), dim3(256), 0, s, (const double*)A,
                          (const int64_t*)doff.ptr, (const int64_t*)dlen.ptr, (double*)dout.ptr);
  DMRGX_HIP(hipGetLastError());
  std::vector<double> out(nb);
  DMRGX_HIP(hipMemcpyAsync(out.data(), dout.ptr, nb * 8, hipMemcpyDeviceToHost, s));
  DMRGX_HIP(hipStreamSynchronize(s));
  return out;
}

}  // namespace

static thread_local int g_last_sweeps = 0;
static thread_local double g_last_seconds = 0;

int block_svd_last_sweeps() { return g_last_sweeps; }
double block_svd_last_seconds() { return g_last_seconds; }

static int block_svd_impl(int dtype, const void* A, void* U, double* S, void* V, const std::vector<SvdBlock>& blocks,
                          const std::vector<int>& need_u, const std::vector<int>& need_v, hipStream_t s);

int block_svd(int dtype, const void* A, void* U, double* S, void* V, const std::vector<SvdBlock>& blocks,
              const std::vector<int>& need_u, const std::vector<int>& need_v, hipStream_t s) {
  const auto t0 = std::chrono::steady_clock::now();
  g_last_sweeps = block_svd_impl(dtype, A, U, S, V, blocks, need_u, need_v, s);
  g_last_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return g_last_sweeps;
}

static int block_svd_impl(int dtype, const void* A, void* U, double* S, void* V, const std::vector<SvdBlock>& blocks,
                          const std::vector<int>& need_u, const std::vector<int>& need_v, hipStream_t s) {
  const bool cplx = dtype == DT_C64;
  const int cw = cplx ? 2 : 1;
  const int nb = (int)blocks.size();
  if (!nb) return 0;
  // qb: QR view (W = A or A^H, m x n); jb: Jacobi view (X = R^H, n x n)
  std::vector<JBlk> qb(nb), jb(nb);
  int64_t woff = 0, xoff = 0, coff = 0, toff = 0, boff = 0, gbase = 0, abase = 0, gpoff = 0, ebase = 0, xbase = 0;
  int64_t pbase = 0;
  int max_m = 0, max_pan = 0, max_nbc = 0, max_trail = 0;
  std::vector<double> amax = block_maxabs(cplx, A, blocks, s);
  for (int b = 0; b < nb; ++b) {
    const SvdBlock& k = blocks[b];
    JBlk& q = qb[b];
    q = JBlk{};
    q.rows = k.rows;
    q.cols = k.cols;
    q.transposed = k.rows < k.cols;
    q.m = q.transposed ? k.cols : k.rows;
    q.n = q.transposed ? k.rows : k.cols;
    // all-zero and 1x1 special cases of src/tensor.cpp:764-782
    q.special = (k.rows * k.cols == 0) ? 1 : (amax[b] <= TOLZERO ? 1 : (k.rows * k.cols == 1 ? 2 : 0));
    q.tol = std::max(k.rows, k.cols) * EPS;
    {
      // sigma_max lies in [amax, sqrt(rows cols) amax]
      const double lim = NEGL * std::sqrt((double)k.rows * k.cols) * amax[b];
      q.negl2 = lim * lim;
    }
    q.a_off = k.a_off;
    q.w_off = woff;
    q.c_off = coff;
    q.t_off = toff;
    q.b_off = boff;
    q.elem_base = ebase;
    q.npan = (q.n + NB - 1) / NB;
    const int nu = need_u.empty() ? q.rows : std::min(need_u[b], q.rows);
    const int nv = need_v.empty() ? q.cols : std::min(need_v[b], q.cols);
    q.nbc = q.special ? 0 : (q.transposed ? nv : nu);  // columns of B = Q_h [Qj; 0]
    woff += (int64_t)q.n * q.m;
    coff += q.n;
    toff += (int64_t)q.npan * NB * NB;
    boff += (int64_t)q.m * q.nbc;
    ebase += (int64_t)q.n * q.m;
    JBlk& j = jb[b];
    j = q;
    j.m = q.n;
    j.w_off = xoff;
    j.q_off = xoff;  // X and Qj have the same shape and offsets in their own arrays
    j.elem_base = xbase;
    j.nbk = (j.n + BJ - 1) / BJ;
    j.nplay = j.nbk + (j.nbk & 1);
    j.nchg = (j.m + RCG - 1) / RCG;
    j.ncha = (j.m + RCA - 1) / RCA + (j.n + RCA - 1) / RCA;
    j.gram_base = gbase;
    j.apply_base = abase;
    j.gp_off = gpoff;
    j.pair_base = pbase;
    xoff += (int64_t)j.n * j.n;
    xbase += 2 * (int64_t)j.n * j.n;
    if (!q.special) {
      const int64_t pairs = j.nplay / 2;
      gbase += pairs * j.nchg;
      abase += pairs * j.ncha;
      gpoff += pairs * j.nchg * PJ * PJ;
      pbase += pairs;
      max_m = std::max(max_m, q.m);
      max_pan = std::max(max_pan, q.npan);
      max_nbc = std::max(max_nbc, q.nbc);
      max_trail = std::max(max_trail, q.n);
    }
  }
  DMRGX_CHECK(max_m <= 2048, "block_svd: sector blocks with more than 2048 rows and columns are not supported");
  DevBuf dW(std::max<int64_t>(woff, 1) * 8 * cw), dX(std::max<int64_t>(xoff, 1) * 8 * cw),
      dQ(std::max<int64_t>(xoff, 1) * 8 * cw), dB(std::max<int64_t>(boff, 1) * 8 * cw);
  DevBuf dBeta(std::max<int64_t>(coff, 1) * 8 * cw), dTau(std::max<int64_t>(coff, 1) * 8),
      dT(std::max<int64_t>(toff, 1) * 8 * cw), dCN(std::max<int64_t>(coff, 1) * 8),
      dPIV(std::max<int64_t>(coff, 1) * 4), dPINV(std::max<int64_t>(coff, 1) * 4), dCB(nb * 8);
  DevBuf dqb(nb * sizeof(JBlk)), djb(nb * sizeof(JBlk)), dflag(nb * 8), dGP(std::max<int64_t>(gpoff, 1) * 8 * cw);
  DevBuf dZW(std::max<int64_t>(pbase, 1) * PJ * PJ * 8 * cw), dZF(std::max<int64_t>(pbase, 1) * 4);
  upload(dqb.ptr, qb.data(), nb * sizeof(JBlk), s);
  upload(djb.ptr, jb.data(), nb * sizeof(JBlk), s);
  const JBlk* QB = (const JBlk*)dqb.ptr;
  const JBlk* JB = (const JBlk*)djb.ptr;
  if (ebase > 0) {
    unsigned g = (unsigned)((ebase + 255) / 256);
    if (cplx) hipLaunchKernelGGL(k_jinit<true>, dim3(g), dim3(256), 0, s, (const double*)A, (double*)dW.ptr, QB, nb,
                                 ebase);
    else hipLaunchKernelGGL(k_jinit<false>, dim3(g), dim3(256), 0, s, (const double*)A, (double*)dW.ptr, QB, nb,
                            ebase);
    DMRGX_HIP(hipGetLastError());
  }
  // 1. Householder QR with panel column pivoting of every block, panel by panel
  if (coff > 0) {
    std::vector<int64_t> cb(nb);
    for (int b = 0; b < nb; ++b) cb[b] = qb[b].c_off;
    upload(dCB.ptr, cb.data(), nb * 8, s);
    unsigned g = (unsigned)((coff + 3) / 4);
    if (cplx) hipLaunchKernelGGL(k_qr_colnorm<true>, dim3(g), dim3(256), 0, s, (const double*)dW.ptr, QB,
                                 (const int64_t*)dCB.ptr, nb, coff, (double*)dCN.ptr, (int*)dPIV.ptr);
    else hipLaunchKernelGGL(k_qr_colnorm<false>, dim3(g), dim3(256), 0, s, (const double*)dW.ptr, QB,
                            (const int64_t*)dCB.ptr, nb, coff, (double*)dCN.ptr, (int*)dPIV.ptr);
    DMRGX_HIP(hipGetLastError());
  }
  const int rpt = max_m <= 256 ? 1 : (max_m <= 512 ? 2 : 4);
  const int pnt = max_m <= 1024 ? 256 : 512;
  for (int kp = 0; kp < max_pan; ++kp) {
    auto panel = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(nb), dim3(pnt), 0, s, (double*)dW.ptr, QB, kp, (double*)dBeta.ptr,
                         (double*)dTau.ptr, (double*)dT.ptr, (double*)dCN.ptr, (int*)dPIV.ptr);
    };
    if (cplx) {
      if (rpt == 1) panel(k_qr_panel<true, 1>);
      else if (rpt == 2) panel(k_qr_panel<true, 2>);
      else panel(k_qr_panel<true, 4>);
    } else {
      if (rpt == 1) panel(k_qr_panel<false, 1>);
      else if (rpt == 2) panel(k_qr_panel<false, 2>);
      else panel(k_qr_panel<false, 4>);
    }
    const int trail = max_trail - (kp + 1) * NB;
    if (trail > 0) {
      const int tiles = (trail + CT - 1) / CT;
      if (cplx) hipLaunchKernelGGL(k_qr_apply<true>, dim3(nb * tiles), dim3(256), 0, s, (const double*)dW.ptr,
                                   (double*)dW.ptr, QB, tiles, kp, (const double*)dT.ptr, 1, (double*)dCN.ptr);
      else hipLaunchKernelGGL(k_qr_apply<false>, dim3(nb * tiles), dim3(256), 0, s, (const double*)dW.ptr,
                              (double*)dW.ptr, QB, tiles, kp, (const double*)dT.ptr, 1, (double*)dCN.ptr);
    }
  }
  DMRGX_HIP(hipGetLastError());
  // inverse of the QR column permutation
  if (coff > 0) {
    std::vector<int> piv(coff), pinv(coff);
    DMRGX_HIP(hipMemcpyAsync(piv.data(), dPIV.ptr, coff * 4, hipMemcpyDeviceToHost, s));
    DMRGX_HIP(hipStreamSynchronize(s));
    for (int b = 0; b < nb; ++b) {
      const int64_t o = qb[b].c_off;
      for (int j = 0; j < qb[b].n; ++j) {
        const int p = piv[o + j];
        DMRGX_CHECK(p >= 0 && p < qb[b].n, "block_svd: corrupt column permutation");
        pinv[o + p] = j;
      }
    }
    upload(dPINV.ptr, pinv.data(), coff * 4, s);
  }
  // 2. X = R^H, Qj = I
  if (xbase > 0) {
    unsigned g = (unsigned)((xbase + 255) / 256);
    if (cplx) hipLaunchKernelGGL(k_qr_rh<true>, dim3(g), dim3(256), 0, s, (const double*)dW.ptr,
                                 (const double*)dBeta.ptr, (double*)dX.ptr, (double*)dQ.ptr, QB, JB, nb, xbase);
    else hipLaunchKernelGGL(k_qr_rh<false>, dim3(g), dim3(256), 0, s, (const double*)dW.ptr,
                            (const double*)dBeta.ptr, (double*)dX.ptr, (double*)dQ.ptr, QB, JB, nb, xbase);
    DMRGX_HIP(hipGetLastError());
  }
  // 3. blocked one-sided Jacobi on X, rotations accumulated in Qj
  int max_rounds = 0;
  for (auto& j : jb)
    if (!j.special) max_rounds = std::max(max_rounds, j.nplay - 1);
  int sweeps = 0;
  std::vector<unsigned long long> flags(nb);
  if (max_rounds > 0 && gbase > 0) {
    DMRGX_CHECK(gbase < (1ll << 31) && abase < (1ll << 31), "block_svd: too many work-groups");
    for (sweeps = 1; sweeps <= MAX_SWEEPS; ++sweeps) {
      fill_zero(dflag.ptr, nb * 8, s);
      for (int r = 0; r < max_rounds; ++r) {
        if (cplx) {
          hipLaunchKernelGGL(k_bgram<true>, dim3((unsigned)gbase), dim3(256), 0, s, (const double*)dX.ptr, JB, nb,
                             gbase, r, (double*)dGP.ptr);
          hipLaunchKernelGGL(k_beig<true>, dim3((unsigned)pbase), dim3(256), 0, s, JB, nb, pbase, r,
                             (const double*)dGP.ptr, (double*)dZW.ptr, (int*)dZF.ptr, (unsigned long long*)dflag.ptr);
          hipLaunchKernelGGL(k_bapply<true>, dim3((unsigned)abase), dim3(256), 0, s, (double*)dX.ptr,
                             (double*)dQ.ptr, JB, nb, abase, r, (const double*)dZW.ptr, (const int*)dZF.ptr);
        } else {
          hipLaunchKernelGGL(k_bgram<false>, dim3((unsigned)gbase), dim3(256), 0, s, (const double*)dX.ptr, JB, nb,
                             gbase, r, (double*)dGP.ptr);
          hipLaunchKernelGGL(k_beig<false>, dim3((unsigned)pbase), dim3(256), 0, s, JB, nb, pbase, r,
                             (const double*)dGP.ptr, (double*)dZW.ptr, (int*)dZF.ptr, (unsigned long long*)dflag.ptr);
          hipLaunchKernelGGL(k_bapply<false>, dim3((unsigned)abase), dim3(256), 0, s, (double*)dX.ptr,
                             (double*)dQ.ptr, JB, nb, abase, r, (const double*)dZW.ptr, (const int*)dZF.ptr);
        }
      }
      DMRGX_HIP(hipGetLastError());
      DMRGX_HIP(hipMemcpyAsync(flags.data(), dflag.ptr, nb * 8, hipMemcpyDeviceToHost, s));
      DMRGX_HIP(hipStreamSynchronize(s));
      bool done = true;
      for (int b = 0; b < nb; ++b) {
        double d;
        memcpy(&d, &flags[b], 8);
        if (!jb[b].special && d > jb[b].tol) done = false;
      }
      if (done) break;
    }
  }
  // 4. singular values (column norms of X) and per-block descending order
  std::vector<int64_t> col_base(nb);
  int64_t ncol = 0;
  for (int b = 0; b < nb; ++b) { col_base[b] = ncol; ncol += jb[b].n; }
  std::vector<double> sig(std::max<int64_t>(ncol, 1), 0.0);
  DevBuf dsig(std::max<int64_t>(ncol, 1) * 8), dcb(nb * 8), dperm(std::max<int64_t>(ncol, 1) * 4);
  if (ncol > 0) {
    upload(dcb.ptr, col_base.data(), nb * 8, s);
    unsigned g = (unsigned)((ncol + 3) / 4);
    if (cplx) hipLaunchKernelGGL(k_jnorms<true>, dim3(g), dim3(256), 0, s, (const double*)dX.ptr, JB, nb, ncol,
                                 (const int64_t*)dcb.ptr, (double*)dsig.ptr);
    else hipLaunchKernelGGL(k_jnorms<false>, dim3(g), dim3(256), 0, s, (const double*)dX.ptr, JB, nb, ncol,
                            (const int64_t*)dcb.ptr, (double*)dsig.ptr);
    DMRGX_HIP(hipGetLastError());
    DMRGX_HIP(hipMemcpyAsync(sig.data(), dsig.ptr, ncol * 8, hipMemcpyDeviceToHost, s));
    DMRGX_HIP(hipStreamSynchronize(s));
  }
  std::vector<int> perm(std::max<int64_t>(ncol, 1), 0);
  std::vector<FinBlk> fb(nb);
  int64_t fbase = 0;
  std::vector<int64_t> bbase(nb);
  for (int b = 0; b < nb; ++b) {
    const JBlk& j = jb[b];
    const JBlk& q = qb[b];
    const int mn = std::min(j.rows, j.cols);
    const int64_t c0 = col_base[b];
    std::vector<int> p(j.n);
    std::iota(p.begin(), p.end(), 0);
    // stable: equal values keep column order
    std::stable_sort(p.begin(), p.end(), [&](int x, int y) { return sig[c0 + x] > sig[c0 + y]; });
    for (int i = 0; i < j.n; ++i) perm[c0 + i] = p[i];
    std::vector<double> sv(mn, 0.0);
    if (j.special == 2) sv[0] = sig[c0];  // |a| of the 1x1 block
    else if (j.special == 0)
      for (int i = 0; i < mn; ++i) sv[i] = sig[c0 + p[i]];
    upload(S + blocks[b].s_off, sv.data(), mn * 8, s);
    FinBlk& f = fb[b];
    f.rows = j.rows;
    f.cols = j.cols;
    f.m = q.m;
    f.n = j.n;
    f.transposed = j.transposed;
    f.special = j.special;
    f.nbc = q.nbc;
    f.pad = 0;
    f.x_off = j.w_off;
    f.b_off = q.b_off;
    f.u_off = blocks[b].u_off;
    f.v_off = blocks[b].v_off;
    f.a_off = blocks[b].a_off;
    f.c_off = q.c_off;
    f.col0 = c0;
    f.elem_base = fbase;
    fbase += (int64_t)j.rows * j.rows + (int64_t)j.cols * j.cols;
    bbase[b] = q.b_off;
  }
  if (ncol > 0) upload(dperm.ptr, perm.data(), ncol * 4, s);
  // 5. B = Q_h [Qj[:, perm]; 0 | e_n ...]: reflectors applied last panel first
  if (boff > 0) {
    DevBuf dbb(nb * 8);
    upload(dbb.ptr, bbase.data(), nb * 8, s);
    unsigned g = (unsigned)((boff + 255) / 256);
    if (cplx) hipLaunchKernelGGL(k_binit<true>, dim3(g), dim3(256), 0, s, (const double*)dQ.ptr, (double*)dB.ptr,
                                 QB, JB, (const int64_t*)dbb.ptr, (const int64_t*)dcb.ptr, (const int*)dperm.ptr, nb,
                                 boff);
    else hipLaunchKernelGGL(k_binit<false>, dim3(g), dim3(256), 0, s, (const double*)dQ.ptr, (double*)dB.ptr,
                            QB, JB, (const int64_t*)dbb.ptr, (const int64_t*)dcb.ptr, (const int*)dperm.ptr, nb,
                            boff);
    DMRGX_HIP(hipGetLastError());
    const int tiles = (max_nbc + CT - 1) / CT;
    for (int kp = max_pan - 1; kp >= 0; --kp) {
      if (cplx) hipLaunchKernelGGL(k_qr_apply<true>, dim3(nb * tiles), dim3(256), 0, s, (const double*)dW.ptr,
                                   (double*)dB.ptr, QB, tiles, kp, (const double*)dT.ptr, 0, (double*)dCN.ptr);
      else hipLaunchKernelGGL(k_qr_apply<false>, dim3(nb * tiles), dim3(256), 0, s, (const double*)dW.ptr,
                              (double*)dB.ptr, QB, tiles, kp, (const double*)dT.ptr, 0, (double*)dCN.ptr);
    }
    DMRGX_HIP(hipGetLastError());
  }
  // 6. U, V in the reference layout
  DevBuf dF(nb * sizeof(FinBlk));
  upload(dF.ptr, fb.data(), nb * sizeof(FinBlk), s);
  if (fbase > 0) {
    unsigned g = (unsigned)((fbase + 255) / 256);
    if (cplx) hipLaunchKernelGGL(k_jfinal<true>, dim3(g), dim3(256), 0, s, (const double*)dX.ptr, (const double*)dB.ptr,
                                 (const double*)A, (const FinBlk*)dF.ptr, nb, fbase, (const int*)dperm.ptr,
                                 (const double*)dsig.ptr, (const int*)dPINV.ptr, (double*)U, (double*)V);
    else hipLaunchKernelGGL(k_jfinal<false>, dim3(g), dim3(256), 0, s, (const double*)dX.ptr, (const double*)dB.ptr,
                            (const double*)A, (const FinBlk*)dF.ptr, nb, fbase, (const int*)dperm.ptr,
                            (const double*)dsig.ptr, (const int*)dPINV.ptr, (double*)U, (double*)V);
    DMRGX_HIP(hipGetLastError());
  }
  // 7. the X / sig side beyond the numerical rank: converged one-sided Jacobi
  // leaves the non-negligible columns of X orthogonal relative to their own
  // norms, so X / sig is orthonormal there; the negligible ones are replaced
  // by the Householder complement of the others.
  std::vector<CompTask> tu, tv;
  int64_t wu = 0, wv = 0, hu = 0, hv = 0;
  for (int b = 0; b < nb; ++b) {
    const JBlk& j = jb[b];
    if (j.special) continue;
    const int64_t c0 = col_base[b];
    const double smax = j.n ? sig[c0 + perm[c0]] : 0.0;
    const double lim = std::max(std::sqrt(j.negl2), 1e-100 * smax);
    int good = 0;
    while (good < j.n && sig[c0 + perm[c0 + good]] > lim) ++good;
    const int nu = need_u.empty() ? j.rows : std::min(need_u[b], j.rows);
    const int nv = need_v.empty() ? j.cols : std::min(need_v[b], j.cols);
    const int need = j.transposed ? nu : nv;  // X / sig gives U when transposed, V otherwise
    if (need > good) {
      const int k0 = good;
      if (j.transposed) {
        tu.push_back(CompTask{blocks[b].u_off, 1, j.rows, wu, hu, j.rows, k0, need, good});
        wu += (int64_t)j.rows * std::max(k0, 1);
        hu += 3 * (int64_t)std::max(k0, 1);
      } else {
        tv.push_back(CompTask{blocks[b].v_off, j.cols, 1, wv, hv, j.cols, k0, need, good});
        wv += (int64_t)j.cols * std::max(k0, 1);
        hv += 3 * (int64_t)std::max(k0, 1);
      }
    }
  }
  auto launch = [&](std::vector<CompTask>& v, int64_t wlen, int64_t hlen, void* Xo) {
    if (v.empty()) return;
    std::vector<CompItem> items, cols;
    int k0max = 0;
    for (int q = 0; q < (int)v.size(); ++q) {
      for (int j = v[q].good; j < v[q].k1; j += 4) items.push_back(CompItem{q, j});
      for (int c = 0; c < v[q].k0; c += 4) cols.push_back(CompItem{q, c});
      k0max = std::max(k0max, v[q].k0);
    }
    DevBuf d(v.size() * sizeof(CompTask)), di(std::max<size_t>(items.size(), 1) * sizeof(CompItem));
    DevBuf dc(std::max<size_t>(cols.size(), 1) * sizeof(CompItem));
    DevBuf wk(std::max<int64_t>(wlen, 1) * 8 * cw), hk(std::max<int64_t>(hlen, 1) * 8);
    upload(d.ptr, v.data(), v.size() * sizeof(CompTask), s);
    if (!items.empty()) upload(di.ptr, items.data(), items.size() * sizeof(CompItem), s);
    if (!cols.empty()) upload(dc.ptr, cols.data(), cols.size() * sizeof(CompItem), s);
    const unsigned nt = (unsigned)v.size();
    auto* T = (const CompTask*)d.ptr;
    if (k0max > 0) {
      if (cplx) hipLaunchKernelGGL(k_cqr_copy<true>, dim3(nt, 64), dim3(256), 0, s, (const double*)Xo,
                                   (double*)wk.ptr, T);
      else hipLaunchKernelGGL(k_cqr_copy<false>, dim3(nt, 64), dim3(256), 0, s, (const double*)Xo,
                              (double*)wk.ptr, T);
      DMRGX_HIP(hipGetLastError());
    }
    for (int p0 = 0; p0 < k0max; p0 += CPB) {
      if (cplx) hipLaunchKernelGGL(k_cqr_panel<true>, dim3(nt), dim3(CNT), 0, s, (double*)wk.ptr, (double*)hk.ptr,
                                   T, p0);
      else hipLaunchKernelGGL(k_cqr_panel<false>, dim3(nt), dim3(CNT), 0, s, (double*)wk.ptr, (double*)hk.ptr,
                              T, p0);
      DMRGX_HIP(hipGetLastError());
      if (p0 + CPB < k0max) {
        if (cplx) hipLaunchKernelGGL(k_cqr_trail<true>, dim3((unsigned)cols.size()), dim3(256), 0, s,
                                     (double*)wk.ptr, (const double*)hk.ptr, T, (const CompItem*)dc.ptr, p0);
        else hipLaunchKernelGGL(k_cqr_trail<false>, dim3((unsigned)cols.size()), dim3(256), 0, s,
                                (double*)wk.ptr, (const double*)hk.ptr, T, (const CompItem*)dc.ptr, p0);
        DMRGX_HIP(hipGetLastError());
      }
    }
    if (!items.empty()) {
      if (cplx) hipLaunchKernelGGL(k_complete_apply<true>, dim3((unsigned)items.size()), dim3(256), 0, s,
                                   (double*)Xo, (const double*)wk.ptr, (const double*)hk.ptr, T,
                                   (const CompItem*)di.ptr);
      else hipLaunchKernelGGL(k_complete_apply<false>, dim3((unsigned)items.size()), dim3(256), 0, s,
                              (double*)Xo, (const double*)wk.ptr, (const double*)hk.ptr, T,
                              (const CompItem*)di.ptr);
      DMRGX_HIP(hipGetLastError());
    }
  };
  launch(tu, wu, hu, U);
  launch(tv, wv, hv, V);
  DMRGX_HIP(hipStreamSynchronize(s));
  return sweeps;
}

}  // namespace dmrgx

using namespace dmrgx;

extern "C" int dmrgx_block_svd_sweeps(void) { return block_svd_last_sweeps(); }

extern "C" int dmrgx_block_svd(int dtype, int n_blk, const int32_t* rows, const int32_t* cols, const void* A,
                               const int64_t* a_off, void* U, const int64_t* u_off, double* S,
                               const int64_t* s_off, void* V, const int64_t* v_off) {
  try {
    ensure_init();
    std::vector<SvdBlock> b(n_blk);
    for (int i = 0; i < n_blk; ++i) {
      DMRGX_CHECK(rows[i] >= 0 && cols[i] >= 0, "block_svd: negative shape");
      b[i] = SvdBlock{rows[i], cols[i], a_off[i], u_off[i], s_off[i], v_off[i]};
    }
    block_svd(dtype, A, U, S, V, b, {}, {}, rt().stream);
    return 0;
  } catch (const dmrgx::Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return -2;
  }
}
