// Batched FP64 one-sided (Hestenes) Jacobi SVD.
//
// Each block works on W = A (rows >= cols) or W = A^H (rows < cols), stored
// with one contiguous row per column of W, and accumulates the rotations in
// Q.  One launch per round of the round-robin ordering processes every
// column pair of every block (one wavefront per pair), so all sectors of a
// qtensor::svd run together.  Convergence is tested once per sweep from the
// largest |w_p^H w_q| / (|w_p||w_q|) seen.  Replaces LAPACK ?gesvd in the
// reference's per-sector loop (src/qtensor.cpp:1167-1176, src/tensor.cpp:736-785).
#include <hip/hip_runtime.h>

#include <algorithm>
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

struct JBlk {
  int m, n;         // W has n columns of length m
  int npl;          // players in the round robin (n rounded up to even)
  int transposed;   // W = A^H
  int rows, cols;   // original A shape
  int special;      // 0 normal, 1 all-zero, 2 one-by-one
  int pad;
  double tol;
  int64_t a_off, w_off, q_off;
  int64_t pair_base;  // first global pair slot
  int64_t elem_base;  // first flat element for init / finalize kernels (n*m + n*n)
};

__device__ __forceinline__ int find_blk(const JBlk* b, int nb, int64_t x, bool by_pair) {
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    int64_t s = by_pair ? b[mid].pair_base : b[mid].elem_base;
    if (s <= x) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <bool CPLX>
__global__ __launch_bounds__(256) void k_jinit(const double* __restrict__ A, double* __restrict__ W,
                                               double* __restrict__ Q, const JBlk* __restrict__ blks, int nb,
                                               int64_t total) {
  int64_t x = blockIdx.x * (int64_t)256 + threadIdx.x;
  if (x >= total) return;
  const int b = find_blk(blks, nb, x, false);
  const JBlk k = blks[b];
  int64_t e = x - k.elem_base;
  const int64_t nw = (int64_t)k.n * k.m;
  if (e < nw) {
    const int j = (int)(e / k.m), i = (int)(e % k.m);  // W column j, row i
    int64_t src = k.transposed ? k.a_off + (int64_t)j * k.cols + i : k.a_off + (int64_t)i * k.cols + j;
    if (CPLX) {
      W[2 * (k.w_off + e)] = A[2 * src];
      W[2 * (k.w_off + e) + 1] = k.transposed ? -A[2 * src + 1] : A[2 * src + 1];
    } else {
      W[k.w_off + e] = A[src];
    }
  } else {
    e -= nw;
    if (e >= (int64_t)k.n * k.n) return;
    const int r = (int)(e / k.n), c = (int)(e % k.n);
    if (CPLX) {
      Q[2 * (k.q_off + e)] = r == c ? 1.0 : 0.0;
      Q[2 * (k.q_off + e) + 1] = 0.0;
    } else {
      Q[k.q_off + e] = r == c ? 1.0 : 0.0;
    }
  }
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <bool CPLX>
__global__ __launch_bounds__(256) void k_jround(double* __restrict__ W, double* __restrict__ Q,
                                                const JBlk* __restrict__ blks, int nb, int64_t total_pairs,
                                                int round, unsigned long long* __restrict__ off_flag) {
  const int lane = threadIdx.x & 63;
  const int64_t g = blockIdx.x * (int64_t)4 + (threadIdx.x >> 6);
  if (g >= total_pairs) return;
  const int b = find_blk(blks, nb, g, true);
  const JBlk k = blks[b];
  if (k.special) return;
  const int N = k.npl;
  if (round >= N - 1) return;
  const int p = (int)(g - k.pair_base);
  int i, j;
  if (p == 0) { i = N - 1; j = round; }
  else { i = (round + p) % (N - 1); j = (round - p + N - 1) % (N - 1); }
  if (i >= k.n || j >= k.n) return;
  if (i > j) { int t = i; i = j; j = t; }
  double* wi = W + (CPLX ? 2 : 1) * (k.w_off + (int64_t)i * k.m);
  double* wj = W + (CPLX ? 2 : 1) * (k.w_off + (int64_t)j * k.m);
  double al = 0, be = 0, gr = 0, gi = 0;
  for (int x = lane; x < k.m; x += 64) {
    if (CPLX) {
      double ar = wi[2 * x], ai = wi[2 * x + 1], br = wj[2 * x], bi = wj[2 * x + 1];
      al += ar * ar + ai * ai;
      be += br * br + bi * bi;
      gr += ar * br + ai * bi;  // conj(a) * b
      gi += ar * bi - ai * br;
    } else {
      double a = wi[x], c = wj[x];
      al += a * a;
      be += c * c;
      gr += a * c;
    }
  }
  al = wave_sum(al);
  be = wave_sum(be);
  gr = wave_sum(gr);
  if (CPLX) gi = wave_sum(gi);
  const double ag = CPLX ? hypot(gr, gi) : fabs(gr);
  if (al <= 0 || be <= 0 || ag == 0) return;
  const double ratio = ag / sqrt(al * be);
  if (lane == 0) atomicMax(&off_flag[b], (unsigned long long)__double_as_longlong(ratio));
  if (ratio <= k.tol) return;
  // rotation: w_i' = c w_i - s e^{-i phi} w_j,  w_j' = s e^{i phi} w_i + c w_j
  const double zeta = (be - al) / (2 * ag);
  const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
  const double c = 1 / sqrt(1 + t * t), s = c * t;
  const double pr = CPLX ? gr / ag : (gr >= 0 ? 1.0 : -1.0), pi = CPLX ? gi / ag : 0.0;  // e^{i phi}
  for (int x = lane; x < k.m; x += 64) {
    if (CPLX) {
      double ar = wi[2 * x], ai = wi[2 * x + 1], br = wj[2 * x], bi = wj[2 * x + 1];
      // e^{-i phi} w_j = (pr - i pi)(br + i bi)
      double er = pr * br + pi * bi, ei = pr * bi - pi * br;
      // e^{i phi} w_i = (pr + i pi)(ar + i ai)
      double fr = pr * ar - pi * ai, fi = pr * ai + pi * ar;
      wi[2 * x] = c * ar - s * er;
      wi[2 * x + 1] = c * ai - s * ei;
      wj[2 * x] = s * fr + c * br;
      wj[2 * x + 1] = s * fi + c * bi;
    } else {
      double a = wi[x], bb = wj[x];
      wi[x] = c * a - s * pr * bb;
      wj[x] = s * pr * a + c * bb;
    }
  }
  double* qi = Q + (CPLX ? 2 : 1) * (k.q_off + (int64_t)i * k.n);
  double* qj = Q + (CPLX ? 2 : 1) * (k.q_off + (int64_t)j * k.n);
  for (int x = lane; x < k.n; x += 64) {
    if (CPLX) {
      double ar = qi[2 * x], ai = qi[2 * x + 1], br = qj[2 * x], bi = qj[2 * x + 1];
      double er = pr * br + pi * bi, ei = pr * bi - pi * br;
      double fr = pr * ar - pi * ai, fi = pr * ai + pi * ar;
      qi[2 * x] = c * ar - s * er;
      qi[2 * x + 1] = c * ai - s * ei;
      qj[2 * x] = s * fr + c * br;
      qj[2 * x + 1] = s * fi + c * bi;
    } else {
      double a = qi[x], bb = qj[x];
      qi[x] = c * a - s * pr * bb;
      qj[x] = s * pr * a + c * bb;
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
  int rows, cols, m, n, transposed, special;
  int64_t w_off, q_off, u_off, v_off, s_off, a_off;
  int64_t col0;        // offset into perm / sig arrays
  int64_t elem_base;   // flat element base (rows*rows + cols*cols)
};

// U, V in the reference layout (A = U diag(S) V), sorted by perm.
template <bool CPLX>
__global__ __launch_bounds__(256) void k_jfinal(const double* __restrict__ W, const double* __restrict__ Q,
                                                const double* __restrict__ A, const FinBlk* __restrict__ fb, int nb,
                                                int64_t total, const int* __restrict__ perm,
                                                const double* __restrict__ sig, double* __restrict__ U,
                                                double* __restrict__ V) {
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
  bool isU = e < nu;
  int64_t out;
  if (isU) {
    const int r = (int)(e / f.rows), c = (int)(e % f.rows);  // U[r][c]
    out = f.u_off + e;
    if (f.special == 1) { vr = r == c ? 1.0 : 0.0; }
    else if (f.special == 2) { vr = 1.0; }
    else if (!f.transposed) {
      if (c < f.n) {
        int src = perm[f.col0 + c];
        double sg = sig[f.col0 + src];
        if (sg > 0) {
          const double* w = W + (CPLX ? 2 : 1) * (f.w_off + (int64_t)src * f.m);
          vr = (CPLX ? w[2 * r] : w[r]) / sg;
          vi = CPLX ? w[2 * r + 1] / sg : 0;
        }
      }
    } else {
      int src = perm[f.col0 + c];  // U = Q, column c = Q row src
      const double* q = Q + (CPLX ? 2 : 1) * (f.q_off + (int64_t)src * f.n);
      vr = CPLX ? q[2 * r] : q[r];
      vi = CPLX ? q[2 * r + 1] : 0;
    }
  } else {
    e -= nu;
    if (e >= (int64_t)f.cols * f.cols) return;
    const int r = (int)(e / f.cols), c = (int)(e % f.cols);  // V[r][c]
    out = f.v_off + e;
    if (f.special == 1) { vr = r == c ? 1.0 : 0.0; }
    else if (f.special == 2) {
      double ar = CPLX ? A[2 * f.a_off] : A[f.a_off], ai = CPLX ? A[2 * f.a_off + 1] : 0;
      double z = CPLX ? hypot(ar, ai) : fabs(ar);
      vr = ar / z;
      vi = ai / z;
    } else if (!f.transposed) {
      int src = perm[f.col0 + r];  // V = Q^H: row r = conj(Q row src)
      const double* q = Q + (CPLX ? 2 : 1) * (f.q_off + (int64_t)src * f.n);
      vr = CPLX ? q[2 * c] : q[c];
      vi = CPLX ? -q[2 * c + 1] : 0;
    } else {
      if (r < f.n) {
        int src = perm[f.col0 + r];
        double sg = sig[f.col0 + src];
        if (sg > 0) {
          const double* w = W + (CPLX ? 2 : 1) * (f.w_off + (int64_t)src * f.m);
          vr = (CPLX ? w[2 * c] : w[c]) / sg;
          vi = CPLX ? -w[2 * c + 1] / sg : 0;
        }
      }
    }
  }
  double* D = isU ? U : V;
  if (CPLX) { D[2 * out] = vr; D[2 * out + 1] = vi; }
  else D[out] = vr;
}

// Re-orthonormalisation and completion of vectors [good, k1) of X (length L
// each; element (k, i) at off + k*ks + i*is): Householder QR of the first k0
// vectors (X = Q R); vector j in [good, k0) becomes Q e_j times the phase of
// R_jj (the Gram-Schmidt orthonormalisation of the original direction, same
// phase), and vector j in [k0, k1) becomes the complement vector Q e_j.  The
// reflectors live in a workspace of L*k0 elements at w_off.  One workgroup of
// 1024 threads per task; target vectors are built one per wavefront.
struct CompTask {
  int64_t off, ks, is, w_off;
  int L, k0, k1, good;  // QR of vectors [0,k0); rewrite vectors [good,k1)
};

constexpr int CNT = 1024;

template <bool CPLX>
__global__ __launch_bounds__(CNT) void k_complete(double* __restrict__ X, double* __restrict__ Wk,
                                                   const CompTask* __restrict__ tasks) {
  const CompTask t = tasks[blockIdx.x];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = CNT / 64;
  constexpr int C = CPLX ? 2 : 1;
  __shared__ double red[2][CNT / 64];
  __shared__ double vn2_s;
  double* W = Wk + C * t.w_off;  // column-major L x k0 (vector j at W + C*j*L)
  auto xel = [&](int k, int i) -> double* { return X + C * (t.off + (int64_t)k * t.ks + (int64_t)i * t.is); };
  // copy the kept vectors
  for (int64_t e = tid; e < (int64_t)t.L * t.k0; e += CNT) {
    int j = (int)(e / t.L), i = (int)(e % t.L);
    const double* x = xel(j, i);
    W[C * e] = x[0];
    if (CPLX) W[C * e + 1] = x[1];
  }
  __syncthreads();
  // Householder QR; reflector i stored in W[i:, i] with its squared norm in vn2[i] (LDS-free: recomputed)
  extern __shared__ double vn2_all[];
  for (int i = 0; i < t.k0; ++i) {
    double* col = W + C * (int64_t)i * t.L;
    // |x|^2 over rows i..L-1
    double a = 0;
    for (int r = i + tid; r < t.L; r += CNT) a += CPLX ? col[2 * r] * col[2 * r] + col[2 * r + 1] * col[2 * r + 1] : col[r] * col[r];
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    if (lane == 0) red[0][wave] = a;
    __syncthreads();
    if (tid == 0) {
      double nrm2 = 0;
      for (int q = 0; q < nw; ++q) nrm2 += red[0][q];
      double nrm = sqrt(nrm2);
      double x0r = col[C * i], x0i = CPLX ? col[C * i + 1] : 0;
      double ax0 = CPLX ? hypot(x0r, x0i) : fabs(x0r);
      double pr = ax0 > 0 ? x0r / ax0 : 1, pi = ax0 > 0 ? x0i / ax0 : 0;  // e^{i arg x0}
      // beta = -e^{i arg x0} |x|;  v = x - beta e_1  =>  v_0 = x0 + e^{i arg x0}|x|
      col[C * i] = x0r + pr * nrm;
      if (CPLX) col[C * i + 1] = x0i + pi * nrm;
      // |v|^2 = 2|x|(|x| + |x0|); R_ii = beta = -e^{i arg x0}|x|
      vn2_all[i] = 2 * nrm * (nrm + ax0);
      vn2_all[t.k0 + 2 * i] = -pr;
      vn2_all[t.k0 + 2 * i + 1] = -pi;
    }
    __syncthreads();
    const double vn2 = vn2_all[i];
    if (vn2 <= 0) continue;
    // apply H_i to the remaining kept vectors, one wavefront per column
    for (int c = i + 1 + wave; c < t.k0; c += nw) {
      double* y = W + C * (int64_t)c * t.L;
      double dr = 0, di = 0;  // v^H y
      for (int r = i + lane; r < t.L; r += 64) {
        if (CPLX) {
          dr += col[2 * r] * y[2 * r] + col[2 * r + 1] * y[2 * r + 1];
          di += col[2 * r] * y[2 * r + 1] - col[2 * r + 1] * y[2 * r];
        } else {
          dr += col[r] * y[r];
        }
      }
      for (int o = 32; o > 0; o >>= 1) { dr += __shfl_xor(dr, o, 64); if (CPLX) di += __shfl_xor(di, o, 64); }
      const double fr = 2 * dr / vn2, fi = 2 * di / vn2;
      for (int r = i + lane; r < t.L; r += 64) {
        if (CPLX) {
          y[2 * r] -= col[2 * r] * fr - col[2 * r + 1] * fi;
          y[2 * r + 1] -= col[2 * r] * fi + col[2 * r + 1] * fr;
        } else {
          y[r] -= col[r] * fr;
        }
      }
    }
    __syncthreads();
  }
  // Q e_j = H_0 H_1 ... H_{k0-1} e_j for j = good .. k1-1
  for (int j = t.good + wave; j < t.k1; j += nw) {
    for (int r = lane; r < t.L; r += 64) {
      double* x = xel(j, r);
      x[0] = r == j ? 1.0 : 0.0;
      if (CPLX) x[1] = 0.0;
    }
    for (int i = t.k0 - 1; i >= 0; --i) {
      const double vn2 = vn2_all[i];
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
      const double qr = vn2_all[t.k0 + 2 * j], qi = vn2_all[t.k0 + 2 * j + 1];
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
  (void)vn2_s;
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
  DMRGX_HIP(hipMemcpyAsync(doff.ptr, off.data(), nb * 8, hipMemcpyHostToDevice, s));
  DMRGX_HIP(hipMemcpyAsync(dlen.ptr, len.data(), nb * 8, hipMemcpyHostToDevice, s));
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

int block_svd(int dtype, const void* A, void* U, double* S, void* V, const std::vector<SvdBlock>& blocks,
              const std::vector<int>& need_u, const std::vector<int>& need_v, hipStream_t s) {
  const bool cplx = dtype == DT_C64;
  const int cw = cplx ? 2 : 1;
  const int nb = (int)blocks.size();
  if (!nb) return 0;
  // classify blocks: all-zero and 1x1 special cases of src/tensor.cpp:764-782
  std::vector<JBlk> jb(nb);
  int64_t woff = 0, qoff = 0, pbase = 0, ebase = 0;
  std::vector<double> amax = block_maxabs(cplx, A, blocks, s);
  for (int b = 0; b < nb; ++b) {
    const SvdBlock& k = blocks[b];
    JBlk& j = jb[b];
    j.rows = k.rows;
    j.cols = k.cols;
    j.transposed = k.rows < k.cols;
    j.m = j.transposed ? k.cols : k.rows;
    j.n = j.transposed ? k.rows : k.cols;
    j.npl = j.n + (j.n & 1);
    j.special = (k.rows * k.cols == 0) ? 1 : (amax[b] <= TOLZERO ? 1 : (k.rows * k.cols == 1 ? 2 : 0));
    j.tol = std::max(k.rows, k.cols) * EPS;
    j.a_off = k.a_off;
    j.w_off = woff;
    j.q_off = qoff;
    j.pair_base = pbase;
    j.elem_base = ebase;
    woff += (int64_t)j.n * j.m;
    qoff += (int64_t)j.n * j.n;
    pbase += j.npl / 2;
    ebase += (int64_t)j.n * j.m + (int64_t)j.n * j.n;
  }
  DevBuf dW(std::max<int64_t>(woff, 1) * 8 * cw), dQ(std::max<int64_t>(qoff, 1) * 8 * cw);
  DevBuf dB(nb * sizeof(JBlk)), dflag(nb * 8);
  DMRGX_HIP(hipMemcpyAsync(dB.ptr, jb.data(), nb * sizeof(JBlk), hipMemcpyHostToDevice, s));
  if (ebase > 0) {
    unsigned g = (unsigned)((ebase + 255) / 256);
    if (cplx) hipLaunchKernelGGL(k_jinit<true>, dim3(g), dim3(256), 0, s, (const double*)A, (double*)dW.ptr,
                                 (double*)dQ.ptr, (const JBlk*)dB.ptr, nb, ebase);
    else hipLaunchKernelGGL(k_jinit<false>, dim3(g), dim3(256), 0, s, (const double*)A, (double*)dW.ptr,
                            (double*)dQ.ptr, (const JBlk*)dB.ptr, nb, ebase);
    DMRGX_HIP(hipGetLastError());
  }
  int max_rounds = 0;
  for (auto& j : jb)
    if (!j.special) max_rounds = std::max(max_rounds, j.npl - 1);
  int sweeps = 0;
  std::vector<unsigned long long> flags(nb);
  if (max_rounds > 0 && pbase > 0) {
    unsigned g = (unsigned)((pbase + 3) / 4);
    for (sweeps = 1; sweeps <= MAX_SWEEPS; ++sweeps) {
      fill_zero(dflag.ptr, nb * 8, s);
      for (int r = 0; r < max_rounds; ++r) {
        if (cplx) hipLaunchKernelGGL(k_jround<true>, dim3(g), dim3(256), 0, s, (double*)dW.ptr, (double*)dQ.ptr,
                                     (const JBlk*)dB.ptr, nb, pbase, r, (unsigned long long*)dflag.ptr);
        else hipLaunchKernelGGL(k_jround<false>, dim3(g), dim3(256), 0, s, (double*)dW.ptr, (double*)dQ.ptr,
                                (const JBlk*)dB.ptr, nb, pbase, r, (unsigned long long*)dflag.ptr);
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
  // singular values (column norms) and per-block descending order
  std::vector<int64_t> col_base(nb);
  int64_t ncol = 0;
  for (int b = 0; b < nb; ++b) { col_base[b] = ncol; ncol += jb[b].n; }
  std::vector<double> sig(std::max<int64_t>(ncol, 1), 0.0);
  DevBuf dsig(std::max<int64_t>(ncol, 1) * 8), dcb(nb * 8), dperm(std::max<int64_t>(ncol, 1) * 4);
  if (ncol > 0) {
    DMRGX_HIP(hipMemcpyAsync(dcb.ptr, col_base.data(), nb * 8, hipMemcpyHostToDevice, s));
    unsigned g = (unsigned)((ncol + 3) / 4);
    if (cplx) hipLaunchKernelGGL(k_jnorms<true>, dim3(g), dim3(256), 0, s, (const double*)dW.ptr,
                                 (const JBlk*)dB.ptr, nb, ncol, (const int64_t*)dcb.ptr, (double*)dsig.ptr);
    else hipLaunchKernelGGL(k_jnorms<false>, dim3(g), dim3(256), 0, s, (const double*)dW.ptr,
                            (const JBlk*)dB.ptr, nb, ncol, (const int64_t*)dcb.ptr, (double*)dsig.ptr);
    DMRGX_HIP(hipGetLastError());
    DMRGX_HIP(hipMemcpyAsync(sig.data(), dsig.ptr, ncol * 8, hipMemcpyDeviceToHost, s));
    DMRGX_HIP(hipStreamSynchronize(s));
  }
  std::vector<int> perm(std::max<int64_t>(ncol, 1), 0);
  std::vector<double> Sout;
  std::vector<FinBlk> fb(nb);
  int64_t fbase = 0;
  std::vector<int64_t> s_off(nb);
  for (int b = 0; b < nb; ++b) {
    const JBlk& j = jb[b];
    int mn = std::min(j.rows, j.cols);
    int64_t c0 = col_base[b];
    std::vector<int> p(j.n);
    std::iota(p.begin(), p.end(), 0);
    // stable: equal values keep column order
    std::stable_sort(p.begin(), p.end(), [&](int x, int y) { return sig[c0 + x] > sig[c0 + y]; });
    for (int q = 0; q < j.n; ++q) perm[c0 + q] = p[q];
    std::vector<double> sv(mn, 0.0);
    if (j.special == 2) {
      // |a| of the 1x1 block equals its column norm
      sv[0] = sig[c0];
    } else if (j.special == 0) {
      for (int q = 0; q < mn; ++q) sv[q] = sig[c0 + p[q]];
    }
    DMRGX_HIP(hipMemcpyAsync(S + blocks[b].s_off, sv.data(), mn * 8, hipMemcpyHostToDevice, s));
    FinBlk& f = fb[b];
    f.rows = j.rows;
    f.cols = j.cols;
    f.m = j.m;
    f.n = j.n;
    f.transposed = j.transposed;
    f.special = j.special;
    f.w_off = j.w_off;
    f.q_off = j.q_off;
    f.u_off = blocks[b].u_off;
    f.v_off = blocks[b].v_off;
    f.s_off = blocks[b].s_off;
    f.a_off = blocks[b].a_off;
    f.col0 = c0;
    f.elem_base = fbase;
    fbase += (int64_t)j.rows * j.rows + (int64_t)j.cols * j.cols;
  }
  DevBuf dF(nb * sizeof(FinBlk));
  DMRGX_HIP(hipMemcpyAsync(dF.ptr, fb.data(), nb * sizeof(FinBlk), hipMemcpyHostToDevice, s));
  if (ncol > 0) DMRGX_HIP(hipMemcpyAsync(dperm.ptr, perm.data(), ncol * 4, hipMemcpyHostToDevice, s));
  if (fbase > 0) {
    unsigned g = (unsigned)((fbase + 255) / 256);
    if (cplx) hipLaunchKernelGGL(k_jfinal<true>, dim3(g), dim3(256), 0, s, (const double*)dW.ptr, (const double*)dQ.ptr,
                                 (const double*)A, (const FinBlk*)dF.ptr, nb, fbase, (const int*)dperm.ptr,
                                 (const double*)dsig.ptr, (double*)U, (double*)V);
    else hipLaunchKernelGGL(k_jfinal<false>, dim3(g), dim3(256), 0, s, (const double*)dW.ptr, (const double*)dQ.ptr,
                            (const double*)A, (const FinBlk*)dF.ptr, nb, fbase, (const int*)dperm.ptr,
                            (const double*)dsig.ptr, (double*)U, (double*)V);
    DMRGX_HIP(hipGetLastError());
  }
  // completion of vectors beyond the numerical rank (singular values below
  // 1e-8 of the largest, or zero): Householder complement of the kept ones
  std::vector<CompTask> tu, tv;
  int64_t wu = 0, wv = 0;
  int maxk = 0;
  for (int b = 0; b < nb; ++b) {
    const JBlk& j = jb[b];
    if (j.special) continue;
    int64_t c0 = col_base[b];
    double smax = j.n ? sig[c0 + perm[c0]] : 0.0;
    int good = 0;
    while (good < j.n && sig[c0 + perm[c0 + good]] > 1e-8 * smax && sig[c0 + perm[c0 + good]] > 0) ++good;
    int npos = 0;
    while (npos < j.n && sig[c0 + perm[c0 + npos]] > 0) ++npos;
    int nu = need_u.empty() ? j.rows : std::min(need_u[b], j.rows);
    int nv = need_v.empty() ? j.cols : std::min(need_v[b], j.cols);
    const int need = j.transposed ? nv : nu;
    if (need > good) {
      const int k0 = std::min(npos, need);
      if (!j.transposed) {
        tu.push_back(CompTask{blocks[b].u_off, 1, j.rows, wu, j.rows, k0, need, good});
        wu += (int64_t)j.rows * std::max(k0, 1);
      } else {
        tv.push_back(CompTask{blocks[b].v_off, j.cols, 1, wv, j.cols, k0, need, good});
        wv += (int64_t)j.cols * std::max(k0, 1);
      }
      maxk = std::max(maxk, k0);
    }
  }
  auto launch = [&](std::vector<CompTask>& v, int64_t wlen, void* X) {
    if (v.empty()) return;
    DevBuf d(v.size() * sizeof(CompTask)), wk(std::max<int64_t>(wlen, 1) * 8 * cw);
    DMRGX_HIP(hipMemcpyAsync(d.ptr, v.data(), v.size() * sizeof(CompTask), hipMemcpyHostToDevice, s));
    size_t shm = 3 * (size_t)std::max(maxk, 1) * sizeof(double);
    DMRGX_CHECK(shm <= 64 * 1024, "block_svd: too many kept vectors for the completion kernel");
    if (cplx) hipLaunchKernelGGL(k_complete<true>, dim3((unsigned)v.size()), dim3(CNT), shm, s, (double*)X,
                                 (double*)wk.ptr, (const CompTask*)d.ptr);
    else hipLaunchKernelGGL(k_complete<false>, dim3((unsigned)v.size()), dim3(CNT), shm, s, (double*)X,
                            (double*)wk.ptr, (const CompTask*)d.ptr);
    DMRGX_HIP(hipGetLastError());
    DMRGX_HIP(hipStreamSynchronize(s));
  };
  launch(tu, wu, U);
  launch(tv, wv, V);
  DMRGX_HIP(hipStreamSynchronize(s));
  return sweeps;
}

}  // namespace dmrgx

using namespace dmrgx;

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
