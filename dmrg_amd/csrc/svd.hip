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
constexpr double JTOL = 1e-15;      // off-diagonal convergence threshold
constexpr int MAX_SWEEPS = 40;

struct JBlk {
  int m, n;         // W has n columns of length m
  int npl;          // players in the round robin (n rounded up to even)
  int transposed;   // W = A^H
  int rows, cols;   // original A shape
  int special;      // 0 normal, 1 all-zero, 2 one-by-one
  int pad;
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
  if (ratio <= JTOL) return;
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

// Orthonormal completion: vectors k in [k0, k1) of X (length L each; element
// (k, i) at off + k*ks + i*is) are re-orthogonalised against 0..k-1 (modified
// Gram-Schmidt, twice); failures are replaced by unit vectors.  One WG per task.
struct CompTask {
  int64_t off, ks, is;
  int L, k0, k1, pad;
};

template <bool CPLX>
__global__ __launch_bounds__(256) void k_complete(double* __restrict__ X, const CompTask* __restrict__ tasks) {
  const CompTask t = tasks[blockIdx.x];
  __shared__ double red[2][256];
  __shared__ double coef[2];
  int next_unit = 0;
  const int tid = threadIdx.x;
  auto el = [&](int k, int i) -> double* { return X + (CPLX ? 2 : 1) * (t.off + (int64_t)k * t.ks + (int64_t)i * t.is); };
  auto dot = [&](int a, int b, double& rr, double& ri) {  // <x_a, x_b>
    double sr = 0, si = 0;
    for (int i = tid; i < t.L; i += 256) {
      double* pa = el(a, i);
      double* pb = el(b, i);
      if (CPLX) { sr += pa[0] * pb[0] + pa[1] * pb[1]; si += pa[0] * pb[1] - pa[1] * pb[0]; }
      else sr += pa[0] * pb[0];
    }
    red[0][tid] = sr;
    red[1][tid] = si;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (tid < s) { red[0][tid] += red[0][tid + s]; red[1][tid] += red[1][tid + s]; }
      __syncthreads();
    }
    rr = red[0][0];
    ri = red[1][0];
    __syncthreads();
  };
  for (int k = t.k0; k < t.k1; ++k) {
    double n0r, n0i;
    dot(k, k, n0r, n0i);
    bool use_unit = !(n0r > 0.25);
    for (int attempt = 0; attempt < t.L + 1; ++attempt) {
      if (use_unit) {
        if (next_unit >= t.L) break;
        for (int i = tid; i < t.L; i += 256) {
          double* p = el(k, i);
          p[0] = i == next_unit ? 1.0 : 0.0;
          if (CPLX) p[1] = 0.0;
        }
        __syncthreads();
        ++next_unit;
      }
      double before;
      {
        double a, b2;
        dot(k, k, a, b2);
        before = a;
      }
      for (int pass = 0; pass < 2; ++pass)
        for (int j = 0; j < k; ++j) {
          double cr, ci;
          dot(j, k, cr, ci);  // <x_j, x_k>
          for (int i = tid; i < t.L; i += 256) {
            double* pk = el(k, i);
            double* pj = el(j, i);
            if (CPLX) {
              pk[0] -= cr * pj[0] - ci * pj[1];
              pk[1] -= cr * pj[1] + ci * pj[0];
            } else {
              pk[0] -= cr * pj[0];
            }
          }
          __syncthreads();
        }
      double a, b2;
      dot(k, k, a, b2);
      // accept unless the candidate lies (numerically) inside the span so far
      if (a > 1e-6 * before && a > 1e-24) {
        double inv = 1.0 / sqrt(a);
        for (int i = tid; i < t.L; i += 256) {
          double* p = el(k, i);
          p[0] *= inv;
          if (CPLX) p[1] *= inv;
        }
        __syncthreads();
        break;
      }
      use_unit = true;
    }
  }
  (void)coef;
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
      double worst = 0;
      for (auto f : flags) {
        double d;
        memcpy(&d, &f, 8);
        worst = std::max(worst, d);
      }
      if (worst <= JTOL) break;
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
  // completion / cleanup of vectors beyond the numerical rank
  std::vector<CompTask> ct;
  for (int b = 0; b < nb; ++b) {
    const JBlk& j = jb[b];
    if (j.special) continue;
    int64_t c0 = col_base[b];
    double smax = j.n ? sig[c0 + perm[c0]] : 0.0;
    int good = 0;
    while (good < j.n && sig[c0 + perm[c0 + good]] > 1e-6 * smax && sig[c0 + perm[c0 + good]] > 0) ++good;
    // vectors built from W: U columns (not transposed) or V rows (transposed)
    int nu = need_u.empty() ? j.rows : std::min(need_u[b], j.rows);
    int nv = need_v.empty() ? j.cols : std::min(need_v[b], j.cols);
    if (!j.transposed) {
      if (nu > good) ct.push_back(CompTask{blocks[b].u_off, 1, j.rows, j.rows, good, nu, 0});
    } else {
      if (nv > good) ct.push_back(CompTask{blocks[b].v_off, j.cols, 1, j.cols, good, nv, 0});
    }
  }
  if (!ct.empty()) {
    // U tasks and V tasks live in different arrays: launch per array
    std::vector<CompTask> tu, tv;
    for (auto& c : ct) (c.ks == 1 ? tu : tv).push_back(c);
    auto launch = [&](std::vector<CompTask>& v, void* X) {
      if (v.empty()) return;
      DevBuf d(v.size() * sizeof(CompTask));
      DMRGX_HIP(hipMemcpyAsync(d.ptr, v.data(), v.size() * sizeof(CompTask), hipMemcpyHostToDevice, s));
      if (cplx) hipLaunchKernelGGL(k_complete<true>, dim3((unsigned)v.size()), dim3(256), 0, s, (double*)X,
                                   (const CompTask*)d.ptr);
      else hipLaunchKernelGGL(k_complete<false>, dim3((unsigned)v.size()), dim3(256), 0, s, (double*)X,
                              (const CompTask*)d.ptr);
      DMRGX_HIP(hipGetLastError());
      DMRGX_HIP(hipStreamSynchronize(s));
    };
    launch(tu, U);
    launch(tv, V);
  }
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
