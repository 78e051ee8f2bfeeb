// FP64 thick-restart Lanczos (Wu & Simon) with full reorthogonalisation.
// Krylov basis, reorthogonalisation and restarts stay on the device; only the
// small projected matrix (<= ncv x ncv, real symmetric) is diagonalised on the
// host.  Replaces the ARPACK dsaupd/znaupd reverse-communication loop of the
// reference (src/dsaupd.cpp:65-155, src/znaupd.cpp:78-197) with the same
// wanted set (nev lowest), ncv and tolerance semantics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <type_traits>

#include "common.hpp"
#include "gemm.hpp"
#include "lanczos.hpp"

namespace dmrgx {
namespace {

constexpr int G = 8;  // Krylov columns handled per pass of the dot / combine kernels
constexpr int DOT_BLOCKS_MAX = 4096;  // work-groups of one dot-product launch (all column groups)
constexpr int PARTIAL_DOUBLES = (DOT_BLOCKS_MAX + 64) * 2 * G;

// element type of one complex entry of a T array, loaded as one vector
template <typename T>
using cplx_of = typename std::conditional<std::is_same<T, float>::value, float2, double2>::type;

// Partial <X_c, w> (conjugate-linear in X) for columns c of one group of G:
// X_c = V_c for c < k, and X_k = w itself when with_self (the squared norm).
// V holds TV elements (double, or float for the basis of a low-precision
// solve), w TW elements; the sums are FP64.  The launch is one-dimensional
// with the column group fastest, so the groups that read one range of w run
// side by side; complex entries load as one vector each.
template <bool CPLX, typename TV = double, typename TW = double>
__global__ __launch_bounds__(256) void k_vdots(const TV* __restrict__ V, int64_t ld, int k, int with_self,
                                               const TW* __restrict__ w, int64_t n, int ng,
                                               double* __restrict__ partial) {
  const int gy = blockIdx.x % ng, bx = blockIdx.x / ng, nbx = gridDim.x / ng;
  const int j0 = gy * G;
  const int ktot = k + with_self;
  const int g = min(G, ktot - j0);
  double re[G], im[G];
#pragma unroll
  for (int c = 0; c < G; ++c) re[c] = im[c] = 0;
  int64_t i = bx * (int64_t)blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)nbx * blockDim.x;
  for (; i < n; i += stride) {
    if (CPLX) {
      const cplx_of<TW> wv = reinterpret_cast<const cplx_of<TW>*>(w)[i];
      const double wr = wv.x, wi = wv.y;
#pragma unroll
      for (int c = 0; c < G; ++c)
        if (c < g) {
          double vr = wr, vi = wi;
          if (j0 + c < k) {
            const cplx_of<TV> v = reinterpret_cast<const cplx_of<TV>*>(V + (int64_t)(j0 + c) * ld)[i];
            vr = v.x;
            vi = v.y;
          }
          re[c] += vr * wr + vi * wi;
          im[c] += vr * wi - vi * wr;
        }
    } else {
      const double wr = w[i];
#pragma unroll
      for (int c = 0; c < G; ++c)
        if (c < g) re[c] += ((j0 + c < k) ? (double)V[(int64_t)(j0 + c) * ld + i] : wr) * wr;
    }
  }
  __shared__ double red[2 * G][256];
#pragma unroll
  for (int c = 0; c < G; ++c) {
    red[2 * c][threadIdx.x] = re[c];
    red[2 * c + 1][threadIdx.x] = im[c];
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int q = 0; q < 2 * G; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 2 * G) partial[((int64_t)gy * nbx + bx) * 2 * G + threadIdx.x] = red[threadIdx.x][0];
}

// one workgroup per column group: sum the per-block partials
__global__ void k_reduce_partials(const double* __restrict__ partial, int nblk, int ktot, double* __restrict__ out) {
  const int j0 = blockIdx.x * G;
  const int q = threadIdx.x / 64, lane = threadIdx.x & 63;
  const int g = min(G, ktot - j0);
  if (q >= 2 * g) return;
  double s = 0;
  for (int b = lane; b < nblk; b += 64) s += partial[((int64_t)blockIdx.x * nblk + b) * 2 * G + q];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) out[2 * j0 + q] = s;  // out[2c] = re, out[2c+1] = im
}

// w -= sum_j V_j h_j
template <bool CPLX, typename TV = double, typename TW = double>
__global__ __launch_bounds__(256) void k_gemv_sub(const TV* __restrict__ V, int64_t ld, int k,
                                                  const double* __restrict__ h, TW* __restrict__ w,
                                                  int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    if (CPLX) {
      double ar = 0, ai = 0;
      for (int j = 0; j < k; ++j) {
        const TV* v = V + (int64_t)j * ld;
        const double vr = v[2 * i], vi = v[2 * i + 1], hr = h[2 * j], hi = h[2 * j + 1];
        ar += vr * hr - vi * hi;
        ai += vr * hi + vi * hr;
      }
      w[2 * i] = (TW)((double)w[2 * i] - ar);
      w[2 * i + 1] = (TW)((double)w[2 * i + 1] - ai);
    } else {
      double a = 0;
      for (int j = 0; j < k; ++j) a += (double)V[(int64_t)j * ld + i] * h[2 * j];
      w[i] = (TW)((double)w[i] - a);
    }
  }
}

// out = (w - sum_j V_j h_j) * a
template <bool CPLX, typename TV = double>
__global__ __launch_bounds__(256) void k_gemv_sub_scale(const TV* __restrict__ V, int64_t ld, int k,
                                                        const double* __restrict__ h, const double* __restrict__ w,
                                                        TV* __restrict__ out, int64_t n, double a) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    if (CPLX) {
      double ar = 0, ai = 0;
      for (int j = 0; j < k; ++j) {
        const TV* v = V + (int64_t)j * ld;
        const double vr = v[2 * i], vi = v[2 * i + 1], hr = h[2 * j], hi = h[2 * j + 1];
        ar += vr * hr - vi * hi;
        ai += vr * hi + vi * hr;
      }
      out[2 * i] = (TV)((w[2 * i] - ar) * a);
      out[2 * i + 1] = (TV)((w[2 * i + 1] - ai) * a);
    } else {
      double s = 0;
      for (int j = 0; j < k; ++j) s += (double)V[(int64_t)j * ld + i] * h[2 * j];
      out[i] = (TV)((w[i] - s) * a);
    }
  }
}

// out_c = sum_j V_j Y[j][c] for c in [c0, c0+g), real coefficients Y (m x ky row-major)
template <typename TV = double, typename TO = double>
__global__ __launch_bounds__(256) void k_combine(const TV* __restrict__ V, int64_t ld, int m,
                                                 const double* __restrict__ Y, int ky, int c0, int g,
                                                 TO* __restrict__ out, int64_t ldo, int64_t nr) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < nr; i += stride) {
    double acc[G];
#pragma unroll
    for (int c = 0; c < G; ++c) acc[c] = 0;
    for (int j = 0; j < m; ++j) {
      const double v = V[(int64_t)j * ld + i];
#pragma unroll
      for (int c = 0; c < G; ++c)
        if (c < g) acc[c] += v * Y[j * ky + c0 + c];
    }
#pragma unroll
    for (int c = 0; c < G; ++c)
      if (c < g) out[(int64_t)(c0 + c) * ldo + i] = (TO)acc[c];
  }
}

template <typename T>
__global__ void k_scale(T* __restrict__ v, int64_t nr, double a) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < nr; i += stride) v[i] = (T)((double)v[i] * a);
}

template <typename TS, typename TD>
__global__ void k_scale_copy(const TS* __restrict__ s, TD* __restrict__ d, int64_t nr, double a) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < nr; i += stride) d[i] = (TD)((double)s[i] * a);
}

__device__ __forceinline__ uint64_t splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

template <typename T>
__global__ void k_random(T* __restrict__ v, int64_t nr, uint64_t seed) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < nr; i += stride) {
    uint64_t x = splitmix64(seed ^ splitmix64((uint64_t)i));
    v[i] = (T)((double)(x >> 11) * (2.0 / 9007199254740992.0) - 1.0);
  }
}

inline unsigned grid_for(int64_t n) {
  int64_t b = (n + 255) / 256;
  return (unsigned)std::max<int64_t>(1, std::min<int64_t>(b, 4096));
}

struct Work {
  DevBuf* partial = nullptr;
  DevBuf* h = nullptr;   // 2*(ncv+2) doubles
  ~Work() { delete partial; delete h; }
};

// h[0..k) = V[0..k)^H w, plus h[k] = <w, w> when with_self (device, complex pairs)
template <typename TV, typename TW>
void vdots(const TV* V, int64_t ld, int k, const TW* w, int64_t n, bool cplx, double* h, double* partial,
           hipStream_t s, bool with_self = false) {
  const int ktot = k + (with_self ? 1 : 0);
  if (ktot <= 0) return;
  const int ng = (ktot + G - 1) / G;
  int nblk = (int)std::min<int64_t>(DOT_BLOCKS_MAX / ng + 1, std::max<int64_t>(1, (n + 1023) / 1024));
  nblk = std::max(1, std::min(nblk, DOT_BLOCKS_MAX));
  dim3 grid(nblk * ng);
  if (cplx)
    hipLaunchKernelGGL((k_vdots<true, TV, TW>), grid, dim3(256), 0, s, V, ld, k, with_self ? 1 : 0, w, n, ng,
                       partial);
  else
    hipLaunchKernelGGL((k_vdots<false, TV, TW>), grid, dim3(256), 0, s, V, ld, k, with_self ? 1 : 0, w, n, ng,
                       partial);
  hipLaunchKernelGGL(k_reduce_partials, dim3(ng), dim3(64 * 2 * G), 0, s, partial, nblk, ktot, h);
}

template <typename TV, typename TW>
void gemv_sub(const TV* V, int64_t ld, int k, const double* h, TW* w, int64_t n, bool cplx, hipStream_t s) {
  if (k <= 0) return;
  if (cplx) hipLaunchKernelGGL((k_gemv_sub<true, TV, TW>), dim3(grid_for(n)), dim3(256), 0, s, V, ld, k, h, w, n);
  else hipLaunchKernelGGL((k_gemv_sub<false, TV, TW>), dim3(grid_for(n)), dim3(256), 0, s, V, ld, k, h, w, n);
}

template <typename T>
void dev_random_t(T* v, int64_t n, bool cplx, uint64_t seed, hipStream_t s) {
  int64_t nr = n * (cplx ? 2 : 1);
  hipLaunchKernelGGL(k_random<T>, dim3(grid_for(nr)), dim3(256), 0, s, v, nr, seed);
}

template <typename T>
void dev_scale_t(T* v, int64_t n, bool cplx, double a, hipStream_t s) {
  int64_t nr = n * (cplx ? 2 : 1);
  hipLaunchKernelGGL(k_scale<T>, dim3(grid_for(nr)), dim3(256), 0, s, v, nr, a);
}

template <typename T>
double dev_norm2_t(const T* v, int64_t n, bool cplx, hipStream_t s) {
  DevBuf partial(PARTIAL_DOUBLES * sizeof(double));
  DevBuf h(2 * G * sizeof(double));
  vdots(v, 0, 1, v, n, cplx, (double*)h.ptr, (double*)partial.ptr, s);
  double o[2];
  DMRGX_HIP(hipMemcpyAsync(o, h.ptr, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  DMRGX_HIP(hipStreamSynchronize(s));
  return std::sqrt(std::max(o[0], 0.0));
}

}  // namespace

void dev_random(double* v, int64_t n, bool cplx, uint64_t seed, hipStream_t s) { dev_random_t(v, n, cplx, seed, s); }

void dev_scale(double* v, int64_t n, bool cplx, double a, hipStream_t s) { dev_scale_t(v, n, cplx, a, s); }

void dev_dot(const double* a, const double* b, int64_t n, bool cplx, double out[2], hipStream_t s) {
  DevBuf partial(PARTIAL_DOUBLES * sizeof(double));
  DevBuf h(2 * G * sizeof(double));
  vdots(a, 0, 1, b, n, cplx, (double*)h.ptr, (double*)partial.ptr, s);
  DMRGX_HIP(hipMemcpyAsync(out, h.ptr, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  DMRGX_HIP(hipStreamSynchronize(s));
}

double dev_norm2(const double* v, int64_t n, bool cplx, hipStream_t s) {
  double o[2];
  dev_dot(v, v, n, cplx, o, s);
  return std::sqrt(std::max(o[0], 0.0));
}

void dev_orthogonalise(const double* V, int64_t ld, int k, double* v, int64_t n, bool cplx, hipStream_t s) {
  if (k <= 0) return;
  DevBuf partial(PARTIAL_DOUBLES * sizeof(double));
  DevBuf h(2 * (k + G) * sizeof(double));
  for (int pass = 0; pass < 2; ++pass) {
    vdots(V, ld, k, v, n, cplx, (double*)h.ptr, (double*)partial.ptr, s);
    gemv_sub(V, ld, k, (const double*)h.ptr, v, n, cplx, s);
  }
}

namespace {

// The Krylov iteration of eigsh with basis vectors stored as TV (double, or
// float for a low-precision solve); the operator's input and output, the
// Gram-Schmidt sums and the projected problem stay FP64.
template <typename TV>
EigshResult eigsh_krylov(const MatVec& op, int64_t n, bool cplx, int nev, int ncv, double tol, int max_restart,
                         uint64_t seed, const double* v0, double* evec, int n_vec, bool early) {
  hipStream_t s = rt().stream;
  EigshResult res;
  constexpr bool F64 = std::is_same<TV, double>::value;
  const int64_t nr = n * (cplx ? 2 : 1);
  const int64_t ld = nr;
  const double eps23 = std::pow(std::numeric_limits<double>::epsilon(), 2.0 / 3.0);

  DevBuf Vb((size_t)(ncv + 1) * nr * sizeof(TV)), Tmp((size_t)ncv * nr * sizeof(TV));
  // FP64 operator input (a float basis column widened) and output
  DevBuf win(F64 ? 0 : (size_t)nr * 8), wout(F64 ? 0 : (size_t)nr * 8);
  DevBuf partial(PARTIAL_DOUBLES * sizeof(double));
  DevBuf hbuf(2 * (ncv + 2) * sizeof(double)), h2buf(2 * (ncv + 2) * sizeof(double));
  DevBuf Yb((size_t)ncv * ncv * sizeof(double));
  TV* V = (TV*)Vb.ptr;
  // scratch vector w (the operator's output): the first column of Tmp between
  // restarts for an FP64 basis, its own buffer otherwise
  double* w = F64 ? (double*)Tmp.ptr : (double*)wout.ptr;
  double* h = (double*)hbuf.ptr;
  double* h2 = (double*)h2buf.ptr;

  // start vector
  if (v0) {
    if (F64) DMRGX_HIP(hipMemcpyAsync(V, v0, nr * 8, hipMemcpyDeviceToDevice, s));
    else convert_f64_to_f32(v0, (float*)V, nr, s);
  } else {
    dev_random_t(V, n, cplx, seed, s);
  }
  double nv = dev_norm2_t(V, n, cplx, s);
  DMRGX_CHECK(nv > 0, "eigsh: zero start vector");
  dev_scale_t(V, n, cplx, 1.0 / nv, s);

  const int m = ncv;
  std::vector<double> T((size_t)m * m, 0.0);
  int j0 = 0;
  std::vector<double> theta, Y;
  double beta_last = 0;
  int mm = m;  // basis vectors behind theta / Y (fewer when an early test passed)
  bool done = false;
  std::vector<double> hh(2 * (m + 2)), hh2(2 * (m + 2));
  const int keep = std::min(m - 2, nev + std::max(2, (m - nev) / 2));

  for (int restart = 0; restart <= max_restart; ++restart) {
    res.n_restart = restart;
    for (int j = j0; j < m; ++j) {
      TV* vj = V + (int64_t)j * ld;
      if (F64) {
        op((const double*)vj, w);
      } else {
        convert_f32_to_f64((const float*)vj, (double*)win.ptr, nr, s);
        op((const double*)win.ptr, w);
      }
      res.n_matvec++;
      // classical Gram-Schmidt against v_0..v_j, always two passes ("twice
      // is enough", Giraud, Langou and Rozloznik 2005).  One pass amplifies
      // the basis' loss of orthogonality by about |w| / |w'| per step, and
      // DMRG operators have |alpha| >> beta (a spectrum far from zero), so a
      // single pass loses orthogonality within a dozen steps.  The second
      // pass's |w'|^2 comes with its coefficients; |w''|^2 = |w'|^2 - |h2|^2;
      // both coefficient sets are fetched with one synchronisation and the
      // last subtraction is fused with the normalisation of v_{j+1}.
      vdots(V, ld, j + 1, w, n, cplx, h, (double*)partial.ptr, s, false);
      gemv_sub(V, ld, j + 1, h, w, n, cplx, s);
      vdots(V, ld, j + 1, w, n, cplx, h2, (double*)partial.ptr, s, true);
      DMRGX_HIP(hipMemcpyAsync(hh.data(), h, 2 * (j + 1) * sizeof(double), hipMemcpyDeviceToHost, s));
      DMRGX_HIP(hipMemcpyAsync(hh2.data(), h2, 2 * (j + 2) * sizeof(double), hipMemcpyDeviceToHost, s));
      DMRGX_HIP(hipStreamSynchronize(s));
      const double alpha = hh[2 * j] + hh2[2 * j];
      const double nw2 = hh2[2 * (j + 1)];
      double hsq = 0;
      for (int q = 0; q <= j; ++q) hsq += hh2[2 * q] * hh2[2 * q] + hh2[2 * q + 1] * hh2[2 * q + 1];
      const double b2 = nw2 - hsq;
      double w2 = nw2;  // |w|^2 = |h|^2 + |w'|^2
      for (int q = 0; q <= j; ++q) w2 += hh[2 * q] * hh[2 * q] + hh[2 * q + 1] * hh[2 * q + 1];
      const double* hfin = h2;
      res.n_reorth++;
      double beta = std::sqrt(std::max(b2, 0.0));
      T[(size_t)j * m + j] = alpha;
      if (j + 1 < m) {
        T[(size_t)j * m + j + 1] = T[(size_t)(j + 1) * m + j] = beta;
      }
      beta_last = beta;
      TV* vn = V + (int64_t)(j + 1) * ld;
      if (beta > 1e-13 * std::max(std::fabs(alpha), std::sqrt(w2)) && beta > 1e-300) {
        if (cplx) hipLaunchKernelGGL((k_gemv_sub_scale<true, TV>), dim3(grid_for(n)), dim3(256), 0, s, V, ld, j + 1,
                                     hfin, (const double*)w, vn, n, 1.0 / beta);
        else hipLaunchKernelGGL((k_gemv_sub_scale<false, TV>), dim3(grid_for(n)), dim3(256), 0, s, V, ld, j + 1,
                                hfin, (const double*)w, vn, n, 1.0 / beta);
      } else {
        // invariant subspace: continue with a fresh random direction
        dev_random_t(vn, n, cplx, seed + 7919ull * (j + 1) + restart, s);
        for (int pass = 0; pass < 2; ++pass) {
          vdots(V, ld, j + 1, (const TV*)vn, n, cplx, h, (double*)partial.ptr, s);
          gemv_sub(V, ld, j + 1, h, vn, n, cplx, s);
        }
        double z = dev_norm2_t(vn, n, cplx, s);
        dev_scale_t(vn, n, cplx, 1.0 / z, s);
        if (j + 1 < m) T[(size_t)j * m + j + 1] = T[(size_t)(j + 1) * m + j] = 0;
        beta_last = 0;
      }
      // early test: Rayleigh-Ritz on the leading (j+1) x (j+1) projected
      // matrix after every step (the same test as at the end of a cycle)
      if (early && j + 1 < m && j + 1 > nev) {
        const int k = j + 1;
        std::vector<double> A((size_t)k * k), th, Yk;
        for (int r = 0; r < k; ++r)
          for (int c = 0; c < k; ++c) A[(size_t)r * k + c] = T[(size_t)r * m + c];
        host_symeig(k, A, th, Yk);
        bool ok = true;
        double mr = 0;
        for (int i = 0; i < nev; ++i) {
          double r = std::fabs(beta_last * Yk[(size_t)(k - 1) * k + i]);
          mr = std::max(mr, r);
          if (r > tol * std::max(std::fabs(th[i]), eps23)) ok = false;
        }
        if (ok) {
          theta = th;
          Y = Yk;
          mm = k;
          res.max_resid = mr;
          res.converged = done = true;
          break;
        }
      }
    }
    if (done) break;
    // Rayleigh-Ritz on the projected matrix
    std::vector<double> A = T;
    host_symeig(m, A, theta, Y);
    bool all = true;
    res.max_resid = 0;
    for (int i = 0; i < nev; ++i) {
      double r = std::fabs(beta_last * Y[(size_t)(m - 1) * m + i]);
      res.max_resid = std::max(res.max_resid, r);
      if (r > tol * std::max(std::fabs(theta[i]), eps23)) all = false;
    }
    if (all || restart == max_restart) {
      res.converged = all;
      break;
    }
    // thick restart: keep the `keep` lowest Ritz vectors and the residual direction
    std::vector<double> Yk((size_t)m * keep);
    for (int r = 0; r < m; ++r)
      for (int c = 0; c < keep; ++c) Yk[(size_t)r * keep + c] = Y[(size_t)r * m + c];
    DMRGX_HIP(hipMemcpyAsync(Yb.ptr, Yk.data(), Yk.size() * 8, hipMemcpyHostToDevice, s));
    for (int c0 = 0; c0 < keep; c0 += G)
      hipLaunchKernelGGL((k_combine<TV, TV>), dim3(grid_for(nr)), dim3(256), 0, s, V, ld, m, (const double*)Yb.ptr,
                         keep, c0, std::min(G, keep - c0), (TV*)Tmp.ptr, ld, nr);
    DMRGX_HIP(hipMemcpyAsync(V, Tmp.ptr, (size_t)keep * nr * sizeof(TV), hipMemcpyDeviceToDevice, s));
    DMRGX_HIP(hipMemcpyAsync(V + (int64_t)keep * ld, V + (int64_t)m * ld, nr * sizeof(TV), hipMemcpyDeviceToDevice,
                             s));
    std::fill(T.begin(), T.end(), 0.0);
    for (int i = 0; i < keep; ++i) {
      T[(size_t)i * m + i] = theta[i];
      double sgl = beta_last * Y[(size_t)(m - 1) * m + i];
      T[(size_t)i * m + keep] = T[(size_t)keep * m + i] = sgl;
    }
    j0 = keep;
    // the arrow row is consumed by the first step after the restart: its
    // alpha comes from the Gram-Schmidt coefficients like any other step
  }
  res.evals.assign(theta.begin(), theta.begin() + nev);
  // the n_vec lowest Ritz vectors (FP64), the lowest normalised
  const int nvec = std::max(1, std::min(n_vec, mm));
  std::vector<double> y0((size_t)mm * nvec);
  for (int r = 0; r < mm; ++r)
    for (int c = 0; c < nvec; ++c) y0[(size_t)r * nvec + c] = Y[(size_t)r * mm + c];
  DMRGX_HIP(hipMemcpyAsync(Yb.ptr, y0.data(), y0.size() * 8, hipMemcpyHostToDevice, s));
  for (int c0 = 0; c0 < nvec; c0 += G)
    hipLaunchKernelGGL((k_combine<TV, double>), dim3(grid_for(nr)), dim3(256), 0, s, V, ld, mm,
                       (const double*)Yb.ptr, nvec, c0, std::min(G, nvec - c0), evec, nr, nr);
  double z = dev_norm2_t((const double*)evec, n, cplx, s);
  dev_scale_t(evec, n, cplx, 1.0 / z, s);
  DMRGX_HIP(hipStreamSynchronize(s));
  DMRGX_HIP(hipGetLastError());
  return res;
}

}  // namespace

EigshResult eigsh(const MatVec& op, int64_t n, bool cplx, int nev, int ncv, double tol,
                  int max_restart, uint64_t seed, const double* v0, double* evec, int n_vec, bool early,
                  bool fp32_basis) {
  ensure_init();
  hipStream_t s = rt().stream;
  EigshResult res;
  DMRGX_CHECK(n > 0, "eigsh: empty operator");
  const int64_t nr = n * (cplx ? 2 : 1);
  nev = (int)std::min<int64_t>(nev, n);
  ncv = (int)std::min<int64_t>(std::max(ncv, nev + 2), n);

  if (n <= 64 || ncv <= nev + 1) {
    // materialise the operator column by column and diagonalise on the host
    DevBuf e(nr * 8), y(nr * 8);
    std::vector<double> ycol(nr);
    std::vector<dmrgx::cplx> H((size_t)n * n);
    for (int64_t j = 0; j < n; ++j) {
      fill_zero(e.ptr, nr * 8, s);
      double one = 1.0;
      DMRGX_HIP(hipMemcpyAsync((double*)e.ptr + (cplx ? 2 * j : j), &one, 8, hipMemcpyHostToDevice, s));
      op((const double*)e.ptr, (double*)y.ptr);
      res.n_matvec++;
      DMRGX_HIP(hipMemcpyAsync(ycol.data(), y.ptr, nr * 8, hipMemcpyDeviceToHost, s));
      DMRGX_HIP(hipStreamSynchronize(s));
      for (int64_t i = 0; i < n; ++i)
        H[(size_t)i * n + j] = cplx ? dmrgx::cplx(ycol[2 * i], ycol[2 * i + 1]) : dmrgx::cplx(ycol[i], 0);
    }
    std::vector<double> w;
    std::vector<dmrgx::cplx> V;
    host_hermitian_eig((int)n, H, !cplx, w, V);
    res.evals.assign(w.begin(), w.begin() + nev);
    const int nv = (int)std::min<int64_t>(std::max(n_vec, 1), n);
    std::vector<double> x((size_t)std::max(n_vec, 1) * nr, 0.0);
    for (int c = 0; c < nv; ++c)
      for (int64_t i = 0; i < n; ++i) {
        double* xc = x.data() + (size_t)c * nr;
        if (cplx) { xc[2 * i] = V[(size_t)i * n + c].real(); xc[2 * i + 1] = V[(size_t)i * n + c].imag(); }
        else xc[i] = V[(size_t)i * n + c].real();
      }
    DMRGX_HIP(hipMemcpyAsync(evec, x.data(), x.size() * 8, hipMemcpyHostToDevice, s));
    DMRGX_HIP(hipStreamSynchronize(s));
    res.converged = true;
    return res;
  }

  // a float basis holds orthogonality to about 1e-7: only for solves whose
  // tolerance it does not limit
  if (fp32_basis && tol >= 1e-6)
    return eigsh_krylov<float>(op, n, cplx, nev, ncv, tol, max_restart, seed, v0, evec, n_vec, early);
  return eigsh_krylov<double>(op, n, cplx, nev, ncv, tol, max_restart, seed, v0, evec, n_vec, early);
}

}  // namespace dmrgx
