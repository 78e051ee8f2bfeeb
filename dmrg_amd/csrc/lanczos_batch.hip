// Thick-restart Lanczos on many independent operators at once.  The chains'
// vectors are concatenated (chain c owns elements [off_c, off_c + n_c) of
// every Krylov vector), one matvec call applies every chain's operator, and
// the Gram-Schmidt reductions and vector updates are segmented by chain in
// single launches.  The projected matrices, Ritz values and restarts are per
// chain on the host.  Each chain follows eigsh's algorithm and stopping test
// (the ARPACK test of src/dsaupd.cpp:65-155 / src/znaupd.cpp:78-197); all
// chains take the same number of steps, until every chain has converged.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <atomic>
#include <functional>
#include <memory>

#include "common.hpp"
#include "gemm.hpp"
#include "host_pool.hpp"
#include "lanczos.hpp"

namespace dmrgx {
namespace {

constexpr int G = 8;      // Krylov columns per pass of the dot / combine kernels
constexpr int COLS = G;   // the same, for code where G names a projected matrix
// elements of one chain handled by one work-group: enough work-groups to
// fill the chip (about 1024), between 256 and 4096 elements each
inline int64_t seg_elems(int64_t total) {
  int64_t c = (total / 1024 + 255) / 256 * 256;
  return std::max<int64_t>(256, std::min<int64_t>(4096, c));
}
constexpr int64_t DENSE_MAX = 64;

// element range [b, e) of the concatenated vector, owned by chain ordinal `chain`
struct SegBlk {
  int64_t b, e;
  int chain, pad;
};

// partial[(y, blk)][2c + {re, im}] = <X_c, w> over the block, X_c = V_{j0+c} (j0+c < k) or w
template <bool CPLX>
__global__ __launch_bounds__(256) void k_vdots_seg(const double* __restrict__ V, int64_t ld, int k,
                                                   const double* __restrict__ w, const SegBlk* __restrict__ blks,
                                                   double* __restrict__ partial) {
  const int j0 = blockIdx.y * G;
  const int g = min(G, k + 1 - j0);
  const SegBlk B = blks[blockIdx.x];
  double re[G], im[G];
#pragma unroll
  for (int c = 0; c < G; ++c) re[c] = im[c] = 0;
  for (int64_t i = B.b + threadIdx.x; i < B.e; i += 256) {
    if (CPLX) {
      const double wr = w[2 * i], wi = w[2 * i + 1];
#pragma unroll
      for (int c = 0; c < G; ++c)
        if (c < g) {
          const double* v = (j0 + c < k) ? V + (int64_t)(j0 + c) * ld : w;
          const double vr = v[2 * i], vi = v[2 * i + 1];
          re[c] += vr * wr + vi * wi;
          im[c] += vr * wi - vi * wr;
        }
    } else {
      const double wr = w[i];
#pragma unroll
      for (int c = 0; c < G; ++c)
        if (c < g) re[c] += ((j0 + c < k) ? V[(int64_t)(j0 + c) * ld + i] : wr) * wr;
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
  if (threadIdx.x < 2 * G)
    partial[((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * 2 * G + threadIdx.x] = red[threadIdx.x][0];
}

// h[chain * hs + 2 j + {re, im}] = sum of the chain's block partials;
// work-group (chain, column group), one wavefront per real component
__global__ void k_reduce_seg(const double* __restrict__ partial, int nblk, int ktot, const int* __restrict__ cb,
                             const int* __restrict__ ce, double* __restrict__ h, int hs) {
  const int chain = blockIdx.x, j0 = blockIdx.y * G;
  const int q = threadIdx.x / 64, lane = threadIdx.x & 63;
  const int g = min(G, ktot - j0);
  if (q >= 2 * g) return;
  double s = 0;
  for (int b = cb[chain] + lane; b < ce[chain]; b += 64) s += partial[((int64_t)blockIdx.y * nblk + b) * 2 * G + q];
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if (lane == 0) h[(int64_t)chain * hs + 2 * j0 + q] = s;
}

// out = (w - sum_{j<k} V_j h_chain[j]) * scale[chain]; out may alias w; scale null = 1
template <bool CPLX>
__global__ __launch_bounds__(256) void k_gemv_sub_seg(const double* __restrict__ V, int64_t ld, int k,
                                                      const double* __restrict__ h, int hs, const double* w,
                                                      double* out, const double* __restrict__ scale,
                                                      const SegBlk* __restrict__ blks) {
  const SegBlk B = blks[blockIdx.x];
  const double* hc = h + (int64_t)B.chain * hs;
  const double a = scale ? scale[B.chain] : 1.0;
  for (int64_t i = B.b + threadIdx.x; i < B.e; i += 256) {
    if (CPLX) {
      double ar = 0, ai = 0;
      for (int j = 0; j < k; ++j) {
        const double* v = V + (int64_t)j * ld;
        const double vr = v[2 * i], vi = v[2 * i + 1], hr = hc[2 * j], hi = hc[2 * j + 1];
        ar += vr * hr - vi * hi;
        ai += vr * hi + vi * hr;
      }
      const double xr = (w[2 * i] - ar) * a, xi = (w[2 * i + 1] - ai) * a;
      out[2 * i] = xr;
      out[2 * i + 1] = xi;
    } else {
      double acc = 0;
      for (int j = 0; j < k; ++j) acc += V[(int64_t)j * ld + i] * hc[2 * j];
      out[i] = (w[i] - acc) * a;
    }
  }
}

// out column (c0 + c) = sum_{j<m} V_j Y_chain[j][c0 + c]; Y_chain is m x ky, real
template <bool CPLX>
__global__ __launch_bounds__(256) void k_combine_seg(const double* __restrict__ V, int64_t ld, int m,
                                                     const double* __restrict__ Y, int ky, int c0, int g,
                                                     double* __restrict__ out, const SegBlk* __restrict__ blks) {
  const SegBlk B = blks[blockIdx.x];
  const double* Yc = Y + (int64_t)B.chain * m * ky;
  constexpr int C = CPLX ? 2 : 1;
  for (int64_t i = C * B.b + threadIdx.x; i < C * B.e; i += 256) {
    double acc[G];
#pragma unroll
    for (int c = 0; c < G; ++c) acc[c] = 0;
    for (int j = 0; j < m; ++j) {
      const double v = V[(int64_t)j * ld + i];
#pragma unroll
      for (int c = 0; c < G; ++c)
        if (c < g) acc[c] += v * Yc[j * ky + c0 + c];
    }
#pragma unroll
    for (int c = 0; c < G; ++c)
      if (c < g) out[(int64_t)(c0 + c) * ld + i] = acc[c];
  }
}

// out column c = sum_{j<m} V_j Z_chain[j][c] for c < g, Z_chain m x g complex
// pairs (the real part only for real vectors)
template <bool CPLX>
__global__ __launch_bounds__(256) void k_combine_z_seg(const double* __restrict__ V, int64_t ld, int m,
                                                       const double* __restrict__ Z, int g,
                                                       double* __restrict__ out, const SegBlk* __restrict__ blks) {
  const SegBlk B = blks[blockIdx.x];
  const double* Zc = Z + (int64_t)B.chain * m * g * 2;
  for (int64_t i = B.b + threadIdx.x; i < B.e; i += 256) {
    double ar[G], ai[G];
#pragma unroll
    for (int c = 0; c < G; ++c) ar[c] = ai[c] = 0;
    for (int j = 0; j < m; ++j) {
      const double* v = V + (int64_t)j * ld;
      if (CPLX) {
        const double vr = v[2 * i], vi = v[2 * i + 1];
#pragma unroll
        for (int c = 0; c < G; ++c)
          if (c < g) {
            const double zr = Zc[2 * (j * g + c)], zi = Zc[2 * (j * g + c) + 1];
            ar[c] += vr * zr - vi * zi;
            ai[c] += vr * zi + vi * zr;
          }
      } else {
        const double vr = v[i];
#pragma unroll
        for (int c = 0; c < G; ++c)
          if (c < g) ar[c] += vr * Zc[2 * (j * g + c)];
      }
    }
#pragma unroll
    for (int c = 0; c < G; ++c)
      if (c < g) {
        if (CPLX) {
          out[(int64_t)c * ld + 2 * i] = ar[c];
          out[(int64_t)c * ld + 2 * i + 1] = ai[c];
        } else {
          out[(int64_t)c * ld + i] = ar[c];
        }
      }
  }
}

// The conjugate-gradient inner solve of the Jacobi-Davidson refinement keeps
// its scalars on the device, one set per chain (cg[field * nc + chain]), so
// that its steps queue without host synchronisations.
enum { CG_THETA, CG_RHO, CG_RR, CG_LIVE, CG_ALPHA, CG_BETA, CG_FIELDS };

// the work-group sums of (re, im) into partial[blk][0..1] (the layout
// k_reduce_seg reads for one column)
__device__ __forceinline__ void seg_store2(double re, double im, double* __restrict__ partial) {
  __shared__ double red[2][256];
  red[0][threadIdx.x] = re;
  red[1][threadIdx.x] = im;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) {
      red[0][threadIdx.x] += red[0][threadIdx.x + s];
      red[1][threadIdx.x] += red[1][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x < 2) partial[(int64_t)blockIdx.x * 2 * G + threadIdx.x] = red[threadIdx.x][0];
}

// k_vdots_seg of x = w - theta_chain d, formed on the fly: partials of
// <V_j, x> for j < k and of <x, x>
template <bool CPLX, typename TV>
__global__ __launch_bounds__(256) void k_vdots_shift_seg(const TV* __restrict__ V, int64_t ld, int k,
                                                         const double* __restrict__ w, const double* __restrict__ d,
                                                         const double* __restrict__ cg, int nc,
                                                         const SegBlk* __restrict__ blks,
                                                         double* __restrict__ partial) {
  const int j0 = blockIdx.y * G;
  const int g = min(G, k + 1 - j0);
  const SegBlk B = blks[blockIdx.x];
  const double th = cg[CG_THETA * nc + B.chain];
  double re[G], im[G];
#pragma unroll
  for (int c = 0; c < G; ++c) re[c] = im[c] = 0;
  for (int64_t i = B.b + threadIdx.x; i < B.e; i += 256) {
    if (CPLX) {
      const double xr = w[2 * i] - th * d[2 * i], xi = w[2 * i + 1] - th * d[2 * i + 1];
#pragma unroll
      for (int c = 0; c < G; ++c)
        if (c < g) {
          double vr = xr, vi = xi;
          if (j0 + c < k) {
            const TV* v = V + (int64_t)(j0 + c) * ld;
            vr = v[2 * i];
            vi = v[2 * i + 1];
          }
          re[c] += vr * xr + vi * xi;
          im[c] += vr * xi - vi * xr;
        }
    } else {
      const double xr = w[i] - th * d[i];
#pragma unroll
      for (int c = 0; c < G; ++c)
        if (c < g) re[c] += ((j0 + c < k) ? (double)V[(int64_t)(j0 + c) * ld + i] : xr) * xr;
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
  if (threadIdx.x < 2 * G)
    partial[((int64_t)blockIdx.y * gridDim.x + blockIdx.x) * 2 * G + threadIdx.x] = red[threadIdx.x][0];
}

// w = P (w - theta d) = (w - theta d) - sum_{j<k} V_j h_j (h = V^H (w - theta d)),
// with the partials of <d, w> after it
template <bool CPLX, typename TV>
__global__ __launch_bounds__(256) void k_cg_project(const TV* __restrict__ V, int64_t ld, int k,
                                                    const double* __restrict__ h, int hs, double* __restrict__ w,
                                                    const double* __restrict__ d, const double* __restrict__ cg,
                                                    int nc, const SegBlk* __restrict__ blks,
                                                    double* __restrict__ partial) {
  const SegBlk B = blks[blockIdx.x];
  const double* hc = h + (int64_t)B.chain * hs;
  const double th = cg[CG_THETA * nc + B.chain];
  double re = 0, im = 0;
  for (int64_t i = B.b + threadIdx.x; i < B.e; i += 256) {
    if (CPLX) {
      double ar = 0, ai = 0;
      for (int j = 0; j < k; ++j) {
        const TV* v = V + (int64_t)j * ld;
        const double vr = v[2 * i], vi = v[2 * i + 1], hr = hc[2 * j], hi = hc[2 * j + 1];
        ar += vr * hr - vi * hi;
        ai += vr * hi + vi * hr;
      }
      const double dr = d[2 * i], di = d[2 * i + 1];
      const double xr = (w[2 * i] - th * dr) - ar, xi = (w[2 * i + 1] - th * di) - ai;
      w[2 * i] = xr;
      w[2 * i + 1] = xi;
      re += dr * xr + di * xi;
      im += dr * xi - di * xr;
    } else {
      double acc = 0;
      for (int j = 0; j < k; ++j) acc += (double)V[(int64_t)j * ld + i] * hc[2 * j];
      const double x = (w[i] - th * d[i]) - acc;
      w[i] = x;
      re += d[i] * x;
    }
  }
  seg_store2(re, im, partial);
}

// per chain: alpha = rho / <d, w> while the chain is live and <d, w> > 0,
// else alpha = 0 and the chain stops
__global__ __launch_bounds__(256) void k_cg_alpha(int nc, const double* __restrict__ dw, int hs,
                                                  double* __restrict__ cg) {
  for (int c = threadIdx.x; c < nc; c += 256) {
    const double x = dw[(int64_t)c * hs];
    const bool live = cg[CG_LIVE * nc + c] != 0;
    if (live && x > 0) {
      cg[CG_ALPHA * nc + c] = cg[CG_RHO * nc + c] / x;
    } else {
      cg[CG_ALPHA * nc + c] = 0;
      cg[CG_LIVE * nc + c] = 0;
    }
  }
}

// s += alpha d, res -= alpha w, with the partials of |res|^2 after it
template <bool CPLX>
__global__ __launch_bounds__(256) void k_cg_update(const double* __restrict__ d, double* __restrict__ sv,
                                                   const double* __restrict__ w, double* __restrict__ res,
                                                   const double* __restrict__ cg, int nc,
                                                   const SegBlk* __restrict__ blks, double* __restrict__ partial) {
  constexpr int C = CPLX ? 2 : 1;
  const SegBlk B = blks[blockIdx.x];
  const double a = cg[CG_ALPHA * nc + B.chain];
  double acc = 0;
  for (int64_t i = C * B.b + threadIdx.x; i < C * B.e; i += 256) {
    sv[i] += a * d[i];
    const double r = res[i] - a * w[i];
    res[i] = r;
    acc += r * r;
  }
  seg_store2(acc, 0.0, partial);
}

// per chain: beta = |res|^2 / rho (0 once the chain stopped), rho = |res|^2,
// and the chain stops once |res|^2 <= 1e-6 |r|^2; live[0] = the chains still live
__global__ __launch_bounds__(256) void k_cg_beta(int nc, const double* __restrict__ rn, int hs,
                                                 double* __restrict__ cg, int* __restrict__ live_out) {
  __shared__ int cnt[256];
  int n = 0;
  for (int c = threadIdx.x; c < nc; c += 256) {
    const double x = fmax(rn[(int64_t)c * hs], 0.0);
    const double rho = cg[CG_RHO * nc + c];
    bool live = cg[CG_LIVE * nc + c] != 0;
    cg[CG_BETA * nc + c] = (live && rho > 0) ? x / rho : 0.0;
    cg[CG_RHO * nc + c] = x;
    if (live && x <= 1e-6 * cg[CG_RR * nc + c]) live = false;
    cg[CG_LIVE * nc + c] = live ? 1.0 : 0.0;
    n += live ? 1 : 0;
  }
  cnt[threadIdx.x] = n;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) cnt[threadIdx.x] += cnt[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) live_out[0] = cnt[0];
}

// dn = res + beta d
template <bool CPLX>
__global__ __launch_bounds__(256) void k_cg_dir(const double* __restrict__ res, const double* __restrict__ d,
                                                double* __restrict__ dn, const double* __restrict__ cg, int nc,
                                                const SegBlk* __restrict__ blks) {
  constexpr int C = CPLX ? 2 : 1;
  const SegBlk B = blks[blockIdx.x];
  const double b = cg[CG_BETA * nc + B.chain];
  for (int64_t i = C * B.b + threadIdx.x; i < C * B.e; i += 256) dn[i] = res[i] + b * d[i];
}

// x holds e_j (and no longer e_{j-1}) in every listed chain with j < len;
// one thread per listed chain
__global__ void k_unit_seg(double* __restrict__ x, const int64_t* __restrict__ off, const int64_t* __restrict__ len,
                           int nl, int j, int cw) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nl) return;
  if (j > 0 && j - 1 < len[c]) x[cw * (off[c] + j - 1)] = 0;
  if (j < len[c]) x[cw * (off[c] + j)] = 1;
}

// H[c][i][j] (complex pairs) = y_c[i] for the listed chains; work-group per chain
__global__ void k_gather_col(const double* __restrict__ y, const int64_t* __restrict__ off,
                             const int64_t* __restrict__ len, int nd, int j, int cw, double* __restrict__ H) {
  const int c = blockIdx.x;
  for (int i = threadIdx.x; i < len[c]; i += blockDim.x) {
    double* h = H + (((int64_t)c * nd + i) * nd + j) * 2;
    h[0] = y[cw * (off[c] + i)];
    h[1] = cw == 2 ? y[cw * (off[c] + i) + 1] : 0.0;
  }
}

// Executable graphs indexed by Krylov position: slot j holds the launches of
// step j, captured from the stream at first use and replayed afterwards.  The
// captured launches read their scalars from device or pinned memory written
// before each replay, so a replay equals the eager sequence.
struct GraphSet {
  struct Slot {
    hipStream_t s = nullptr;
    hipGraphExec_t exec = nullptr;
    void run(const std::function<void()>& body) {
      if (!exec) {
        hipGraph_t g = nullptr;
        DMRGX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        try {
          body();
        } catch (...) {
          (void)hipStreamEndCapture(s, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        DMRGX_HIP(hipStreamEndCapture(s, &g));
        const hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        DMRGX_HIP(e);
      }
      DMRGX_HIP(hipGraphLaunch(exec, s));
    }
    ~Slot() {
      if (exec) (void)hipGraphExecDestroy(exec);
    }
  };
  std::vector<Slot> slot;
  GraphSet(int n, hipStream_t s) : slot(n) {
    for (auto& x : slot) x.s = s;
  }
  Slot& operator[](int j) { return slot[j]; }
};

// The segments of a set of chains: work-group blocks of at most CH elements
// and each chain's range of blocks; `chain` in a block is the ordinal within
// the set.
struct Segs {
  int n = 0, nblk = 0;
  std::unique_ptr<DevBuf> blk, cb, ce;
  Segs(const std::vector<int64_t>& off, const std::vector<int64_t>& len, const std::vector<int>& set,
       hipStream_t s) {
    n = (int)set.size();
    int64_t total = 0;
    for (int c : set) total += len[c];
    const int64_t CH = seg_elems(total);
    std::vector<SegBlk> b;
    std::vector<int> hb(n), he(n);
    for (int a = 0; a < n; ++a) {
      const int c = set[a];
      hb[a] = (int)b.size();
      for (int64_t x = off[c]; x < off[c] + len[c]; x += CH)
        b.push_back(SegBlk{x, std::min<int64_t>(x + CH, off[c] + len[c]), a, 0});
      he[a] = (int)b.size();
    }
    nblk = (int)b.size();
    blk = std::make_unique<DevBuf>(std::max(nblk, 1) * sizeof(SegBlk));
    cb = std::make_unique<DevBuf>(std::max(n, 1) * 4);
    ce = std::make_unique<DevBuf>(std::max(n, 1) * 4);
    DMRGX_HIP(hipMemcpyAsync(blk->ptr, b.data(), nblk * sizeof(SegBlk), hipMemcpyHostToDevice, s));
    DMRGX_HIP(hipMemcpyAsync(cb->ptr, hb.data(), n * 4, hipMemcpyHostToDevice, s));
    DMRGX_HIP(hipMemcpyAsync(ce->ptr, he.data(), n * 4, hipMemcpyHostToDevice, s));
    DMRGX_HIP(hipStreamSynchronize(s));  // the host arrays end here
  }
  const SegBlk* B() const { return (const SegBlk*)blk->ptr; }
};

// Segmented Gram-Schmidt helpers over one set of chains; coefficient arrays
// hold hs doubles (complex pairs) per chain.
struct SegOps {
  const Segs& sg;
  bool cplx;
  int64_t ld;
  int hs;
  hipStream_t s;
  DevBuf partial;
  int sets;
  double* pin = nullptr;  // pinned staging: `sets` coefficient sets, then one scale per chain
  SegOps(const Segs& g, bool c, int64_t l, int maxk, hipStream_t st, int nsets = 2)
      : sg(g), cplx(c), ld(l), hs(2 * (maxk + G)), s(st),
        partial((size_t)std::max(g.nblk, 1) * ((maxk + G) / G + 1) * 2 * G * 8), sets(std::max(nsets, 2)) {
    DMRGX_HIP(hipHostMalloc((void**)&pin, ((size_t)sets * sg.n * hs + sg.n) * 8, hipHostMallocDefault));
  }
  ~SegOps() {
    if (pin) (void)hipHostFree(pin);
  }
  SegOps(const SegOps&) = delete;
  SegOps& operator=(const SegOps&) = delete;
  // h[c] = (V[0..k)^H x_c, |x_c|^2), k <= maxk - 1
  void vdots(const double* V, int k, const double* x, double* h) {
    const int ng = (k + 1 + G - 1) / G;
    if (cplx) hipLaunchKernelGGL(k_vdots_seg<true>, dim3(sg.nblk, ng), dim3(256), 0, s, V, ld, k, x, sg.B(),
                                 (double*)partial.ptr);
    else hipLaunchKernelGGL(k_vdots_seg<false>, dim3(sg.nblk, ng), dim3(256), 0, s, V, ld, k, x, sg.B(),
                            (double*)partial.ptr);
    hipLaunchKernelGGL(k_reduce_seg, dim3(sg.n, ng), dim3(64 * 2 * G), 0, s, (const double*)partial.ptr, sg.nblk,
                       k + 1, (const int*)sg.cb->ptr, (const int*)sg.ce->ptr, h, hs);
  }
  void sub(const double* V, int k, const double* h, const double* x, double* out, const double* scale) {
    if (cplx) hipLaunchKernelGGL(k_gemv_sub_seg<true>, dim3(sg.nblk), dim3(256), 0, s, V, ld, k, h, hs, x, out,
                                 scale, sg.B());
    else hipLaunchKernelGGL(k_gemv_sub_seg<false>, dim3(sg.nblk), dim3(256), 0, s, V, ld, k, h, hs, x, out, scale,
                            sg.B());
  }
  // out columns [0, ky) = V[0..m) Y_c with Y on the device (chains x m x ky)
  void combine(const double* V, int m, const double* Y, int ky, double* out) {
    for (int c0 = 0; c0 < ky; c0 += G) {
      if (cplx) hipLaunchKernelGGL(k_combine_seg<true>, dim3(sg.nblk), dim3(256), 0, s, V, ld, m, Y, ky, c0,
                                   std::min(G, ky - c0), out, sg.B());
      else hipLaunchKernelGGL(k_combine_seg<false>, dim3(sg.nblk), dim3(256), 0, s, V, ld, m, Y, ky, c0,
                              std::min(G, ky - c0), out, sg.B());
    }
  }
  // out columns [0, g) = V[0..m) Z_c, Z on the device (chains x m x g complex pairs), g <= 8
  void combine_z(const double* V, int m, const double* Z, int g, double* out) {
    if (cplx) hipLaunchKernelGGL(k_combine_z_seg<true>, dim3(sg.nblk), dim3(256), 0, s, V, ld, m, Z, g, out, sg.B());
    else hipLaunchKernelGGL(k_combine_z_seg<false>, dim3(sg.nblk), dim3(256), 0, s, V, ld, m, Z, g, out, sg.B());
  }
  // the coefficient arrays of `count` (<= sets) consecutive sets starting at
  // h, in pinned memory valid until the next fetch
  const double* fetch(const double* h, int count = 1) {
    DMRGX_CHECK(count <= sets, "batched Lanczos: staging too small");
    DMRGX_HIP(hipMemcpyAsync(pin, h, (size_t)count * sg.n * hs * 8, hipMemcpyDeviceToHost, s));
    DMRGX_HIP(hipStreamSynchronize(s));
    return pin;
  }
  // the copy of fetch without the synchronisation (for graph capture)
  void fetch_async(const double* h, int count) {
    DMRGX_CHECK(count <= sets, "batched Lanczos: staging too small");
    DMRGX_HIP(hipMemcpyAsync(pin, h, (size_t)count * sg.n * hs * 8, hipMemcpyDeviceToHost, s));
  }
  // put() in two halves: the host write into pinned staging, then the copy
  void stage(const std::vector<double>& sc) { std::copy(sc.begin(), sc.begin() + sg.n, pin + (size_t)sets * sg.n * hs); }
  void put_staged(double* d) {
    DMRGX_HIP(hipMemcpyAsync(d, pin + (size_t)sets * sg.n * hs, sg.n * 8, hipMemcpyHostToDevice, s));
  }
  // one scale per chain to the device array d
  void put(const std::vector<double>& sc, double* d) {
    double* p = pin + (size_t)sets * sg.n * hs;
    std::copy(sc.begin(), sc.begin() + sg.n, p);
    DMRGX_HIP(hipMemcpyAsync(d, p, sg.n * 8, hipMemcpyHostToDevice, s));
  }
  // x_c /= |x_c| for every chain of the set (a zero segment stays zero when
  // allow_zero, else it is an error)
  void normalise(double* x, double* h, double* scl, bool allow_zero = false) {
    vdots(nullptr, 0, x, h);
    const double* hh = fetch(h);
    std::vector<double> sc(sg.n);
    for (int a = 0; a < sg.n; ++a) {
      const double n2 = hh[(size_t)a * hs];
      DMRGX_CHECK(n2 > 0 || allow_zero, "batched Lanczos: zero vector");
      sc[a] = n2 > 0 ? 1.0 / std::sqrt(n2) : 0.0;
    }
    put(sc, scl);
    sub(nullptr, 0, h, x, x, scl);
    DMRGX_HIP(hipStreamSynchronize(s));
  }
};

std::vector<int64_t> offsets(const std::vector<int64_t>& len) {
  std::vector<int64_t> off(len.size());
  int64_t o = 0;
  for (size_t c = 0; c < len.size(); ++c) {
    DMRGX_CHECK(len[c] > 0, "batched Lanczos: empty chain");
    off[c] = o;
    o += len[c];
  }
  return off;
}

std::atomic<bool> g_graphs{false};

}  // namespace

bool lanczos_graphs() { return g_graphs.load(); }
void set_lanczos_graphs(bool on) { g_graphs.store(on); }

std::vector<EigshResult> eigsh_batched(const MatVec& op, const std::vector<int64_t>& len, bool cplx, int nev,
                                       int ncv, double tol, int max_restart, const std::vector<uint64_t>& seed,
                                       const double* v0, double* evec, int n_vec, const RetireFn& retire) {
  ensure_init();
  hipStream_t s = rt().stream;
  const int nc = (int)len.size();
  DMRGX_CHECK(nc > 0 && (int)seed.size() == nc, "eigsh_batched: one length and one seed per chain");
  const std::vector<int64_t> off = offsets(len);
  const int64_t N = off.back() + len.back();
  const int cw = cplx ? 2 : 1;
  const int64_t nr = N * cw, ld = nr;
  const double eps23 = std::pow(std::numeric_limits<double>::epsilon(), 2.0 / 3.0);
  std::vector<EigshResult> res(nc);

  // chains diagonalised densely (eigsh's small-operator branch) and Lanczos chains
  std::vector<int> small, big;
  for (int c = 0; c < nc; ++c) {
    const int64_t nev_c = std::min<int64_t>(nev, len[c]);
    const int64_t ncv_c = std::min<int64_t>(std::max<int64_t>(ncv, nev_c + 2), len[c]);
    (len[c] <= DENSE_MAX || ncv_c <= nev_c + 1 ? small : big).push_back(c);
  }
  if (!small.empty()) {
    // materialise each small chain's operator column by column, all chains
    // through the same matvec (the other chains see zero vectors)
    const int ns = (int)small.size();
    std::vector<int64_t> so(ns), sl(ns);
    int nd = 0;
    for (int a = 0; a < ns; ++a) {
      so[a] = off[small[a]];
      sl[a] = len[small[a]];
      nd = std::max<int>(nd, (int)sl[a]);
    }
    DevBuf xb(nr * 8), yb(nr * 8), dso(ns * 8), dsl(ns * 8), Hd((size_t)ns * nd * nd * 16);
    DMRGX_HIP(hipMemcpyAsync(dso.ptr, so.data(), ns * 8, hipMemcpyHostToDevice, s));
    DMRGX_HIP(hipMemcpyAsync(dsl.ptr, sl.data(), ns * 8, hipMemcpyHostToDevice, s));
    fill_zero(xb.ptr, nr * 8, s);
    for (int j = 0; j < nd; ++j) {
      hipLaunchKernelGGL(k_unit_seg, dim3((ns + 63) / 64), dim3(64), 0, s, (double*)xb.ptr,
                         (const int64_t*)dso.ptr, (const int64_t*)dsl.ptr, ns, j, cw);
      op((const double*)xb.ptr, (double*)yb.ptr);
      hipLaunchKernelGGL(k_gather_col, dim3(ns), dim3(64), 0, s, (const double*)yb.ptr, (const int64_t*)dso.ptr,
                         (const int64_t*)dsl.ptr, nd, j, cw, (double*)Hd.ptr);
    }
    std::vector<double> Hh((size_t)ns * nd * nd * 2);
    DMRGX_HIP(hipMemcpyAsync(Hh.data(), Hd.ptr, Hh.size() * 8, hipMemcpyDeviceToHost, s));
    DMRGX_HIP(hipStreamSynchronize(s));
    const int nvs = std::max(n_vec, 1);
    std::vector<double> x((size_t)nvs * nr, 0.0);
    for (int a = 0; a < ns; ++a) {
      const int c = small[a], n = (int)sl[a];
      std::vector<dmrgx::cplx> H((size_t)n * n);
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
          const double* q = Hh.data() + (((size_t)a * nd + i) * nd + j) * 2;
          H[(size_t)i * n + j] = dmrgx::cplx(q[0], q[1]);
        }
      std::vector<double> w;
      std::vector<dmrgx::cplx> Vh;
      host_hermitian_eig(n, H, !cplx, w, Vh);
      res[c].evals.assign(w.begin(), w.begin() + std::min(nev, n));
      res[c].n_matvec = nd;
      res[c].converged = true;
      for (int q = 0; q < std::min(nvs, n); ++q) {
        double* xq = x.data() + (size_t)q * nr;
        for (int i = 0; i < n; ++i) {
          if (cplx) {
            xq[2 * (off[c] + i)] = Vh[(size_t)i * n + q].real();
            xq[2 * (off[c] + i) + 1] = Vh[(size_t)i * n + q].imag();
          } else {
            xq[off[c] + i] = Vh[(size_t)i * n + q].real();
          }
        }
      }
    }
    // the small chains' eigenvectors; the Lanczos chains' segments follow below
    DMRGX_HIP(hipMemcpyAsync(evec, x.data(), x.size() * 8, hipMemcpyHostToDevice, s));
    DMRGX_HIP(hipStreamSynchronize(s));
    if (big.empty()) return res;
  }

  const int nb = (int)big.size();
  int64_t nmin = len[big[0]];
  for (int c : big) nmin = std::min(nmin, len[c]);
  nev = (int)std::min<int64_t>(nev, nmin);
  const int m = (int)std::min<int64_t>(std::max(ncv, nev + 2), nmin);
  const int keep = std::min(m - 2, nev + std::max(2, (m - nev) / 2));
  Segs sg(off, len, big, s);
  SegOps ops(sg, cplx, ld, m + 1, s);
  const int hs = ops.hs;
  DevBuf Vb((size_t)(m + 1) * nr * 8), Tmp((size_t)m * nr * 8);
  DevBuf hbuf((size_t)2 * nb * hs * 8), scl((size_t)nb * 8);
  const int nv = std::max(1, std::min(n_vec, m));
  DevBuf Yb((size_t)nb * m * std::max(keep, nv) * 8);
  double* V = (double*)Vb.ptr;
  double* w = (double*)Tmp.ptr;
  double* h = (double*)hbuf.ptr;
  double* h2 = h + (size_t)nb * hs;
  // the small chains' segments stay zero in every Krylov vector
  if (!small.empty()) {
    fill_zero(V, (size_t)(m + 1) * nr * 8, s);
    fill_zero(Tmp.ptr, (size_t)m * nr * 8, s);
  }
  if (v0) {
    copy_d2d(V, v0, nr * 8, s);
    for (int c : small) fill_zero(V + cw * off[c], len[c] * cw * 8, s);
  } else {
    for (int c : big) dev_random(V + cw * off[c], len[c], cplx, seed[c], s);
  }
  ops.normalise(V, h, (double*)scl.ptr);

  std::vector<std::vector<double>> T(nb, std::vector<double>((size_t)m * m, 0.0)), theta(nb), Y(nb);
  std::vector<double> beta_last(nb, 0.0), alpha(nb), b2(nb), nw2(nb), sc(nb);
  std::vector<int> n_reorth(nb, 0);
  std::vector<char> conv(nb, 0), retired(nb, 0);
  int j0 = 0, n_matvec = 0, n_restart = 0;
  std::unique_ptr<GraphSet> gstep = std::make_unique<GraphSet>(m, s), gfin = std::make_unique<GraphSet>(m, s);
  auto run_graph = [&](GraphSet::Slot& g, const std::function<void()>& body) {
    if (lanczos_graphs()) g.run(body);
    else body();
  };
  for (int restart = 0; restart <= max_restart; ++restart) {
    n_restart = restart;
    for (int j = j0; j < m; ++j) {
      // the matvec and two classical Gram-Schmidt passes (as in eigsh); the
      // second pass's |w'|^2 comes with its coefficients, both sets in one
      // synchronisation.  The same launches recur at index j in every restart
      // cycle, so they run as one hipGraph per j, captured at its first use.
      auto body = [&] {
        op(V + (int64_t)j * ld, w);
        ops.vdots(V, j + 1, w, h);
        ops.sub(V, j + 1, h, w, w, nullptr);
        ops.vdots(V, j + 1, w, h2);
        ops.fetch_async(h, 2);  // h and h2 are adjacent
      };
      run_graph((*gstep)[j], body);
      ++n_matvec;
      DMRGX_HIP(hipStreamSynchronize(s));
      const double* hh = ops.pin;
      for (int a = 0; a < nb; ++a) {
        if (retired[a]) continue;
        const double* q = hh + (size_t)a * hs;
        const double* q2 = hh + (size_t)(nb + a) * hs;
        double hsq = 0, w2 = q2[2 * (j + 1)];
        for (int i = 0; i <= j; ++i) {
          hsq += q2[2 * i] * q2[2 * i] + q2[2 * i + 1] * q2[2 * i + 1];
          w2 += q[2 * i] * q[2 * i] + q[2 * i + 1] * q[2 * i + 1];
        }
        alpha[a] = q[2 * j] + q2[2 * j];
        nw2[a] = w2;
        b2[a] = q2[2 * (j + 1)] - hsq;
        n_reorth[a]++;
      }
      const double* hfin = h2;
      std::vector<int> lost;
      for (int a = 0; a < nb; ++a) {
        if (retired[a]) {  // a finished chain's later vectors are zero
          sc[a] = 0;
          continue;
        }
        double beta = std::sqrt(std::max(b2[a], 0.0));
        if (!(beta > 1e-13 * std::max(std::fabs(alpha[a]), std::sqrt(std::max(nw2[a], 0.0))) && beta > 1e-300)) {
          beta = 0;
          lost.push_back(a);
        }
        T[a][(size_t)j * m + j] = alpha[a];
        if (j + 1 < m) T[a][(size_t)j * m + j + 1] = T[a][(size_t)(j + 1) * m + j] = beta;
        beta_last[a] = beta;
        sc[a] = beta > 0 ? 1.0 / beta : 0.0;
      }
      double* vn = V + (int64_t)(j + 1) * ld;
      ops.stage(sc);
      run_graph((*gfin)[j], [&] {
        ops.put_staged((double*)scl.ptr);
        ops.sub(V, j + 1, hfin, w, vn, (const double*)scl.ptr);
      });
      for (int a : lost) {
        // invariant subspace of this chain: continue with a fresh random
        // direction orthogonal to its basis (as in eigsh)
        const int c = big[a];
        double* seg = vn + cw * off[c];
        dev_random(seg, len[c], cplx, seed[c] + 7919ull * (j + 1) + restart, s);
        dev_orthogonalise(V + cw * off[c], ld, j + 1, seg, len[c], cplx, s);
        const double z = dev_norm2(seg, len[c], cplx, s);
        dev_scale(seg, len[c], cplx, 1.0 / z, s);
      }
      DMRGX_HIP(hipStreamSynchronize(s));  // sc is rewritten by the next step
    }
    // Rayleigh-Ritz per chain, the chains spread over host threads
    parallel_for(nb, [&](int a) {
      if (retired[a]) return;
      host_symeig(m, T[a], theta[a], Y[a]);
      bool ok = true;
      double mr = 0;
      for (int i = 0; i < nev; ++i) {
        const double r = std::fabs(beta_last[a] * Y[a][(size_t)(m - 1) * m + i]);
        mr = std::max(mr, r);
        if (r > tol * std::max(std::fabs(theta[a][i]), eps23)) ok = false;
      }
      res[big[a]].max_resid = mr;
      conv[a] = ok;
    });
    bool all = true;
    for (int a = 0; a < nb; ++a) all = all && conv[a];
    if (all || restart == max_restart) break;
    // chains that converged in this cycle finish here: their Ritz vectors
    // now, and the caller may drop them from the operator
    if (retire) {
      std::vector<int> done;
      for (int a = 0; a < nb; ++a)
        if (conv[a] && !retired[a]) done.push_back(a);
      if (!done.empty()) {
        std::vector<double> y0((size_t)nb * m * nv, 0.0);
        for (int a : done)
          for (int r = 0; r < m; ++r)
            for (int q = 0; q < nv; ++q) y0[((size_t)a * m + r) * nv + q] = Y[a][(size_t)r * m + q];
        DMRGX_HIP(hipMemcpyAsync(Yb.ptr, y0.data(), y0.size() * 8, hipMemcpyHostToDevice, s));
        ops.combine(V, m, (const double*)Yb.ptr, nv, (double*)Tmp.ptr);
        ops.normalise((double*)Tmp.ptr, h, (double*)scl.ptr, true);
        for (int a : done) {
          const int c = big[a];
          for (int q = 0; q < nv; ++q)
            copy_d2d(evec + q * nr + cw * off[c], (double*)Tmp.ptr + q * nr + cw * off[c], len[c] * cw * 8, s);
          EigshResult& r = res[c];
          r.evals.assign(theta[a].begin(), theta[a].begin() + nev);
          r.n_matvec = n_matvec;
          r.n_restart = n_restart;
          r.n_reorth = n_reorth[a];
          r.converged = true;
          retired[a] = 1;
        }
        DMRGX_HIP(hipStreamSynchronize(s));  // y0 ends here
        std::vector<char> active(nc, 0);
        for (int a = 0; a < nb; ++a)
          if (!retired[a]) active[big[a]] = 1;
        retire(active);
        // captured steps hold the previous operator's launches
        gstep = std::make_unique<GraphSet>(m, s);
        gfin = std::make_unique<GraphSet>(m, s);
      }
    }
    // thick restart of every chain: its `keep` lowest Ritz vectors and the residual
    std::vector<double> Yk((size_t)nb * m * keep);
    for (int a = 0; a < nb; ++a)
      for (int r = 0; r < m; ++r)
        for (int k = 0; k < keep; ++k) Yk[((size_t)a * m + r) * keep + k] = Y[a][(size_t)r * m + k];
    DMRGX_HIP(hipMemcpyAsync(Yb.ptr, Yk.data(), Yk.size() * 8, hipMemcpyHostToDevice, s));
    ops.combine(V, m, (const double*)Yb.ptr, keep, (double*)Tmp.ptr);
    DMRGX_HIP(hipMemcpyAsync(V, Tmp.ptr, (size_t)keep * nr * 8, hipMemcpyDeviceToDevice, s));
    DMRGX_HIP(hipMemcpyAsync(V + (int64_t)keep * ld, V + (int64_t)m * ld, nr * 8, hipMemcpyDeviceToDevice, s));
    for (int a = 0; a < nb; ++a) {
      std::fill(T[a].begin(), T[a].end(), 0.0);
      for (int i = 0; i < keep; ++i) {
        T[a][(size_t)i * m + i] = theta[a][i];
        const double sgl = beta_last[a] * Y[a][(size_t)(m - 1) * m + i];
        T[a][(size_t)i * m + keep] = T[a][(size_t)keep * m + i] = sgl;
      }
    }
    j0 = keep;
    DMRGX_HIP(hipStreamSynchronize(s));  // Yk ends here
  }
  // the nv lowest Ritz vectors of every Lanczos chain (the lowest normalised)
  // into its segments of evec's columns
  std::vector<double> y0((size_t)nb * m * nv);
  for (int a = 0; a < nb; ++a)
    for (int r = 0; r < m; ++r)
      for (int q = 0; q < nv; ++q) y0[((size_t)a * m + r) * nv + q] = Y[a][(size_t)r * m + q];
  DMRGX_HIP(hipMemcpyAsync(Yb.ptr, y0.data(), y0.size() * 8, hipMemcpyHostToDevice, s));
  double* out = (double*)Tmp.ptr;
  ops.combine(V, m, (const double*)Yb.ptr, nv, out);
  int n_retired = 0;
  for (char x : retired) n_retired += x;
  ops.normalise(out, h, (double*)scl.ptr, n_retired > 0);
  if (small.empty() && n_retired == 0) {
    copy_d2d(evec, out, (size_t)nv * nr * 8, s);
  } else {
    for (int q = 0; q < nv; ++q)
      for (int a = 0; a < nb; ++a)
        if (!retired[a]) {
          const int c = big[a];
          copy_d2d(evec + q * nr + cw * off[c], out + q * nr + cw * off[c], len[c] * cw * 8, s);
        }
  }
  DMRGX_HIP(hipStreamSynchronize(s));
  for (int a = 0; a < nb; ++a) {
    if (retired[a]) continue;
    EigshResult& r = res[big[a]];
    r.evals.assign(theta[a].begin(), theta[a].begin() + nev);
    r.n_matvec = n_matvec;
    r.n_restart = n_restart;
    r.n_reorth = n_reorth[a];
    r.converged = conv[a];
  }
  DMRGX_HIP(hipGetLastError());
  return res;
}

}  // namespace dmrgx

namespace dmrgx {

std::vector<double> jd_polish_batched(const MatVec& op_lo, const MatVec& op64, const std::vector<int64_t>& len,
                                      bool cplx, int p, double* X, double tol, int max_outer, int inner,
                                      int* n64, int* nlo, std::vector<double>* res_out,
                                      std::vector<char>* conv_out) {
  ensure_init();
  hipStream_t s = rt().stream;
  const int nc = (int)len.size();
  DMRGX_CHECK(nc > 0, "jd_polish: no chains");
  DMRGX_CHECK(p >= 1 && max_outer >= 0 && inner >= 1, "jd_polish: p >= 1, max_outer >= 0, inner >= 1");
  const std::vector<int64_t> off = offsets(len);
  const int64_t N = off.back() + len.back();
  const int64_t nr = N * (cplx ? 2 : 1);
  const int qmax = p + max_outer;
  std::vector<int> all(nc);
  for (int c = 0; c < nc; ++c) all[c] = c;
  Segs sg(off, len, all, s);
  SegOps ops(sg, cplx, nr, qmax + 1, s);
  const int hs = ops.hs;
  const size_t vb = (size_t)nr * 8;
  DevBuf Vb(qmax * vb), HVb(qmax * vb), wk(7 * vb), hb((size_t)2 * nc * hs * 8), scl((size_t)nc * 8),
      Zb((size_t)nc * qmax * 16);
  // an FP32 copy of the basis for the CG projections (the inner solve is a
  // few digits deep; the correction is orthogonalised against V in FP64)
  DevBuf Vfb(qmax * (size_t)nr * 4);
  float* Vf = (float*)Vfb.ptr;
  // the CG scalars (device), their pinned staging and the per-step live counts
  DevBuf cgb((size_t)CG_FIELDS * nc * 8), liveb((size_t)std::max(inner, 1) * sizeof(int));
  double* cgd = (double*)cgb.ptr;
  int* live_dev = (int*)liveb.ptr;
  struct CgHost {
    double* pin = nullptr;
    int* live = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    ~CgHost() {
      if (pin) (void)hipHostFree(pin);
      if (live) (void)hipHostFree(live);
      for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
    }
  } cgh;
  DMRGX_HIP(hipHostMalloc((void**)&cgh.pin, (size_t)CG_FIELDS * nc * 8, hipHostMallocDefault));
  DMRGX_HIP(hipHostMalloc((void**)&cgh.live, (size_t)std::max(inner, 1) * sizeof(int), hipHostMallocDefault));
  for (auto& e : cgh.ev) DMRGX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  double* cg_pin = cgh.pin;
  int* live_pin = cgh.live;
  hipEvent_t* cg_ev = cgh.ev;
  double* V = (double*)Vb.ptr;
  double* HV = (double*)HVb.ptr;
  double* u = (double*)wk.ptr;
  double* t1 = u + nr;
  double* r = t1 + nr;
  double* sv = r + nr;
  double* dv = sv + nr;
  double* wv = dv + nr;
  double* dn = wv + nr;
  double* h = (double*)hb.ptr;
  double* h2 = h + (size_t)nc * hs;
  const double eps23 = std::pow(std::numeric_limits<double>::epsilon(), 2.0 / 3.0);
  // per-chain host state: the projected matrix G = V^H H V and the valid columns
  std::vector<std::vector<dmrgx::cplx>> G(nc, std::vector<dmrgx::cplx>((size_t)qmax * qmax));
  std::vector<std::vector<char>> valid(nc, std::vector<char>(qmax, 0));
  std::vector<double> th0(nc, 0.0), sc(nc), coef((size_t)nc * hs, 0.0);
  auto put_coef = [&](const std::vector<double>& a, double* d) {  // one complex coefficient (a_c, 0) per chain
    std::fill(coef.begin(), coef.end(), 0.0);
    for (int c = 0; c < nc; ++c) coef[(size_t)c * hs] = a[c];
    DMRGX_HIP(hipMemcpyAsync(d, coef.data(), coef.size() * 8, hipMemcpyHostToDevice, s));
    DMRGX_HIP(hipStreamSynchronize(s));
  };
  // V_j = the part of src orthogonal to V[0..j) (two passes), normalised per
  // chain, or zero where it is dependent; then HV_j = H64 V_j and column j of G
  auto add_vector = [&](int j, const double* src) {
    ops.vdots(V, j, src, h);
    ops.sub(V, j, h, src, t1, nullptr);
    ops.vdots(V, j, t1, h2);
    const double* hh = ops.fetch(h, 2);
    for (int c = 0; c < nc; ++c) {
      const double* a = hh + (size_t)c * hs;
      const double* b = hh + (size_t)(nc + c) * hs;
      double hsq = 0;
      for (int i = 0; i < j; ++i) hsq += b[2 * i] * b[2 * i] + b[2 * i + 1] * b[2 * i + 1];
      const double w2 = a[2 * j], b2 = b[2 * j] - hsq;
      const bool ok = w2 > 0 && b2 > 1e-20 * w2 && j < len[c];
      valid[c][j] = ok;
      sc[c] = ok ? 1.0 / std::sqrt(b2) : 0.0;
    }
    ops.put(sc, (double*)scl.ptr);
    ops.sub(V, j, h2, t1, V + (size_t)j * nr, (const double*)scl.ptr);
    convert_f64_to_f32(V + (size_t)j * nr, Vf + (size_t)j * nr, nr, s);
    op64(V + (size_t)j * nr, HV + (size_t)j * nr);
    if (n64) (*n64)++;
    ops.vdots(V, j + 1, HV + (size_t)j * nr, h);
    const double* g = ops.fetch(h, 1);
    for (int c = 0; c < nc; ++c) {
      const double* col = g + (size_t)c * hs;
      for (int i = 0; i <= j; ++i) {
        const dmrgx::cplx x(col[2 * i], cplx ? col[2 * i + 1] : 0.0);
        G[c][(size_t)i * qmax + j] = x;
        G[c][(size_t)j * qmax + i] = std::conj(x);
      }
    }
  };
  int q = 0;
  for (int j = 0; j < p; ++j) add_vector(q++, X + (size_t)j * nr);
  std::vector<char> done(nc, 0);
  std::vector<double> rr(nc, 0.0);
  for (;;) {  // ends when every chain converged or the basis is full
    // Rayleigh-Ritz in span V: the lowest pair, u = V z, r = H u - theta u
    std::vector<double> Z((size_t)nc * q * 2, 0.0);
    parallel_for(nc, [&](int c) {
      std::vector<int> idx;
      for (int j = 0; j < q; ++j)
        if (valid[c][j]) idx.push_back(j);
      const int n = (int)idx.size();
      DMRGX_CHECK(n > 0, "jd_polish: empty basis");
      std::vector<dmrgx::cplx> A((size_t)n * n), W;
      for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) A[(size_t)a * n + b] = G[c][(size_t)idx[a] * qmax + idx[b]];
      std::vector<double> w;
      host_hermitian_eig(n, A, !cplx, w, W);
      th0[c] = w[0];
      for (int a = 0; a < n; ++a) {
        Z[((size_t)c * q + idx[a]) * 2] = W[(size_t)a * n].real();
        Z[((size_t)c * q + idx[a]) * 2 + 1] = W[(size_t)a * n].imag();
      }
    });
    DMRGX_HIP(hipMemcpyAsync(Zb.ptr, Z.data(), Z.size() * 8, hipMemcpyHostToDevice, s));
    ops.combine_z(V, q, (const double*)Zb.ptr, 1, u);
    ops.combine_z(HV, q, (const double*)Zb.ptr, 1, r);
    put_coef(th0, h);
    ops.sub(u, 1, h, r, r, nullptr);  // r = H u - theta u
    ops.vdots(nullptr, 0, r, h2);
    const double* rr0 = ops.fetch(h2, 1);
    bool all_done = true;
    for (int c = 0; c < nc; ++c) {
      rr[c] = std::max(rr0[(size_t)c * hs], 0.0);
      done[c] = std::sqrt(rr[c]) <= tol * std::max(std::fabs(th0[c]), eps23);
      all_done = all_done && done[c];
    }
    if (all_done || q >= qmax) break;
    // correction: CG on P (H_lo - theta) P s = r, P the projector orthogonal
    // to V (its spectrum on range(P) is >= E_p - theta > 0 near convergence);
    // converged chains keep s = 0
    fill_zero(sv, vb, s);
    copy_d2d(dv, r, vb, s);
    // zero the residual of converged chains so their s stays zero
    {
      std::vector<double> keep(nc);
      for (int c = 0; c < nc; ++c) keep[c] = done[c] ? 0.0 : 1.0;
      ops.put(keep, (double*)scl.ptr);
      ops.sub(nullptr, 0, h, dv, dv, (const double*)scl.ptr);
      copy_d2d(t1, dv, vb, s);  // t1: the CG residual
    }
    // the CG scalars start on the device; the steps queue without waiting on
    // the host, which reads each step's count of live chains one step later
    for (int c = 0; c < nc; ++c) {
      const double rho = done[c] ? 0.0 : rr[c];
      cg_pin[CG_THETA * nc + c] = th0[c];
      cg_pin[CG_RHO * nc + c] = rho;
      cg_pin[CG_RR * nc + c] = rr[c];
      cg_pin[CG_LIVE * nc + c] = (!done[c] && rho > 0) ? 1.0 : 0.0;
      cg_pin[CG_ALPHA * nc + c] = cg_pin[CG_BETA * nc + c] = 0.0;
    }
    DMRGX_HIP(hipMemcpyAsync(cgd, cg_pin, (size_t)CG_FIELDS * nc * 8, hipMemcpyHostToDevice, s));
    const int ng = (q + 1 + COLS - 1) / COLS;
    double* part = (double*)ops.partial.ptr;
    const int* cb = (const int*)ops.sg.cb->ptr;
    const int* ce = (const int*)ops.sg.ce->ptr;
    const int nblk = ops.sg.nblk;
    for (int it = 0; it < inner; ++it) {
      op_lo(dv, wv);
      if (nlo) (*nlo)++;
      // w = P (H_lo - theta) d and <d, w>
      if (cplx) hipLaunchKernelGGL((k_vdots_shift_seg<true, float>), dim3(nblk, ng), dim3(256), 0, s, Vf, nr, q, wv,
                                   dv, cgd, nc, ops.sg.B(), part);
      else hipLaunchKernelGGL((k_vdots_shift_seg<false, float>), dim3(nblk, ng), dim3(256), 0, s, Vf, nr, q, wv, dv,
                              cgd, nc, ops.sg.B(), part);
      hipLaunchKernelGGL(k_reduce_seg, dim3(nc, ng), dim3(64 * 2 * COLS), 0, s, part, nblk, q + 1, cb, ce, h, hs);
      if (cplx) hipLaunchKernelGGL((k_cg_project<true, float>), dim3(nblk), dim3(256), 0, s, Vf, nr, q, h, hs, wv, dv,
                                   cgd, nc, ops.sg.B(), part);
      else hipLaunchKernelGGL((k_cg_project<false, float>), dim3(nblk), dim3(256), 0, s, Vf, nr, q, h, hs, wv, dv,
                              cgd, nc, ops.sg.B(), part);
      hipLaunchKernelGGL(k_reduce_seg, dim3(nc, 1), dim3(64 * 2 * COLS), 0, s, part, nblk, 1, cb, ce, h2, hs);
      hipLaunchKernelGGL(k_cg_alpha, dim3(1), dim3(256), 0, s, nc, h2, hs, cgd);
      // s += alpha d, res -= alpha w, |res|^2
      if (cplx) hipLaunchKernelGGL(k_cg_update<true>, dim3(nblk), dim3(256), 0, s, dv, sv, wv, t1, cgd, nc,
                                   ops.sg.B(), part);
      else hipLaunchKernelGGL(k_cg_update<false>, dim3(nblk), dim3(256), 0, s, dv, sv, wv, t1, cgd, nc, ops.sg.B(),
                              part);
      hipLaunchKernelGGL(k_reduce_seg, dim3(nc, 1), dim3(64 * 2 * COLS), 0, s, part, nblk, 1, cb, ce, h2, hs);
      hipLaunchKernelGGL(k_cg_beta, dim3(1), dim3(256), 0, s, nc, h2, hs, cgd, live_dev + it);
      DMRGX_HIP(hipMemcpyAsync(live_pin + it, live_dev + it, sizeof(int), hipMemcpyDeviceToHost, s));
      DMRGX_HIP(hipEventRecord(cg_ev[it & 1], s));
      if (it + 1 == inner) break;
      // an inner solve a few digits deep is enough for the outer step: stop
      // once no chain was live after the previous step (this step then ran
      // with every alpha zero and changed nothing)
      if (it > 0) {
        DMRGX_HIP(hipEventSynchronize(cg_ev[(it - 1) & 1]));
        if (live_pin[it - 1] == 0) break;
      }
      if (cplx) hipLaunchKernelGGL(k_cg_dir<true>, dim3(nblk), dim3(256), 0, s, t1, dv, dn, cgd, nc, ops.sg.B());
      else hipLaunchKernelGGL(k_cg_dir<false>, dim3(nblk), dim3(256), 0, s, t1, dv, dn, cgd, nc, ops.sg.B());
      std::swap(dv, dn);
    }
    add_vector(q++, sv);
  }
  // the polished lowest vector of every chain into X column 0
  copy_d2d(X, u, vb, s);
  ops.normalise(X, h, (double*)scl.ptr);
  DMRGX_HIP(hipStreamSynchronize(s));
  if (res_out) {
    res_out->resize(nc);
    for (int c = 0; c < nc; ++c) (*res_out)[c] = std::sqrt(rr[c]);
  }
  if (conv_out) *conv_out = done;
  return th0;
}

std::vector<double> residual_batched(const MatVec& op, const std::vector<int64_t>& len, bool cplx, const double* x,
                                     const std::vector<double>& theta) {
  ensure_init();
  hipStream_t s = rt().stream;
  const int nc = (int)len.size();
  DMRGX_CHECK(nc > 0 && (int)theta.size() == nc, "residual_batched: one Ritz value per chain");
  const std::vector<int64_t> off = offsets(len);
  const int64_t nr = (off.back() + len.back()) * (cplx ? 2 : 1);
  std::vector<int> all(nc);
  for (int c = 0; c < nc; ++c) all[c] = c;
  Segs sg(off, len, all, s);
  SegOps ops(sg, cplx, nr, 2, s);
  const int hs = ops.hs;
  DevBuf yb((size_t)nr * 8), hb((size_t)nc * hs * 8);
  double* y = (double*)yb.ptr;
  double* h = (double*)hb.ptr;
  op(x, y);
  std::vector<double> coef((size_t)nc * hs, 0.0);
  for (int c = 0; c < nc; ++c) coef[(size_t)c * hs] = theta[c];
  DMRGX_HIP(hipMemcpyAsync(h, coef.data(), coef.size() * 8, hipMemcpyHostToDevice, s));
  ops.sub(x, 1, h, y, y, nullptr);  // y = H x - theta x
  ops.vdots(nullptr, 0, y, h);
  const double* rr = ops.fetch(h, 1);
  std::vector<double> out(nc);
  for (int c = 0; c < nc; ++c) out[c] = std::sqrt(std::max(rr[(size_t)c * hs], 0.0));
  return out;
}

}  // namespace dmrgx
