// Leg contraction (environment x MPO), leg permutation and strided box copies.
// All are HBM-bound; each launch covers every block of a qtensor operation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "dmrgx.h"
#include "common.hpp"
#include "tensor_ops.hpp"

namespace dmrgx {
namespace {

inline unsigned blocks_for(int64_t n) { return (unsigned)std::max<int64_t>(1, (n + 255) / 256); }

// find the task owning flat block index b (start[] ascending, start[0] = 0)
__device__ __forceinline__ int find_task(const int64_t* start, int n, int64_t b) {
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <bool CPLX>
__global__ __launch_bounds__(256) void k_leg(const double* __restrict__ A, const double* __restrict__ B,
                                             double* __restrict__ C, const LegOut* __restrict__ outs, int n_out,
                                             const LegPair* __restrict__ pairs, const int64_t* __restrict__ start) {
  const int t = find_task(start, n_out, blockIdx.x);
  const LegOut o = outs[t];
  const int64_t total = (int64_t)o.al * o.bl * o.br * o.ar;
  const int64_t e = (blockIdx.x - start[t]) * 256 + threadIdx.x;
  if (e >= total) return;
  const int i2 = (int)(e % o.ar);
  int64_t r = e / o.ar;
  const int j2 = (int)(r % o.br);
  r /= o.br;
  const int j1 = (int)(r % o.bl);
  const int i1 = (int)(r / o.bl);
  double sr = 0, si = 0;
  for (int p = o.pair_begin; p < o.pair_end; ++p) {
    const LegPair q = pairs[p];
    double ar_ = 0, ai_ = 0;
    for (int k = 0; k < q.mid; ++k) {
      int64_t ia = q.a_off + ((int64_t)i1 * q.mid + k) * o.ar + i2;
      int64_t ib = q.b_off + ((int64_t)j1 * q.mid + k) * o.br + j2;
      if (CPLX) {
        double xr = A[2 * ia], xi = A[2 * ia + 1], yr = B[2 * ib], yi = B[2 * ib + 1];
        ar_ += xr * yr - xi * yi;
        ai_ += xr * yi + xi * yr;
      } else {
        ar_ += A[ia] * B[ib];
      }
    }
    sr += q.sign * ar_;
    si += q.sign * ai_;
  }
  const int64_t ic = o.c_off + e;
  if (CPLX) {
    if (o.accumulate) { C[2 * ic] += sr; C[2 * ic + 1] += si; }
    else { C[2 * ic] = sr; C[2 * ic + 1] = si; }
  } else {
    if (o.accumulate) C[ic] += sr;
    else C[ic] = sr;
  }
}

template <bool CPLX>
__global__ __launch_bounds__(256) void k_permute(const double* __restrict__ in, double* __restrict__ out,
                                                 const PermTask* __restrict__ tasks, int n,
                                                 const int64_t* __restrict__ start) {
  const int t = find_task(start, n, blockIdx.x);
  const PermTask p = tasks[t];
  int64_t total = 1;
  for (int i = 0; i < p.nleg; ++i) total *= p.dims[i];
  const int64_t e = (blockIdx.x - start[t]) * 256 + threadIdx.x;
  if (e >= total) return;
  int64_t istride[8];
  int64_t s = 1;
  for (int i = p.nleg - 1; i >= 0; --i) { istride[i] = s; s *= p.dims[i]; }
  // decompose the output index (output leg q has dim dims[perm[q]])
  int64_t r = e, src = 0;
  for (int q = p.nleg - 1; q >= 0; --q) {
    int d = p.dims[p.perm[q]];
    int64_t c = r % d;
    r /= d;
    src += c * istride[p.perm[q]];
  }
  if (CPLX) {
    out[2 * (p.out_off + e)] = in[2 * (p.in_off + src)];
    out[2 * (p.out_off + e) + 1] = in[2 * (p.in_off + src) + 1];
  } else {
    out[p.out_off + e] = in[p.in_off + src];
  }
}

template <bool CPLX>
__global__ __launch_bounds__(256) void k_boxes(const double* __restrict__ src, double* __restrict__ dst,
                                               const BoxTask* __restrict__ tasks, int n,
                                               const int64_t* __restrict__ start) {
  const int t = find_task(start, n, blockIdx.x);
  const BoxTask b = tasks[t];
  const int64_t total = (int64_t)b.n0 * b.n1 * b.n2;
  const int64_t e = (blockIdx.x - start[t]) * 256 + threadIdx.x;
  if (e >= total) return;
  const int k = (int)(e % b.n2);
  const int64_t r = e / b.n2;
  const int j = (int)(r % b.n1);
  const int i = (int)(r / b.n1);
  const int64_t is = b.src_off + i * b.src_s0 + j * b.src_s1 + k;
  const int64_t id = b.dst_off + i * b.dst_s0 + j * b.dst_s1 + k;
  if (CPLX) {
    double xr = src[2 * is], xi = src[2 * is + 1];
    if (b.mode == -1) { xr = -xr; xi = -xi; }
    else if (b.mode == 2) xi = -xi;
    dst[2 * id] = xr;
    dst[2 * id + 1] = xi;
  } else {
    double x = src[is];
    dst[id] = b.mode == -1 ? -x : x;
  }
}

template <bool CPLX>
__global__ __launch_bounds__(256) void k_box_maxabs(const double* __restrict__ src, const BoxTask* __restrict__ tasks,
                                                    double* __restrict__ out) {
  const BoxTask b = tasks[blockIdx.x];
  const int64_t total = (int64_t)b.n0 * b.n1 * b.n2;
  double m = 0;
  for (int64_t e = threadIdx.x; e < total; e += 256) {
    const int k = (int)(e % b.n2);
    const int64_t r = e / b.n2;
    const int j = (int)(r % b.n1);
    const int i = (int)(r / b.n1);
    const int64_t is = b.src_off + i * b.src_s0 + j * b.src_s1 + k;
    double v = CPLX ? hypot(src[2 * is], src[2 * is + 1]) : fabs(src[is]);
    m = fmax(m, v);
  }
  __shared__ double red[256];
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = red[0];
}

__global__ void k_conj(double* __restrict__ p, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[2 * i + 1] = -p[2 * i + 1];
}
__global__ void k_scale_r(double* __restrict__ p, int64_t n, double a) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) p[i] *= a;
}

// upload a descriptor array and the per-task block starts; returns #blocks
template <typename T>
int64_t upload_tasks(const std::vector<T>& t, const std::vector<int64_t>& sizes, DevBuf& dt, DevBuf& ds,
                     hipStream_t s) {
  std::vector<int64_t> start(t.size());
  int64_t acc = 0;
  for (size_t i = 0; i < t.size(); ++i) {
    start[i] = acc;
    acc += std::max<int64_t>(1, (sizes[i] + 255) / 256);
  }
  upload(dt.ptr, t.data(), t.size() * sizeof(T), s);
  upload(ds.ptr, start.data(), start.size() * 8, s);
  return acc;
}

}  // namespace

void leg_contract(int dtype, const void* A, const void* B, void* C, const std::vector<LegOut>& outs,
                  const std::vector<LegPair>& pairs, hipStream_t s) {
  if (outs.empty()) return;
  std::vector<int64_t> sizes;
  for (auto& o : outs) sizes.push_back((int64_t)o.al * o.bl * o.br * o.ar);
  DevBuf dt(outs.size() * sizeof(LegOut)), dp(std::max<size_t>(pairs.size(), 1) * sizeof(LegPair)),
      ds(outs.size() * 8);
  int64_t nb = upload_tasks(outs, sizes, dt, ds, s);
  if (!pairs.empty())
    upload(dp.ptr, pairs.data(), pairs.size() * sizeof(LegPair), s);
  DMRGX_CHECK(nb < (int64_t)1 << 31, "leg_contract: grid too large");
  if (dtype == DT_C64)
    hipLaunchKernelGGL(k_leg<true>, dim3((unsigned)nb), dim3(256), 0, s, (const double*)A, (const double*)B,
                       (double*)C, (const LegOut*)dt.ptr, (int)outs.size(), (const LegPair*)dp.ptr,
                       (const int64_t*)ds.ptr);
  else
    hipLaunchKernelGGL(k_leg<false>, dim3((unsigned)nb), dim3(256), 0, s, (const double*)A, (const double*)B,
                       (double*)C, (const LegOut*)dt.ptr, (int)outs.size(), (const LegPair*)dp.ptr,
                       (const int64_t*)ds.ptr);
  DMRGX_HIP(hipGetLastError());
}

void permute(int dtype, const void* in, void* out, const std::vector<PermTask>& t, hipStream_t s) {
  if (t.empty()) return;
  std::vector<int64_t> sizes;
  for (auto& p : t) {
    DMRGX_CHECK(p.nleg >= 1 && p.nleg <= 8, "permute: 1..8 legs");
    int64_t n = 1;
    for (int i = 0; i < p.nleg; ++i) n *= p.dims[i];
    sizes.push_back(n);
  }
  DevBuf dt(t.size() * sizeof(PermTask)), ds(t.size() * 8);
  int64_t nb = upload_tasks(t, sizes, dt, ds, s);
  if (dtype == DT_C64)
    hipLaunchKernelGGL(k_permute<true>, dim3((unsigned)nb), dim3(256), 0, s, (const double*)in, (double*)out,
                       (const PermTask*)dt.ptr, (int)t.size(), (const int64_t*)ds.ptr);
  else
    hipLaunchKernelGGL(k_permute<false>, dim3((unsigned)nb), dim3(256), 0, s, (const double*)in, (double*)out,
                       (const PermTask*)dt.ptr, (int)t.size(), (const int64_t*)ds.ptr);
  DMRGX_HIP(hipGetLastError());
}

void copy_boxes(int dtype, const void* src, void* dst, const std::vector<BoxTask>& t, hipStream_t s) {
  if (t.empty()) return;
  std::vector<int64_t> sizes;
  for (auto& b : t) sizes.push_back((int64_t)b.n0 * b.n1 * b.n2);
  DevBuf dt(t.size() * sizeof(BoxTask)), ds(t.size() * 8);
  int64_t nb = upload_tasks(t, sizes, dt, ds, s);
  if (dtype == DT_C64)
    hipLaunchKernelGGL(k_boxes<true>, dim3((unsigned)nb), dim3(256), 0, s, (const double*)src, (double*)dst,
                       (const BoxTask*)dt.ptr, (int)t.size(), (const int64_t*)ds.ptr);
  else
    hipLaunchKernelGGL(k_boxes<false>, dim3((unsigned)nb), dim3(256), 0, s, (const double*)src, (double*)dst,
                       (const BoxTask*)dt.ptr, (int)t.size(), (const int64_t*)ds.ptr);
  DMRGX_HIP(hipGetLastError());
}

std::vector<double> box_maxabs(int dtype, const void* src, const std::vector<BoxTask>& t, hipStream_t s) {
  std::vector<double> out(t.size(), 0.0);
  if (t.empty()) return out;
  DevBuf dt(t.size() * sizeof(BoxTask)), dout(t.size() * 8);
  upload(dt.ptr, t.data(), t.size() * sizeof(BoxTask), s);
  if (dtype == DT_C64)
    hipLaunchKernelGGL(k_box_maxabs<true>, dim3((unsigned)t.size()), dim3(256), 0, s, (const double*)src,
                       (const BoxTask*)dt.ptr, (double*)dout.ptr);
  else
    hipLaunchKernelGGL(k_box_maxabs<false>, dim3((unsigned)t.size()), dim3(256), 0, s, (const double*)src,
                       (const BoxTask*)dt.ptr, (double*)dout.ptr);
  DMRGX_HIP(hipGetLastError());
  DMRGX_HIP(hipMemcpyAsync(out.data(), dout.ptr, t.size() * 8, hipMemcpyDeviceToHost, s));
  DMRGX_HIP(hipStreamSynchronize(s));
  return out;
}

// out[seg] = alpha A[seg] + beta B[seg] over block segments; a_off / b_off < 0
// means that operand has no such block (it counts as zero)
template <bool CPLX>
__global__ void k_axpby_seg(const double* __restrict__ A, const double* __restrict__ B, double* __restrict__ O,
                            const AxpbySeg* __restrict__ seg, int nseg, int64_t total, double ar, double ai,
                            double br, double bi) {
  int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; e < total; e += st) {
    int lo = 0, hi = nseg - 1;
    while (lo < hi) {
      int mid = (lo + hi + 1) >> 1;
      if (seg[mid].out_off <= e) lo = mid;
      else hi = mid - 1;
    }
    const AxpbySeg g = seg[lo];
    const int64_t k = e - g.out_off;
    if (CPLX) {
      double xr = 0, xi = 0, yr = 0, yi = 0;
      if (g.a_off >= 0) { xr = A[2 * (g.a_off + k)]; xi = A[2 * (g.a_off + k) + 1]; }
      if (g.b_off >= 0) { yr = B[2 * (g.b_off + k)]; yi = B[2 * (g.b_off + k) + 1]; }
      O[2 * e] = ar * xr - ai * xi + br * yr - bi * yi;
      O[2 * e + 1] = ar * xi + ai * xr + br * yi + bi * yr;
    } else {
      const double x = g.a_off >= 0 ? A[g.a_off + k] : 0.0, y = g.b_off >= 0 ? B[g.b_off + k] : 0.0;
      O[e] = ar * x + br * y;
    }
  }
}

void blocks_axpby(bool cplx, const void* A, const void* B, void* out, const std::vector<AxpbySeg>& segs,
                  int64_t total, double ar, double ai, double br, double bi, hipStream_t s) {
  if (total <= 0 || segs.empty()) return;
  DevBuf d(segs.size() * sizeof(AxpbySeg));
  upload(d.ptr, segs.data(), segs.size() * sizeof(AxpbySeg), s);
  const unsigned g = (unsigned)std::min<int64_t>(4096, blocks_for(total));
  if (cplx) hipLaunchKernelGGL(k_axpby_seg<true>, dim3(g), dim3(256), 0, s, (const double*)A, (const double*)B,
                               (double*)out, (const AxpbySeg*)d.ptr, (int)segs.size(), total, ar, ai, br, bi);
  else hipLaunchKernelGGL(k_axpby_seg<false>, dim3(g), dim3(256), 0, s, (const double*)A, (const double*)B,
                          (double*)out, (const AxpbySeg*)d.ptr, (int)segs.size(), total, ar, ai, br, bi);
  DMRGX_HIP(hipGetLastError());
}

__global__ void k_conj_f32(float* __restrict__ p, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int64_t st = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += st) p[2 * i + 1] = -p[2 * i + 1];
}

void conj_inplace_f32(void* p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_conj_f32, dim3((unsigned)std::min<int64_t>(4096, blocks_for(n))), dim3(256), 0, s, (float*)p, n);
  DMRGX_HIP(hipGetLastError());
}

void conj_inplace(void* p, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_conj, dim3((unsigned)std::min<int64_t>(4096, blocks_for(n))), dim3(256), 0, s, (double*)p, n);
  DMRGX_HIP(hipGetLastError());
}

void scale_real(double* p, int64_t n, double a, hipStream_t s) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_scale_r, dim3((unsigned)std::min<int64_t>(4096, blocks_for(n))), dim3(256), 0, s, p, n, a);
  DMRGX_HIP(hipGetLastError());
}

}  // namespace dmrgx

using namespace dmrgx;

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

extern "C" {

int dmrgx_leg_contract(int dtype, const void* A, const void* B, void* C, const int64_t* tasks, int n_task) {
  CAPI_BEGIN
  ensure_init();
  std::vector<LegOut> outs;
  std::vector<LegPair> pairs;
  for (int i = 0; i < n_task; ++i) {
    const int64_t* q = tasks + 10 * i;
    if (outs.empty() || outs.back().c_off != q[2]) {
      LegOut o{};
      o.c_off = q[2];
      o.al = (int)q[3];
      o.ar = (int)q[5];
      o.bl = (int)q[6];
      o.br = (int)q[7];
      o.pair_begin = (int)pairs.size();
      o.accumulate = (int)q[9];
      outs.push_back(o);
    }
    DMRGX_CHECK(outs.back().al == q[3] && outs.back().ar == q[5] && outs.back().bl == q[6] && outs.back().br == q[7],
                "leg_contract: pairs into one output disagree on shape");
    pairs.push_back(LegPair{q[0], q[1], (int)q[4], (int)q[8]});
    outs.back().pair_end = (int)pairs.size();
  }
  leg_contract(dtype, A, B, C, outs, pairs, rt().stream);
  DMRGX_HIP(hipStreamSynchronize(rt().stream));  // complete on return
  CAPI_END
}

int dmrgx_permute(int dtype, const void* in, void* out, const int64_t* tasks, int n_task) {
  CAPI_BEGIN
  ensure_init();
  std::vector<PermTask> t(n_task);
  for (int i = 0; i < n_task; ++i) {
    const int64_t* q = tasks + 19 * i;
    t[i].in_off = q[0];
    t[i].out_off = q[1];
    t[i].nleg = (int)q[2];
    for (int j = 0; j < 8; ++j) {
      t[i].dims[j] = (int)q[3 + j];
      t[i].perm[j] = (int)q[11 + j];
    }
  }
  permute(dtype, in, out, t, rt().stream);
  DMRGX_HIP(hipStreamSynchronize(rt().stream));  // complete on return
  CAPI_END
}

int dmrgx_copy_boxes(int dtype, const void* src, void* dst, const int64_t* tasks, int n_task) {
  CAPI_BEGIN
  ensure_init();
  std::vector<BoxTask> t(n_task);
  for (int i = 0; i < n_task; ++i) {
    const int64_t* q = tasks + 10 * i;
    t[i] = BoxTask{q[0], q[1], q[2], q[3], q[4], q[5], (int)q[6], (int)q[7], (int)q[8], (int)q[9]};
  }
  copy_boxes(dtype, src, dst, t, rt().stream);
  DMRGX_HIP(hipStreamSynchronize(rt().stream));  // complete on return
  CAPI_END
}

}  // extern "C"
