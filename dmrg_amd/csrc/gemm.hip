// Grouped block GEMM on CDNA4 matrix cores (gfx950).
//
// One launch executes every task of a get_map task list (or of a block-pair
// contraction).  A work item is one 64x64 output tile together with a K-chunk
// of the concatenated segments of all tasks that write that output block, so
// the sum over contracted sectors happens in registers, uneven sector shapes
// share a launch, and long K is split across work items (partial tiles in a
// workspace, summed by one reduction pass) to fill all CUs.  Three arithmetic modes:
//   PREC_FP64   v_mfma_f64_16x16x4f64, FP64 storage
//   PREC_FP32   v_mfma_f32_16x16x4f32, FP32 storage
//   PREC_BF16X3 v_mfma_f32_16x16x32_bf16 with each FP32 operand split into
//               hi+lo bfloat16 parts; hi*hi + hi*lo + lo*hi, FP32 accumulate
// Complex operands (interleaved re/im) use four real products per tile.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <map>
#include <numeric>
#include <queue>
#include <type_traits>

#include "common.hpp"
#include "gemm.hpp"

namespace dmrgx {

namespace {

constexpr int BM = 64, BN = 64, NT = 256;

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

template <typename S>
struct Mfma4;
template <>
struct Mfma4<double> {
  typedef d4v acc;
  static __device__ __forceinline__ acc op(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <>
struct Mfma4<float> {
  typedef f4v acc;
  static __device__ __forceinline__ acc op(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

// Sum the K-chunk partial tiles of the split output blocks into C.
// Element e (row-major M x N, compact) of an output block at its place in C:
// the row goes through the block's row map when it has one.
__device__ __forceinline__ int64_t out_index(const RedTask& r, const int32_t* __restrict__ RM, int64_t e) {
  if (r.row_map < 0) return r.c_off + e;
  const int64_t row = e / r.ncol;
  return r.c_off + (int64_t)RM[r.row_map + row] * r.ncol + (e - row * r.ncol);
}

// The same for plane output, which may land in a padded layout (OutPad).
__device__ __forceinline__ int64_t out_index_planes(const RedTask& r, const int32_t* __restrict__ RM, int64_t e) {
  if (r.op.kp == 0) return out_index(r, RM, e);
  int64_t row = e / r.ncol;
  const int64_t col = e - row * r.ncol;
  if (r.row_map >= 0) row = RM[r.row_map + row];
  const int64_t rr = r.op.r0 + row;
  return r.op.base + (rr / r.op.q) * r.op.kp + (rr % r.op.q) * r.ncol + col;
}

template <typename S, bool CPLX>
__global__ __launch_bounds__(256) void k_reduce_ws(const S* __restrict__ W, S* __restrict__ C,
                                                   const RedTask* __restrict__ t, int n,
                                                   const int64_t* __restrict__ start,
                                                   const int32_t* __restrict__ RM) {
  int lo = 0, hi = n - 1;
  const int64_t b = blockIdx.x;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const RedTask r = t[lo];
  const int64_t e = (b - start[lo]) * 256 + threadIdx.x;
  if (e >= r.len) return;
  // the parts in chunk order (unrolled so that four loads are in flight);
  // complex elements read as one vector
  if (CPLX) {
    using V2 = typename std::conditional<sizeof(S) == 8, double2, float2>::type;
    const V2* W2 = reinterpret_cast<const V2*>(W);
    S ar = 0, ai = 0;
#pragma unroll 4
    for (int k = 0; k < r.nch; ++k) {
      const V2 v = W2[r.w_off + k * r.len + e];
      ar += v.x;
      ai += v.y;
    }
    S* o = C + 2 * out_index(r, RM, e);
    if (r.beta) { o[0] += ar; o[1] += ai; }
    else { o[0] = ar; o[1] = ai; }
  } else {
    S acc = 0;
#pragma unroll 4
    for (int k = 0; k < r.nch; ++k) acc += W[r.w_off + k * r.len + e];
    const int64_t o = out_index(r, RM, e);
    if (r.beta) C[o] += acc;
    else C[o] = acc;
  }
}

// FP32 partial tiles of the split output blocks summed into FP64 or bf16 planes.
template <bool CPLX, int OUT>
__global__ __launch_bounds__(256) void k_reduce_ws_out(const float* __restrict__ W, void* __restrict__ Cv,
                                                       int64_t cps, const RedTask* __restrict__ t, int n,
                                                       const int64_t* __restrict__ start,
                                                       const int32_t* __restrict__ RM);

// Walks the (segment, k-tile) sequence of one work item.
struct Cursor {
  int seg, seg_end, k;
  GemmSeg s;
  __device__ void start(const GemmSeg* segs, int b, int e) {
    seg = b;
    seg_end = e;
    if (seg < seg_end) {
      s = segs[seg];
      k = s.k0;
    }
  }
  __device__ bool valid() const { return seg < seg_end; }
  template <int BK>
  __device__ void next(const GemmSeg* segs) {
    k += BK;
    if (k >= s.k1) {
      ++seg;
      if (seg < seg_end) {
        s = segs[seg];
        k = s.k0;
      }
    }
  }
};

// ---------------------------------------------------------------------------
// FP64 / FP32 kernel (MFMA 16x16x4), LDS tiles row-major [row][k] padded.
// AKC: A is k-contiguous (transA 'N'); BKC: B is k-contiguous (transB 'T').
template <typename S, bool CPLX, bool AKC, bool BKC>
__global__ __launch_bounds__(NT) void k_gemm4(const GemmItem* __restrict__ items, int nitems,
                                              int per_xcd, const GemmSeg* __restrict__ segs,
                                              const S* __restrict__ A, const S* __restrict__ B,
                                              S* __restrict__ C, S* __restrict__ W,
                                              const int32_t* __restrict__ RM) {
  constexpr int BK = 16;
  constexpr int LDK = BK + 1;
  constexpr int NP = CPLX ? 2 : 1;
  constexpr int EPT = BM * BK / NT;  // 4
  const int bid = blockIdx.x;
  const int item_id = (bid & 7) * per_xcd + (bid >> 3);
  if (item_id >= nitems) return;
  const GemmItem it = items[item_id];

  __shared__ S sA[NP][BM * LDK];
  __shared__ S sB[NP][BN * LDK];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  using acc_t = typename Mfma4<S>::acc;
  acc_t acc[NP][2][2];
#pragma unroll
  for (int p = 0; p < NP; ++p)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[p][i][j] = acc_t{0, 0, 0, 0};

  S ra[NP][EPT], rb[NP][EPT];
  int lr[EPT], lk[EPT];  // LDS coordinates of A loads
  int mr[EPT], mk[EPT];  // LDS coordinates of B loads
  int ar[EPT];           // A row of each load (through the row map), -1 outside the tile
#pragma unroll
  for (int j = 0; j < EPT; ++j) {
    if (AKC) { lr[j] = t / BK + (NT / BK) * j; lk[j] = t % BK; }
    else     { lr[j] = t % BM; lk[j] = t / BM + (NT / BM) * j; }
    if (BKC) { mr[j] = t / BK + (NT / BK) * j; mk[j] = t % BK; }
    else     { mr[j] = t % BN; mk[j] = t / BN + (NT / BN) * j; }
    const int gr = it.m0 + lr[j];
    ar[j] = gr < it.M ? (it.row_map >= 0 ? RM[it.row_map + gr] : gr) : -1;
  }

  Cursor cur;
  cur.start(segs, it.seg_begin, it.seg_end);

  auto gload = [&](const Cursor& c) {
    const GemmSeg& s = c.s;
    const S al = (S)s.alpha;
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      int gr = ar[j], gk = c.k + lk[j];
      bool ok = gr >= 0 && gk < s.k1;
      int64_t e = AKC ? s.a_off + (int64_t)gr * s.K + gk : s.a_off + (int64_t)gk * it.M + gr;
      if (CPLX) {
        ra[0][j] = ok ? A[2 * e] * al : (S)0;
        ra[NP - 1][j] = ok ? A[2 * e + 1] * al : (S)0;
      } else {
        ra[0][j] = ok ? A[e] * al : (S)0;
      }
      int gn = it.n0 + mr[j], gk2 = c.k + mk[j];
      bool ok2 = gn < it.N && gk2 < s.k1;
      int64_t f = BKC ? s.b_off + (int64_t)gn * s.K + gk2 : s.b_off + (int64_t)gk2 * it.N + gn;
      if (CPLX) {
        rb[0][j] = ok2 ? B[2 * f] : (S)0;
        rb[NP - 1][j] = ok2 ? B[2 * f + 1] : (S)0;
      } else {
        rb[0][j] = ok2 ? B[f] : (S)0;
      }
    }
  };

  if (cur.valid()) gload(cur);
  while (cur.valid()) {
    __syncthreads();
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
      for (int j = 0; j < EPT; ++j) {
        sA[p][lr[j] * LDK + lk[j]] = ra[p][j];
        sB[p][mr[j] * LDK + mk[j]] = rb[p][j];
      }
    __syncthreads();
    cur.next<BK>(segs);
    if (cur.valid()) gload(cur);
#pragma unroll
    for (int ks = 0; ks < BK / 4; ++ks) {
      const int kx = ks * 4 + (lane >> 4);
      S a[NP][2], b[NP][2];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int p = 0; p < NP; ++p) {
          a[p][i] = sA[p][(wm + i * 16 + (lane & 15)) * LDK + kx];
          b[p][i] = sB[p][(wn + i * 16 + (lane & 15)) * LDK + kx];
        }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          if (CPLX) {
            acc[0][i][j] = Mfma4<S>::op(a[0][i], b[0][j], acc[0][i][j]);
            acc[0][i][j] = Mfma4<S>::op(-a[NP - 1][i], b[NP - 1][j], acc[0][i][j]);
            acc[NP - 1][i][j] = Mfma4<S>::op(a[0][i], b[NP - 1][j], acc[NP - 1][i][j]);
            acc[NP - 1][i][j] = Mfma4<S>::op(a[NP - 1][i], b[0][j], acc[NP - 1][i][j]);
          } else {
            acc[0][i][j] = Mfma4<S>::op(a[0][i], b[0][j], acc[0][i][j]);
          }
        }
    }
  }

  // epilogue: lane holds column lane%16 of each 16x16 tile; the row of
  // accumulator r is 4*(lane/16)+r for f32 and r*4+lane/16 for f64 MFMA
  const bool partial = it.flags & ITEM_PARTIAL;
  const bool beta = !partial && (it.flags & ITEM_BETA);
  auto* D = partial ? W : C;
  constexpr bool F64 = sizeof(S) == 8;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int row = it.m0 + wm + i * 16 + (F64 ? r * 4 + (lane >> 4) : (lane >> 4) * 4 + r);
        int col = it.n0 + wn + j * 16 + (lane & 15);
        if (row < it.M && col < it.N) {
          if (!partial && it.row_map >= 0) row = RM[it.row_map + row];
          int64_t e = it.c_off + (int64_t)row * it.N + col;
          if (CPLX) {
            S vr = acc[0][i][j][r], vi = acc[NP - 1][i][j][r];
            if (beta) { D[2 * e] += vr; D[2 * e + 1] += vi; }
            else { D[2 * e] = vr; D[2 * e + 1] = vi; }
          } else {
            S v = acc[0][i][j][r];
            if (beta) D[e] += v;
            else D[e] = v;
          }
        }
      }
}

// ---------------------------------------------------------------------------
// BF16x3 kernel: FP32 storage, operands split on the way into LDS.
__device__ __forceinline__ unsigned short f2bf_rne(float x) {
  unsigned u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

template <bool CPLX, bool AKC, bool BKC>
__global__ __launch_bounds__(NT) void k_gemm_bf16x3(const GemmItem* __restrict__ items, int nitems,
                                                    int per_xcd, const GemmSeg* __restrict__ segs,
                                                    const float* __restrict__ A,
                                                    const float* __restrict__ B,
                                                    float* __restrict__ C, float* __restrict__ W) {
  constexpr int BK = 32;
  constexpr int LDK = BK + 8;  // 80-byte rows keep 16-byte fragment reads aligned
  constexpr int NP = CPLX ? 2 : 1;
  constexpr int EPT = BM * BK / NT;  // 8
  const int bid = blockIdx.x;
  const int item_id = (bid & 7) * per_xcd + (bid >> 3);
  if (item_id >= nitems) return;
  const GemmItem it = items[item_id];

  // planes: [part][hi/lo][rows*LDK]
  __shared__ __attribute__((aligned(16))) unsigned short sA[NP][2][BM * LDK];
  __shared__ __attribute__((aligned(16))) unsigned short sB[NP][2][BN * LDK];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
  f4v acc[NP][2][2];
#pragma unroll
  for (int p = 0; p < NP; ++p)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[p][i][j] = f4v{0, 0, 0, 0};

  float ra[NP][EPT], rb[NP][EPT];
  int lr[EPT], lk[EPT], mr[EPT], mk[EPT];
#pragma unroll
  for (int j = 0; j < EPT; ++j) {
    if (AKC) { lr[j] = t / BK + (NT / BK) * j; lk[j] = t % BK; }
    else     { lr[j] = t % BM; lk[j] = t / BM + (NT / BM) * j; }
    if (BKC) { mr[j] = t / BK + (NT / BK) * j; mk[j] = t % BK; }
    else     { mr[j] = t % BN; mk[j] = t / BN + (NT / BN) * j; }
  }
  Cursor cur;
  cur.start(segs, it.seg_begin, it.seg_end);
  auto gload = [&](const Cursor& c) {
    const GemmSeg& s = c.s;
    const float al = (float)s.alpha;
#pragma unroll
    for (int j = 0; j < EPT; ++j) {
      int gr = it.m0 + lr[j], gk = c.k + lk[j];
      bool ok = gr < it.M && gk < s.k1;
      int64_t e = AKC ? s.a_off + (int64_t)gr * s.K + gk : s.a_off + (int64_t)gk * it.M + gr;
      if (CPLX) {
        ra[0][j] = ok ? A[2 * e] * al : 0.f;
        ra[NP - 1][j] = ok ? A[2 * e + 1] * al : 0.f;
      } else {
        ra[0][j] = ok ? A[e] * al : 0.f;
      }
      int gn = it.n0 + mr[j], gk2 = c.k + mk[j];
      bool ok2 = gn < it.N && gk2 < s.k1;
      int64_t f = BKC ? s.b_off + (int64_t)gn * s.K + gk2 : s.b_off + (int64_t)gk2 * it.N + gn;
      if (CPLX) {
        rb[0][j] = ok2 ? B[2 * f] : 0.f;
        rb[NP - 1][j] = ok2 ? B[2 * f + 1] : 0.f;
      } else {
        rb[0][j] = ok2 ? B[f] : 0.f;
      }
    }
  };
  if (cur.valid()) gload(cur);
  while (cur.valid()) {
    __syncthreads();
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
      for (int j = 0; j < EPT; ++j) {
        unsigned short h = f2bf_rne(ra[p][j]);
        sA[p][0][lr[j] * LDK + lk[j]] = h;
        sA[p][1][lr[j] * LDK + lk[j]] = f2bf_rne(ra[p][j] - bf2f(h));
        unsigned short g = f2bf_rne(rb[p][j]);
        sB[p][0][mr[j] * LDK + mk[j]] = g;
        sB[p][1][mr[j] * LDK + mk[j]] = f2bf_rne(rb[p][j] - bf2f(g));
      }
    __syncthreads();
    cur.next<BK>(segs);
    if (cur.valid()) gload(cur);
    const int kx = (lane >> 4) * 8;
    bf8v a[NP][2][2], b[NP][2][2];  // [part][hi/lo][tile]
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
      for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int i = 0; i < 2; ++i) {
          u16x8 va = *reinterpret_cast<const u16x8*>(&sA[p][h][(wm + i * 16 + (lane & 15)) * LDK + kx]);
          u16x8 vb = *reinterpret_cast<const u16x8*>(&sB[p][h][(wn + i * 16 + (lane & 15)) * LDK + kx]);
          a[p][h][i] = __builtin_bit_cast(bf8v, va);
          b[p][h][i] = __builtin_bit_cast(bf8v, vb);
        }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        // real part (and the whole product when real): Ar*Br (- Ai*Bi)
        acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][1][i], b[0][0][j], acc[0][i][j], 0, 0, 0);
        acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][0][i], b[0][1][j], acc[0][i][j], 0, 0, 0);
        acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][0][i], b[0][0][j], acc[0][i][j], 0, 0, 0);
        if (CPLX) {
          u16x8 m0 = __builtin_bit_cast(u16x8, a[NP - 1][0][i]) ^ (unsigned short)0x8000;
          u16x8 m1 = __builtin_bit_cast(u16x8, a[NP - 1][1][i]) ^ (unsigned short)0x8000;
          bf8v nih = __builtin_bit_cast(bf8v, m0), nil = __builtin_bit_cast(bf8v, m1);
          acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nil, b[NP - 1][0][j], acc[0][i][j], 0, 0, 0);
          acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nih, b[NP - 1][1][j], acc[0][i][j], 0, 0, 0);
          acc[0][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(nih, b[NP - 1][0][j], acc[0][i][j], 0, 0, 0);
          // imaginary part: Ar*Bi + Ai*Br
          acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][1][i], b[1][0][j], acc[1][i][j], 0, 0, 0);
          acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][0][i], b[1][1][j], acc[1][i][j], 0, 0, 0);
          acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][0][i], b[1][0][j], acc[1][i][j], 0, 0, 0);
          acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][1][i], b[0][0][j], acc[1][i][j], 0, 0, 0);
          acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][0][i], b[0][1][j], acc[1][i][j], 0, 0, 0);
          acc[1][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][0][i], b[0][0][j], acc[1][i][j], 0, 0, 0);
        }
      }
  }

  const bool partial = it.flags & ITEM_PARTIAL;
  const bool beta = !partial && (it.flags & ITEM_BETA);
  auto* D = partial ? W : C;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        int row = it.m0 + wm + i * 16 + (lane >> 4) * 4 + r;
        int col = it.n0 + wn + j * 16 + (lane & 15);
        if (row < it.M && col < it.N) {
          int64_t e = it.c_off + (int64_t)row * it.N + col;
          if (CPLX) {
            float vr = acc[0][i][j][r], vi = acc[NP - 1][i][j][r];
            if (beta) { D[2 * e] += vr; D[2 * e + 1] += vi; }
            else { D[2 * e] = vr; D[2 * e + 1] = vi; }
          } else {
            float v = acc[0][i][j][r];
            if (beta) D[e] += v;
            else D[e] = v;
          }
        }
      }
}


// x -> bf16 hi, lo planes (hi = bf16(x), lo = bf16(x - hi)) of part 0 (re) or 1 (im)
__device__ __forceinline__ void store_split(unsigned short* P, int64_t ps, int64_t e, int part, float x) {
  const unsigned short h = f2bf_rne(x);
  const unsigned short l = f2bf_rne((float)((double)x - (double)bf2f(h)));
  P[(2 * part) * ps + e] = h;
  P[(2 * part + 1) * ps + e] = l;
}

template <bool CPLX, int OUT>
__global__ __launch_bounds__(256) void k_reduce_ws_out(const float* __restrict__ W, void* __restrict__ Cv,
                                                       int64_t cps, const RedTask* __restrict__ t, int n,
                                                       const int64_t* __restrict__ start,
                                                       const int32_t* __restrict__ RM) {
  int lo = 0, hi = n - 1;
  const int64_t b = blockIdx.x;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  const RedTask r = t[lo];
  const int64_t e = (b - start[lo]) * 256 + threadIdx.x;
  if (e >= r.len) return;
  constexpr int c = CPLX ? 2 : 1;
  // the parts in chunk order (unrolled so that four loads are in flight);
  // complex elements read as one float2
  float acc[c];
  if (CPLX) {
    const float2* W2 = reinterpret_cast<const float2*>(W);
    float ar = 0, ai = 0;
#pragma unroll 4
    for (int k = 0; k < r.nch; ++k) {
      const float2 v = W2[r.w_off + k * r.len + e];
      ar += v.x;
      ai += v.y;
    }
    acc[0] = ar;
    acc[c - 1] = ai;
  } else {
    float a = 0;
#pragma unroll 4
    for (int k = 0; k < r.nch; ++k) a += W[r.w_off + k * r.len + e];
    acc[0] = a;
  }
  const int64_t o = OUT == OUT_PLANES ? out_index_planes(r, RM, e) : out_index(r, RM, e);
#pragma unroll
  for (int q = 0; q < c; ++q) {
    if (OUT == OUT_F64) {
      double* C = (double*)Cv;
      if (r.beta) C[c * o + q] += acc[q];
      else C[c * o + q] = acc[q];
    } else {
      store_split((unsigned short*)Cv, cps, o, q, acc[q]);
    }
  }
}

// ---------------------------------------------------------------------------
// BF16x3 on pre-split operands: A, B are bfloat16 planes (re_hi, re_lo[, im_hi,
// im_lo]); 128x128 output tiles, eight wavefronts of 32x64 (2x4 MFMA 16x16x32
// tiles); 16-byte vector loads staged through registers into LDS.
constexpr int PT = 128, PBK = 32, PLDK = PBK + 8;
// eight wavefronts per 128x128 tile, each 32 rows x 64 columns (2 x 4 MFMA tiles)
constexpr int PNT = 512;

// Row of the q-th 64-byte row segment a work-group stages per plane: within
// each group of 16 rows the 4 x 4 (q % 4, q / 4) order is transposed, so the
// four rows that 16 consecutive lanes store (4 lanes per row) are 4 rows apart
// (320 bytes with the 80-byte padded LDS rows: disjoint banks), while each
// 4-lane group still loads 64 contiguous bytes of one row from memory.
__device__ __forceinline__ int stage_row(int q) { return (q & ~15) | ((q & 3) << 2) | ((q >> 2) & 3); }

__device__ __forceinline__ uint4 xor_sign(uint4 v) {
  v.x ^= 0x80008000u; v.y ^= 0x80008000u; v.z ^= 0x80008000u; v.w ^= 0x80008000u;
  return v;
}

// HI: the coarse product of the hi planes alone (one bf16 product per real
// product, the lo planes neither loaded nor multiplied), for operators that
// need only a few digits
template <bool CPLX, int OUT, bool HI = false>
__global__ __launch_bounds__(PNT) void k_gemm_bf3p(const GemmItem* __restrict__ items, int nitems, int per_xcd,
                                                  const GemmSeg* __restrict__ segs,
                                                  const unsigned short* __restrict__ A, int64_t as,
                                                  const unsigned short* __restrict__ B, int64_t bs,
                                                  void* __restrict__ Cv, int64_t cps, float* __restrict__ W,
                                                  const int32_t* __restrict__ RM) {
  constexpr int NPL = CPLX ? 4 : 2;
  constexpr int NP = CPLX ? 2 : 1;
  constexpr int CPR = PBK / 8;               // 16-byte chunks per row per plane (4)
  constexpr int CH = PT * CPR * NPL;         // chunks per operand per k-step
  constexpr int CPT = CH / PNT;              // chunks per thread (4 complex, 2 real)
  const int bid = blockIdx.x;
  const int item_id = (bid & 7) * per_xcd + (bid >> 3);
  if (item_id >= nitems) return;
  const GemmItem it = items[item_id];

  // two LDS stages: the next k-step is stored while this one is multiplied,
  // so that one barrier per k-step suffices
  __shared__ __attribute__((aligned(16))) unsigned short sA[2][NPL][PT * PLDK];
  __shared__ __attribute__((aligned(16))) unsigned short sB[2][NPL][PT * PLDK];
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = (wave >> 1) * 32, wn = (wave & 1) * 64;
  f4v acc[NP][2][4];
#pragma unroll
  for (int p = 0; p < NP; ++p)
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[p][i][j] = f4v{0, 0, 0, 0};

  uint4 ra[CPT], rb[CPT];
  // chunk j of thread t is chunk t of plane j (PNT = PT * CPR): one row and
  // k offset per thread, the plane stride added per chunk
  static_assert(PNT == PT * CPR && CPT == NPL, "one chunk per plane and thread");
  const int srow = stage_row(t / CPR), skc = (t % CPR) * 8;
  // element offsets of this thread's chunk in plane 0, recomputed once per segment
  int64_t offA = 0, offB = 0;
  bool okA = false, okB = false;  // the chunk's row is inside the block
  Cursor cur;
  cur.start(segs, it.seg_begin, it.seg_end);
  // this thread's A row (through the row map), -1 outside the tile
  const int arow = it.m0 + srow < it.M ? (it.row_map >= 0 ? RM[it.row_map + it.m0 + srow] : it.m0 + srow) : -1;
  auto seg_setup = [&](const GemmSeg& s) {
    const int gn = it.n0 + srow;
    offA = s.a_off + (int64_t)max(arow, 0) * s.lda + skc;
    offB = s.b_off + (int64_t)gn * s.ldb + skc;
    okA = arow >= 0;
    okB = gn < it.N;
  };
  if (cur.valid()) seg_setup(cur.s);
  // plane plans are 8-element aligned (operand offsets, row strides and K
  // cuts) and their rows are zero-padded to a multiple of 8, so a 16-byte
  // chunk is either inside [k0, k1) or wholly past k1
  auto gload = [&](const Cursor& c) {
    const bool neg = c.s.alpha < 0;
    const bool in = c.k + skc < c.s.k1;
    const bool la = okA && in, lb = okB && in;
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      if (HI && (j & 1)) continue;  // lo planes (re_lo, im_lo)
      uint4 v = make_uint4(0, 0, 0, 0);
      if (la) v = *reinterpret_cast<const uint4*>(A + j * as + offA + c.k);
      ra[j] = neg ? xor_sign(v) : v;
      uint4 w = make_uint4(0, 0, 0, 0);
      if (lb) w = *reinterpret_cast<const uint4*>(B + j * bs + offB + c.k);
      rb[j] = w;
    }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int j = 0; j < CPT; ++j) {
      if (HI && (j & 1)) continue;
      *reinterpret_cast<uint4*>(&sA[buf][j][srow * PLDK + skc]) = ra[j];
      *reinterpret_cast<uint4*>(&sB[buf][j][srow * PLDK + skc]) = rb[j];
    }
  };
  // the cursor depends on the work item only, so every thread of the block
  // takes the same number of trips and meets the same barriers
  if (cur.valid()) {
    gload(cur);
    stage(0);
  }
  __syncthreads();
  int buf = 0;
  while (cur.valid()) {
    const int prev_seg = cur.seg;
    cur.next<PBK>(segs);
    const bool more = cur.valid();
    if (more) {
      if (cur.seg != prev_seg) seg_setup(cur.s);
      gload(cur);
    }
    const int kx = (lane >> 4) * 8;
    bf8v a[NPL][2];
#pragma unroll
    for (int p = 0; p < NPL; ++p)
#pragma unroll
      for (int i = 0; i < 2; ++i)
        a[p][i] = *reinterpret_cast<const bf8v*>(&sA[buf][p][(wm + i * 16 + (lane & 15)) * PLDK + kx]);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      bf8v b[NPL];
#pragma unroll
      for (int p = 0; p < NPL; ++p)
        b[p] = *reinterpret_cast<const bf8v*>(&sB[buf][p][(wn + j * 16 + (lane & 15)) * PLDK + kx]);
      bf8v nbh, nbl;
      if (CPLX) {
        nbh = __builtin_bit_cast(bf8v, __builtin_bit_cast(u16x8, b[2]) ^ (unsigned short)0x8000);
        nbl = __builtin_bit_cast(bf8v, __builtin_bit_cast(u16x8, b[3]) ^ (unsigned short)0x8000);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        if (HI) {
          f4v r = acc[0][i][j];
          if (CPLX) r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2][i], nbh, r, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][i], b[0], r, 0, 0, 0);
          acc[0][i][j] = r;
          if (CPLX) {
            f4v q = acc[NP - 1][i][j];
            q = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][i], b[2], q, 0, 0, 0);
            q = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2][i], b[0], q, 0, 0, 0);
            acc[NP - 1][i][j] = q;
          }
          continue;
        }
        f4v r = acc[0][i][j];
        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][i], b[0], r, 0, 0, 0);
        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][i], b[1], r, 0, 0, 0);
        if (CPLX) {
          r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[3][i], nbh, r, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2][i], nbl, r, 0, 0, 0);
          r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2][i], nbh, r, 0, 0, 0);
        }
        r = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][i], b[0], r, 0, 0, 0);
        acc[0][i][j] = r;
        if (CPLX) {
          f4v q = acc[NP - 1][i][j];
          q = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][i], b[2], q, 0, 0, 0);
          q = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][i], b[3], q, 0, 0, 0);
          q = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[3][i], b[0], q, 0, 0, 0);
          q = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2][i], b[1], q, 0, 0, 0);
          q = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][i], b[2], q, 0, 0, 0);
          q = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2][i], b[0], q, 0, 0, 0);
          acc[NP - 1][i][j] = q;
        }
      }
    }
    // the other stage was last read before the previous barrier
    if (more) stage(buf ^ 1);
    __syncthreads();
    buf ^= 1;
  }

  const bool partial = it.flags & ITEM_PARTIAL;
  const bool beta = !partial && (it.flags & ITEM_BETA);
  // where each of the lane's eight output rows starts (-1: past the block):
  // through the row map, and into the padded layout for plane output
  int64_t rstart[2][4];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      int row = it.m0 + wm + i * 16 + (lane >> 4) * 4 + r;
      int64_t st = -1;
      if (row < it.M) {
        if (!partial && it.row_map >= 0) row = RM[it.row_map + row];
        st = it.c_off + (int64_t)row * it.N;
        if (OUT == OUT_PLANES && !partial && it.op.kp > 0) {
          const int rr = it.op.r0 + row;
          st = it.op.base + (int64_t)(rr / it.op.q) * it.op.kp + (int64_t)(rr % it.op.q) * it.N;
        }
      }
      rstart[i][r] = st;
    }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int col = it.n0 + wn + j * 16 + (lane & 15);
        if (rstart[i][r] >= 0 && col < it.N) {
          const int64_t e = rstart[i][r] + col;
          const float vr = acc[0][i][j][r], vi = CPLX ? acc[NP - 1][i][j][r] : 0.0f;
          if (partial || OUT == OUT_F32) {
            float* D = partial ? W : (float*)Cv;
            if (CPLX) {
              if (beta) { D[2 * e] += vr; D[2 * e + 1] += vi; }
              else { D[2 * e] = vr; D[2 * e + 1] = vi; }
            } else {
              if (beta) D[e] += vr;
              else D[e] = vr;
            }
          } else if (OUT == OUT_F64) {
            double* D = (double*)Cv;
            if (CPLX) {
              if (beta) { D[2 * e] += vr; D[2 * e + 1] += vi; }
              else { D[2 * e] = vr; D[2 * e + 1] = vi; }
            } else {
              if (beta) D[e] += vr;
              else D[e] = vr;
            }
          } else {  // OUT_PLANES: the plan has no accumulating outputs
            unsigned short* P = (unsigned short*)Cv;
            store_split(P, cps, e, 0, vr);
            if (CPLX) store_split(P, cps, e, 1, vi);
          }
        }
      }
}

// hi / lo bf16 parts of x: hi = x rounded to bf16 (through FP32), lo = the
// rest rounded likewise (for an FP32 x the rest is exact in FP32)
__device__ __forceinline__ void split_bf(float x, unsigned short& h, unsigned short& l) {
  h = f2bf_rne(x);
  l = f2bf_rne(x - bf2f(h));
}
__device__ __forceinline__ void split_bf(double x, unsigned short& h, unsigned short& l) {
  h = f2bf_rne((float)x);
  l = f2bf_rne((float)(x - (double)bf2f(h)));
}

// Two consecutive padded elements per thread (Kp and every block start are
// multiples of 8, so the pair lies in one row), consecutive threads on
// consecutive pairs: coalesced loads of the source and 4-byte stores to each
// plane.  A work-group walks `per` chunks of 512 elements (per is chosen so
// that the launch has about a thousand work-groups), finding the block of
// its first pair by binary search and the later ones by stepping forward.
// Padding columns (K <= col < Kp) are written as 0.
template <typename SRC>
__global__ __launch_bounds__(256) void k_split_padded(const SRC* __restrict__ src, unsigned short* __restrict__ dst,
                                                      int64_t ps, const PadBlk* __restrict__ blk, int nb,
                                                      int64_t total, int cplx, int per) {
  using SRC2 = typename std::conditional<std::is_same<SRC, float>::value, float2, double2>::type;
  const int64_t step = (int64_t)gridDim.x * 512 * per;
  for (int64_t c0 = (int64_t)blockIdx.x * 512 * per; c0 < total; c0 += step) {
    int64_t e = c0 + 2 * threadIdx.x;
    int bi = 0, hi = nb - 1;
    while (bi < hi) {
      const int mid = (bi + hi + 1) >> 1;
      if (blk[mid].dst_off <= e) bi = mid;
      else hi = mid - 1;
    }
    for (int i = 0; i < per && e < total; ++i, e += 512) {
      while (bi + 1 < nb && blk[bi + 1].dst_off <= e) ++bi;
      const PadBlk b = blk[bi];
      const int loc = (int)(e - b.dst_off);  // blocks hold < 2^31 elements
      const int row = loc / b.Kp, col = loc - row * b.Kp;
      const bool ok0 = row < b.rows && col < b.K, ok1 = row < b.rows && col + 1 < b.K;
      const int64_t si = b.src_off + (int64_t)row * b.K + col;
      unsigned short h[2][2], l[2][2];  // [part][element]
      if (cplx) {
        const SRC2* s2 = reinterpret_cast<const SRC2*>(src);
        SRC2 x0, x1;
        x0.x = x0.y = x1.x = x1.y = 0;
        if (ok0) x0 = s2[si];
        if (ok1) x1 = s2[si + 1];
        split_bf(x0.x, h[0][0], l[0][0]);
        split_bf(x0.y, h[1][0], l[1][0]);
        split_bf(x1.x, h[0][1], l[0][1]);
        split_bf(x1.y, h[1][1], l[1][1]);
      } else {
        split_bf(ok0 ? src[si] : (SRC)0, h[0][0], l[0][0]);
        split_bf(ok1 ? src[si + 1] : (SRC)0, h[0][1], l[0][1]);
      }
      for (int part = 0; part <= cplx; ++part) {
        *reinterpret_cast<unsigned*>(dst + (2 * part) * ps + e) = h[part][0] | ((unsigned)h[part][1] << 16);
        *reinterpret_cast<unsigned*>(dst + (2 * part + 1) * ps + e) = l[part][0] | ((unsigned)l[part][1] << 16);
      }
    }
  }
}

__global__ void k_f64_to_f32(const double* __restrict__ s, float* __restrict__ d, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) d[i] = (float)s[i];
}
__global__ void k_f32_to_f64(const float* __restrict__ s, double* __restrict__ d, int64_t n) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) d[i] = (double)s[i];
}

template <typename S, bool CPLX>
void launch4(const GemmPlan& p, const void* A, const void* B, void* C, hipStream_t s, int per) {
  auto it = p.items;
  auto sg = p.segs;
  const bool akc = (p.ta == 'N' || p.ta == 'n'), bkc = (p.tb == 'T' || p.tb == 't');
  dim3 g(p.grid), b(NT);
#define L4(AK, BKc) hipLaunchKernelGGL((k_gemm4<S, CPLX, AK, BKc>), g, b, 0, s, it, p.n_items, per, sg, \
                                       (const S*)A, (const S*)B, (S*)C, (S*)p.ws, p.row_maps)
  if (akc && bkc) L4(true, true);
  else if (akc) L4(true, false);
  else if (bkc) L4(false, true);
  else L4(false, false);
#undef L4
}

template <bool CPLX>
void launch_bf(const GemmPlan& p, const void* A, const void* B, void* C, hipStream_t s, int per) {
  auto it = p.items;
  auto sg = p.segs;
  const bool akc = (p.ta == 'N' || p.ta == 'n'), bkc = (p.tb == 'T' || p.tb == 't');
  dim3 g(p.grid), b(NT);
#define LB(AK, BKc) hipLaunchKernelGGL((k_gemm_bf16x3<CPLX, AK, BKc>), g, b, 0, s, it, p.n_items, per, sg, \
                                       (const float*)A, (const float*)B, (float*)C, \
                                       (float*)p.ws)
  if (akc && bkc) LB(true, true);
  else if (akc) LB(true, false);
  else if (bkc) LB(false, true);
  else LB(false, false);
#undef LB
}

}  // namespace

std::unique_ptr<GemmPlan> make_gemm_plan(const std::vector<GemmTask>& tasks_in, int dtype, int prec,
                                         char ta, char tb,
                                         const std::vector<std::pair<int64_t, int64_t>>& extra_zero,
                                         int split_target, bool planes, GemmTables* host_only,
                                         const std::vector<int32_t>* row_maps, const PadLayout* out_pad) {
  if (!host_only) ensure_init();
  auto plan = std::make_unique<GemmPlan>();
  plan->planes = planes;
  plan->tile = planes ? 128 : 64;
  const int BM = plan->tile, BN = plan->tile;
  DMRGX_CHECK(!planes || (prec == PREC_BF16X3 && (ta == 'N' || ta == 'n') && (tb == 'T' || tb == 't')),
              "pre-split plans are BF16x3 with ('N','T') operands");
  plan->dtype = dtype;
  plan->prec = prec;
  plan->ta = ta;
  plan->tb = tb;
  const int BKq = prec == PREC_BF16X3 ? 32 : 16;
  const double cf = dtype == DT_C64 ? 8.0 : 2.0;

  // group valid tasks by output block (and row map), first-appearance order
  struct Group { int64_t c_off; int64_t row_map; int M, N; int beta; std::vector<int> t; int64_t ksum = 0; };
  std::vector<Group> groups;
  std::map<std::pair<int64_t, int64_t>, int> gidx;
  const bool akc_ = (ta == 'N' || ta == 'n');
  for (int i = 0; i < (int)tasks_in.size(); ++i) {
    const GemmTask& k = tasks_in[i];
    const int kb = k.k0, ke = k.k1 > 0 ? k.k1 : k.K;
    if (k.M <= 0 || k.N <= 0 || k.K <= 0 || ke <= kb) continue;  // BLAS wrapper skips empty GEMMs
    DMRGX_CHECK(kb >= 0 && ke <= k.K, "gemm plan: K range outside the task");
    if (k.row_map >= 0) {
      DMRGX_CHECK(row_maps && k.row_map + k.M <= (int64_t)row_maps->size(), "gemm plan: row map out of range");
      DMRGX_CHECK(akc_ && (planes || prec != PREC_BF16X3), "gemm plan: row maps need a k-contiguous A");
    }
    const std::pair<int64_t, int64_t> key{k.c_off, k.row_map};
    auto f = gidx.find(key);
    if (f == gidx.end()) {
      gidx[key] = (int)groups.size();
      groups.push_back(Group{k.c_off, k.row_map, k.M, k.N, k.beta, {}, 0});
      f = gidx.find(key);
    }
    Group& g = groups[f->second];
    DMRGX_CHECK(g.M == k.M && g.N == k.N, "gemm plan: tasks into one output block disagree on shape");
    // the pre-split kernel loads whole 16-byte chunks: operand offsets, row
    // strides and K ranges on 8-element boundaries, rows zero-padded to a
    // multiple of 8
    DMRGX_CHECK(!planes || ((k.a_off | k.b_off | (k.lda > 0 ? k.lda : k.K) | (k.ldb > 0 ? k.ldb : k.K) | kb) & 7) == 0,
                "pre-split gemm plan: operand offsets, row strides and K ranges must be multiples of 8");
    g.t.push_back(i);
    g.ksum += ke - kb;
    plan->flops += cf * (double)k.M * k.N * (ke - kb);
  }
  // split-K so the launch has enough work items for every CU
  double work = 0;
  for (auto& g : groups) {
    int64_t tl = (int64_t)((g.M + BM - 1) / BM) * ((g.N + BN - 1) / BN);
    work += (double)tl * g.ksum;
  }
  int target = split_target > 0 ? split_target : 2 * std::max(rt().num_cu, 1);
  int64_t kchunk = (int64_t)std::max(8.0 * BKq, work / std::max<double>(target, 1.0));
  kchunk = (kchunk + BKq - 1) / BKq * BKq;

  // Balanced mode (split_target < 0, -split_target work-groups resident at
  // once): whole output blocks are cut into two K halves where that shortens
  // the launch.  Work-groups go round-robin to the 8 XCDs and each XCD runs
  // a contiguous range of items, so every XCD gets its share of whole tiles
  // first and then its share of half tiles; a list-scheduling model of those
  // per-XCD queues picks how many blocks (the ones with the most K steps) to
  // cut, charging the partial tiles' extra traffic.
  const bool balanced = split_target < 0;
  const size_t G = groups.size();
  std::vector<int> gnch(G, 1);
  if (balanced && G > 0) {
    const int slots = std::max(1, -split_target / 8);
    std::vector<int64_t> tl(G), ks(G);
    for (size_t q = 0; q < G; ++q) {
      tl[q] = (int64_t)((groups[q].M + BM - 1) / BM) * ((groups[q].N + BN - 1) / BN);
      ks[q] = (groups[q].ksum + BKq - 1) / BKq;
    }
    std::vector<size_t> ord(G);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](size_t a, size_t b) { return ks[a] > ks[b]; });
    // partial elements whose write and read-back cost about one tile K step
    const double red_unit = dtype == DT_C64 ? 6e5 : 1.2e6;
    auto cost = [&](const std::vector<int>& nc) {
      std::vector<int64_t> F, H;  // item lengths in K steps, in emission order
      for (size_t q = 0; q < G; ++q) {
        if (nc[q] == 1) F.insert(F.end(), tl[q], ks[q]);
        else H.insert(H.end(), 2 * tl[q], (ks[q] + 1) / 2);
      }
      int64_t worst = 0;
      for (int x = 0; x < 8; ++x) {
        std::priority_queue<int64_t, std::vector<int64_t>, std::greater<int64_t>> pq;
        for (int i = 0; i < slots; ++i) pq.push(0);
        auto run = [&](const std::vector<int64_t>& L) {
          for (size_t i = L.size() * x / 8; i < L.size() * (x + 1) / 8; ++i) {
            const int64_t t0 = pq.top();
            pq.pop();
            pq.push(t0 + L[i] + 1);  // +1: prologue and epilogue of an item
          }
        };
        run(F);
        run(H);
        while (pq.size() > 1) pq.pop();
        worst = std::max(worst, pq.top());
      }
      double extra = 0;
      for (size_t q = 0; q < G; ++q)
        if (nc[q] == 2) extra += 2.0 * groups[q].M * groups[q].N;
      return (double)worst + extra / red_unit;
    };
    // the model runs at every plan build: cut 0, 1/8, ..., 8/8 of the blocks
    std::vector<int> nc(G, 1), best_nc = nc;
    double best = cost(nc);
    size_t done = 0;
    for (int f = 1; f <= 8; ++f) {
      const size_t x = (G * f + 7) / 8;
      if (x <= done) continue;
      for (size_t q = done; q < x; ++q) nc[ord[q]] = ks[ord[q]] > 1 ? 2 : 1;
      done = x;
      const double c = cost(nc);
      if (c < best) {
        best = c;
        best_nc = nc;
      }
    }
    gnch = best_nc;
  }

  // plane output into a padded layout: each output block must be a run of
  // whole N-element rows of one padded block whose K is a multiple of N
  std::vector<OutPad> gop(G, OutPad{0, 0, 0, 0, 0});
  if (out_pad && planes && !out_pad->identity && !host_only) {
    bool ok = true;
    const auto& PB = out_pad->blk;
    for (size_t gi = 0; gi < G && ok; ++gi) {
      const Group& g = groups[gi];
      int64_t rows = g.M;  // rows of the block the group writes (its row map's largest index + 1)
      if (g.row_map >= 0) {
        rows = 0;
        for (int i = 0; i < g.M; ++i) rows = std::max<int64_t>(rows, (*row_maps)[g.row_map + i] + 1);
      }
      auto it = std::upper_bound(PB.begin(), PB.end(), g.c_off,
                                 [](int64_t v, const PadBlk& b) { return v < b.src_off; });
      if (it == PB.begin()) { ok = false; break; }
      const PadBlk& b = *(it - 1);
      const int64_t rel = g.c_off - b.src_off;
      if (b.K % g.N != 0 || rel % g.N != 0 || rel + rows * g.N > (int64_t)b.rows * b.K) { ok = false; break; }
      gop[gi] = OutPad{b.dst_off, (int)(rel / g.N), b.K / g.N, b.Kp, 0};
    }
    if (ok) plan->out_padded = true;
    else std::fill(gop.begin(), gop.end(), OutPad{0, 0, 0, 0, 0});
  }

  std::vector<GemmItem> items, half;  // balanced mode: whole tiles, then half tiles
  std::vector<GemmSeg> segs;
  std::vector<RedTask> red;
  for (size_t gi = 0; gi < G; ++gi) {
    const Group& g = groups[gi];
    int nch = balanced ? gnch[gi] : (int)std::max<int64_t>(1, (g.ksum + kchunk - 1) / kchunk);
    int64_t ch = (g.ksum + nch - 1) / nch;
    ch = (ch + BKq - 1) / BKq * BKq;
    if (g.beta) plan->any_beta = true;
    // cut the concatenated K of the group's tasks into chunks of about ch; a
    // cut inside a task falls on a multiple of 8 so that every segment starts
    // on a 16-byte boundary of an aligned operand row
    std::vector<std::pair<int, int>> chunk_rng;  // seg range per chunk
    size_t ti = 0;
    int64_t kin = g.t.empty() ? 0 : tasks_in[g.t[0]].k0;  // position inside current task
    while (ti < g.t.size()) {
      int sb = (int)segs.size();
      int64_t need = ch;
      while (need > 0 && ti < g.t.size()) {
        const GemmTask& k = tasks_in[g.t[ti]];
        const int64_t kend = k.k1 > 0 ? k.k1 : k.K;
        int64_t take = std::min<int64_t>(need, kend - kin);
        if (take < kend - kin) {
          take = take / 8 * 8;
          if (take == 0) break;
        }
        GemmSeg s{};
        s.a_off = k.a_off;
        s.b_off = k.b_off;
        s.K = k.K;
        s.lda = k.lda > 0 ? k.lda : k.K;
        s.ldb = k.ldb > 0 ? k.ldb : k.K;
        s.k0 = (int)kin;
        s.k1 = (int)(kin + take);
        s.alpha = k.alpha;
        segs.push_back(s);
        need -= take;
        kin += take;
        if (kin >= kend) {
          ++ti;
          if (ti < g.t.size()) kin = tasks_in[g.t[ti]].k0;
        }
      }
      DMRGX_CHECK((int)segs.size() > sb, "gemm plan: empty K chunk");
      chunk_rng.push_back({sb, (int)segs.size()});
    }
    nch = (int)chunk_rng.size();
    const bool shared = nch > 1;
    const int64_t len = (int64_t)g.M * g.N;
    int64_t w_off = 0;
    if (shared) {
      w_off = plan->ws_elems;
      plan->ws_elems += len * nch;
      red.push_back(RedTask{g.c_off, w_off, len, g.row_map, nch, g.beta, g.N, 0, gop[gi]});
    }
    for (int c = 0; c < nch; ++c)
      for (int n0 = 0; n0 < g.N; n0 += BN)
        for (int m0 = 0; m0 < g.M; m0 += BM) {
          GemmItem itm{};
          itm.c_off = shared ? w_off + (int64_t)c * len : g.c_off;
          // A rows always go through the row map; a partial tile is written
          // compact into the workspace and the reduction maps its rows
          itm.row_map = g.row_map;
          itm.M = g.M;
          itm.N = g.N;
          itm.m0 = m0;
          itm.n0 = n0;
          itm.seg_begin = chunk_rng[c].first;
          itm.seg_end = chunk_rng[c].second;
          itm.flags = shared ? ITEM_PARTIAL : (g.beta ? ITEM_BETA : 0);
          itm.op = gop[gi];
          (balanced && shared ? half : items).push_back(itm);
        }
  }
  if (balanced) {
    // XCD x runs items [x per, (x+1) per): its eighth of the whole tiles, then
    // its eighth of the half tiles, padded with empty items (no rows)
    std::vector<GemmItem> whole;
    whole.swap(items);
    auto share = [](size_t n, int x) { return std::make_pair(n * x / 8, n * (x + 1) / 8); };
    size_t per = 0;
    for (int x = 0; x < 8; ++x) {
      auto f = share(whole.size(), x), h = share(half.size(), x);
      per = std::max(per, (f.second - f.first) + (h.second - h.first));
    }
    GemmItem empty{};
    empty.row_map = -1;
    items.assign(per * 8, empty);
    for (int x = 0; x < 8; ++x) {
      size_t pos = per * x;
      auto f = share(whole.size(), x), h = share(half.size(), x);
      for (size_t i = f.first; i < f.second; ++i) items[pos++] = whole[i];
      for (size_t i = h.first; i < h.second; ++i) items[pos++] = half[i];
    }
  }
  for (auto& z : extra_zero) plan->zero_ranges.push_back(z);
  plan->n_items = (int)items.size();
  plan->n_segs = (int)segs.size();
  int per = (plan->n_items + 7) / 8;
  plan->grid = per * 8;
  std::vector<int64_t> start(red.size());
  int64_t acc = 0;
  for (size_t i = 0; i < red.size(); ++i) {
    start[i] = acc;
    acc += (red[i].len + 255) / 256;
  }
  plan->n_red = (int)red.size();
  plan->red_blocks = (int)acc;
  const bool any_map = std::any_of(groups.begin(), groups.end(), [](const Group& g) { return g.row_map >= 0; });
  if (host_only) {
    DMRGX_CHECK(!any_map, "gemm plan: row maps need device tables");
    host_only->items = std::move(items);
    host_only->segs = std::move(segs);
    host_only->red = std::move(red);
    host_only->red_start = std::move(start);
    return plan;
  }
  hipStream_t s = rt().stream;
  if (any_map) {
    plan->own_row_maps = std::make_unique<DevBuf>(row_maps->size() * sizeof(int32_t));
    upload(plan->own_row_maps->ptr, row_maps->data(), row_maps->size() * sizeof(int32_t), s);
    plan->row_maps = (const int32_t*)plan->own_row_maps->ptr;
  }
  if (plan->n_items) {
    plan->own_items = std::make_unique<DevBuf>(items.size() * sizeof(GemmItem));
    plan->own_segs = std::make_unique<DevBuf>(std::max<size_t>(segs.size(), 1) * sizeof(GemmSeg));
    upload(plan->own_items->ptr, items.data(), items.size() * sizeof(GemmItem), s);
    if (!segs.empty()) upload(plan->own_segs->ptr, segs.data(), segs.size() * sizeof(GemmSeg), s);
    plan->items = (const GemmItem*)plan->own_items->ptr;
    plan->segs = (const GemmSeg*)plan->own_segs->ptr;
  }
  if (!red.empty()) {
    plan->own_ws = std::make_unique<DevBuf>(gemm_ws_bytes(*plan));
    plan->own_red = std::make_unique<DevBuf>(red.size() * sizeof(RedTask));
    plan->own_red_start = std::make_unique<DevBuf>(red.size() * 8);
    upload(plan->own_red->ptr, red.data(), red.size() * sizeof(RedTask), s);
    upload(plan->own_red_start->ptr, start.data(), start.size() * 8, s);
    plan->ws = plan->own_ws->ptr;
    plan->red = (const RedTask*)plan->own_red->ptr;
    plan->red_start = (const int64_t*)plan->own_red_start->ptr;
  }
  return plan;
}

void reduce_partials(const GemmPlan& p, void* C, hipStream_t s) {
  if (!p.n_red) return;
  auto t = p.red;
  auto st = p.red_start;
  dim3 g(p.red_blocks), b(256);
  const bool cx = p.dtype == DT_C64;
  if (p.prec == PREC_FP64) {
    if (cx) hipLaunchKernelGGL((k_reduce_ws<double, true>), g, b, 0, s, (const double*)p.ws, (double*)C, t, p.n_red, st,
                               p.row_maps);
    else hipLaunchKernelGGL((k_reduce_ws<double, false>), g, b, 0, s, (const double*)p.ws, (double*)C, t, p.n_red, st,
                            p.row_maps);
  } else {
    if (cx) hipLaunchKernelGGL((k_reduce_ws<float, true>), g, b, 0, s, (const float*)p.ws, (float*)C, t, p.n_red, st,
                               p.row_maps);
    else hipLaunchKernelGGL((k_reduce_ws<float, false>), g, b, 0, s, (const float*)p.ws, (float*)C, t, p.n_red, st,
                            p.row_maps);
  }
  DMRGX_HIP(hipGetLastError());
}

void reduce_partials(const GemmPlan& p, void* C, hipStream_t s);

void launch_planes_kernel(const GemmPlan& p, const uint16_t* A, int64_t a_stride, const uint16_t* B,
                          int64_t b_stride, void* C, int out, int64_t c_ps, hipStream_t s, bool hi) {
  if (!p.n_items) return;
  float* wsp = (float*)p.ws;
  const int per = p.grid / 8;
  auto it = p.items;
  auto sg = p.segs;
  auto go = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3(p.grid), dim3(PNT), 0, s, it, p.n_items, per, sg, (const unsigned short*)A,
                       a_stride, (const unsigned short*)B, b_stride, C, c_ps, wsp, p.row_maps);
  };
  if (p.dtype == DT_C64) {
    if (hi) {
      if (out == OUT_F32) go(k_gemm_bf3p<true, OUT_F32, true>);
      else if (out == OUT_PLANES) go(k_gemm_bf3p<true, OUT_PLANES, true>);
      else go(k_gemm_bf3p<true, OUT_F64, true>);
    } else {
      if (out == OUT_F32) go(k_gemm_bf3p<true, OUT_F32>);
      else if (out == OUT_PLANES) go(k_gemm_bf3p<true, OUT_PLANES>);
      else go(k_gemm_bf3p<true, OUT_F64>);
    }
  } else {
    if (hi) {
      if (out == OUT_F32) go(k_gemm_bf3p<false, OUT_F32, true>);
      else if (out == OUT_PLANES) go(k_gemm_bf3p<false, OUT_PLANES, true>);
      else go(k_gemm_bf3p<false, OUT_F64, true>);
    } else {
      if (out == OUT_F32) go(k_gemm_bf3p<false, OUT_F32>);
      else if (out == OUT_PLANES) go(k_gemm_bf3p<false, OUT_PLANES>);
      else go(k_gemm_bf3p<false, OUT_F64>);
    }
  }
  DMRGX_HIP(hipGetLastError());
}

void run_gemm_planes(const GemmPlan& p, const uint16_t* A, int64_t a_stride, const uint16_t* B, int64_t b_stride,
                     void* C, int out, int64_t c_ps, hipStream_t s, hipEvent_t ev_begin, hipEvent_t ev_end,
                     bool hi) {
  DMRGX_CHECK(p.planes, "run_gemm_planes needs a pre-split plan");
  DMRGX_CHECK(a_stride % 8 == 0 && b_stride % 8 == 0, "plane strides must be multiples of 8");
  DMRGX_CHECK(out != OUT_PLANES || !p.any_beta, "run_gemm_planes: plane output cannot accumulate");
  const int cw = p.dtype == DT_C64 ? 2 : 1;
  for (auto& z : p.zero_ranges) {
    if (out == OUT_PLANES) {
      // a padded output layout was zeroed once by its owner (the plan never
      // writes these ranges, and their flat offsets are not padded ones)
      if (!p.out_padded)
        for (int q = 0; q < 2 * cw; ++q) fill_zero_u16((uint16_t*)C + q * c_ps + z.first, (size_t)z.second, s);
    } else {
      fill_zero((char*)C + z.first * (out == OUT_F64 ? 8 : 4) * cw, (size_t)z.second * (out == OUT_F64 ? 8 : 4) * cw, s);
    }
  }
  if (ev_begin) DMRGX_HIP(hipEventRecord(ev_begin, s));
  launch_planes_kernel(p, A, a_stride, B, b_stride, C, out, c_ps, s, hi);
  if (ev_end) DMRGX_HIP(hipEventRecord(ev_end, s));
  if (out == OUT_F32) {
    reduce_partials(p, C, s);
  } else if (p.n_red) {
    auto t = p.red;
    auto st = p.red_start;
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(p.red_blocks), dim3(256), 0, s, (const float*)p.ws, C, c_ps, t, p.n_red,
                         st, p.row_maps);
    };
    if (p.dtype == DT_C64) {
      if (out == OUT_PLANES) go(k_reduce_ws_out<true, OUT_PLANES>);
      else go(k_reduce_ws_out<true, OUT_F64>);
    } else {
      if (out == OUT_PLANES) go(k_reduce_ws_out<false, OUT_PLANES>);
      else go(k_reduce_ws_out<false, OUT_F64>);
    }
    DMRGX_HIP(hipGetLastError());
  }
}

int64_t PadLayout::map(int64_t src_off) const {
  auto it = std::lower_bound(blk.begin(), blk.end(), src_off,
                             [](const PadBlk& b, int64_t v) { return b.src_off < v; });
  DMRGX_CHECK(it != blk.end() && it->src_off == src_off, "padded layout: unknown block");
  return it->dst_off;
}

int PadLayout::kp(int64_t src_off) const {
  auto it = std::lower_bound(blk.begin(), blk.end(), src_off,
                             [](const PadBlk& b, int64_t v) { return b.src_off < v; });
  DMRGX_CHECK(it != blk.end() && it->src_off == src_off, "padded layout: unknown block");
  return it->Kp;
}

std::unique_ptr<PadLayout> make_pad_layout(std::vector<PadBlk> blocks) {
  auto L = std::make_unique<PadLayout>();
  std::sort(blocks.begin(), blocks.end(), [](const PadBlk& a, const PadBlk& b) { return a.src_off < b.src_off; });
  int64_t pos = 0;
  for (auto& b : blocks) {
    if (!L->blk.empty() && L->blk.back().src_off == b.src_off) {
      DMRGX_CHECK(L->blk.back().rows == b.rows && L->blk.back().K == b.K,
                  "padded layout: one block seen with two shapes");
      continue;
    }
    PadBlk q = b;
    q.Kp = (b.K + 7) / 8 * 8;
    DMRGX_CHECK((int64_t)q.rows * q.Kp < (1ll << 31), "padded layout: a block of 2^31 elements or more");
    q.dst_off = pos;
    q.pad = 0;
    if (q.dst_off != q.src_off || q.Kp != q.K) L->identity = false;
    pos += (int64_t)q.rows * q.Kp;
    L->blk.push_back(q);
  }
  L->total = pos;
  if (!L->blk.empty()) {
    L->d_blk = std::make_unique<DevBuf>(L->blk.size() * sizeof(PadBlk));
    upload(L->d_blk->ptr, L->blk.data(), L->blk.size() * sizeof(PadBlk), rt().stream);
  }
  return L;
}

template <typename SRC>
static void split_padded(const SRC* src, uint16_t* dst, int64_t ps, const PadLayout& L, bool cplx, hipStream_t s) {
  if (L.total <= 0) return;
  DMRGX_CHECK(ps >= L.total, "split_padded: plane stride below the padded length");
  DMRGX_CHECK(ps % 8 == 0 && L.total % 8 == 0, "split_padded: plane stride and length must be multiples of 8");
  // about 1024 work-groups, each walking at most 8 chunks of 512 elements
  const int per = (int)std::max<int64_t>(1, std::min<int64_t>(8, L.total / (512 * 1024)));
  const int64_t blocks = std::min<int64_t>((L.total + 512 * per - 1) / (512 * per), 8192);
  hipLaunchKernelGGL(k_split_padded<SRC>, dim3((unsigned)blocks), dim3(256), 0, s, src, (unsigned short*)dst, ps,
                     (const PadBlk*)L.d_blk->ptr, (int)L.blk.size(), L.total, cplx ? 1 : 0, per);
  DMRGX_HIP(hipGetLastError());
}

void split_padded_f64(const double* src, uint16_t* dst, int64_t ps, const PadLayout& L, bool cplx, hipStream_t s) {
  split_padded(src, dst, ps, L, cplx, s);
}

void split_padded_f32(const float* src, uint16_t* dst, int64_t ps, const PadLayout& L, bool cplx, hipStream_t s) {
  split_padded(src, dst, ps, L, cplx, s);
}

void launch_plan_kernel(const GemmPlan& p, const void* A, const void* B, void* C, hipStream_t s) {
  if (p.n_items) {
    const int per = p.grid / 8;
    if (p.prec == PREC_FP64) {
      if (p.dtype == DT_C64) launch4<double, true>(p, A, B, C, s, per);
      else launch4<double, false>(p, A, B, C, s, per);
    } else if (p.prec == PREC_FP32) {
      if (p.dtype == DT_C64) launch4<float, true>(p, A, B, C, s, per);
      else launch4<float, false>(p, A, B, C, s, per);
    } else {
      if (p.dtype == DT_C64) launch_bf<true>(p, A, B, C, s, per);
      else launch_bf<false>(p, A, B, C, s, per);
    }
    DMRGX_HIP(hipGetLastError());
  }
}

void run_gemm_plan(const GemmPlan& p, const void* A, const void* B, void* C, hipStream_t s, hipEvent_t ev_begin,
                   hipEvent_t ev_end) {
  const size_t eb = (p.prec == PREC_FP64 ? 8 : 4) * (p.dtype == DT_C64 ? 2 : 1);
  for (auto& z : p.zero_ranges) fill_zero((char*)C + z.first * eb, (size_t)z.second * eb, s);
  // events bracket the GEMM kernel only; the split-K reduction follows ev_end
  if (ev_begin) DMRGX_HIP(hipEventRecord(ev_begin, s));
  launch_plan_kernel(p, A, B, C, s);
  if (ev_end) DMRGX_HIP(hipEventRecord(ev_end, s));
  reduce_partials(p, C, s);
}

void convert_f64_to_f32(const double* src, float* dst, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_f64_to_f32, dim3((unsigned)blocks), dim3(256), 0, s, src, dst, n);
  DMRGX_HIP(hipGetLastError());
}
void convert_f32_to_f64(const float* src, double* dst, int64_t n, hipStream_t s) {
  if (n <= 0) return;
  int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
  hipLaunchKernelGGL(k_f32_to_f64, dim3((unsigned)blocks), dim3(256), 0, s, src, dst, n);
  DMRGX_HIP(hipGetLastError());
}

}  // namespace dmrgx
