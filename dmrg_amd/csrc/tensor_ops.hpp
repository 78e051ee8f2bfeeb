// Memory-bound block kernels behind the qtensor leg operations.
#pragma once

#include <vector>

#include "common.hpp"

namespace dmrgx {

// out[i1,j1,j2,i2] = sum over pairs of sign * sum_k A[i1,k,i2] B[j1,k,j2]
struct LegPair { int64_t a_off, b_off; int mid; int sign; };
struct LegOut {
  int64_t c_off;
  int al, bl, br, ar;
  int pair_begin, pair_end;
  int accumulate;  // 1: add onto existing C
};
void leg_contract(int dtype, const void* A, const void* B, void* C, const std::vector<LegOut>& outs,
                  const std::vector<LegPair>& pairs, hipStream_t s);

struct PermTask {
  int64_t in_off, out_off;
  int nleg;
  int dims[8];  // input leg dims
  int perm[8];  // output leg q = input leg perm[q]
};
void permute(int dtype, const void* in, void* out, const std::vector<PermTask>& t, hipStream_t s);

struct BoxTask {
  int64_t src_off, src_s0, src_s1;
  int64_t dst_off, dst_s0, dst_s1;
  int n0, n1, n2;
  int mode;  // 0 copy, -1 negate, 2 conjugate
};
void copy_boxes(int dtype, const void* src, void* dst, const std::vector<BoxTask>& t, hipStream_t s);
// max |x| over each source box (host result)
std::vector<double> box_maxabs(int dtype, const void* src, const std::vector<BoxTask>& t, hipStream_t s);

struct AxpbySeg {
  int64_t out_off, a_off, b_off, len;  // element offsets; a_off / b_off < 0: block absent
};
// out = alpha A + beta B block by block (complex alpha, beta when cplx)
void blocks_axpby(bool cplx, const void* A, const void* B, void* out, const std::vector<AxpbySeg>& segs,
                  int64_t total, double ar, double ai, double br, double bi, hipStream_t s);
void conj_inplace(void* p, int64_t n, hipStream_t s);  // complex elements
void conj_inplace_f32(void* p, int64_t n, hipStream_t s);  // complex FP32 elements
void scale_real(double* p, int64_t nreal, double a, hipStream_t s);

}  // namespace dmrgx
