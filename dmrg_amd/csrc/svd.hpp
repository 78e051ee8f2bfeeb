// Batched FP64 one-sided Jacobi SVD of row-major blocks.
#pragma once

#include <vector>

#include "common.hpp"

namespace dmrgx {

struct SvdBlock {
  int rows, cols;
  int64_t a_off;  // element offset of A (rows x cols, row-major)
  int64_t u_off;  // U (rows x rows)
  int64_t s_off;  // S (min(rows, cols)), doubles
  int64_t v_off;  // V (cols x cols)
};

// A = U diag(S) V per block; S descending.  `need_u[b]` / `need_v[b]` give how
// many leading columns of U / rows of V must be orthonormal to machine
// precision (columns beyond the numerical rank are completed); pass rows /
// cols for full matrices.  Reports sweeps used.
// Jacobi sweeps and wall seconds of the last block_svd call on this thread.
int block_svd_last_sweeps();
double block_svd_last_seconds();
int block_svd(int dtype, const void* A, void* U, double* S, void* V, const std::vector<SvdBlock>& blocks,
              const std::vector<int>& need_u, const std::vector<int>& need_v, hipStream_t s);

}  // namespace dmrgx
