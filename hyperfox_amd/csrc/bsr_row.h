// Row r of dense row-major BS x BS blocks times BS-vector segments, for the
// thread-per-row BSR kernels.  Columns are taken in pairs,
// acc += a[c] x[c] + a[c+1] x[c+1], which halves the dependent add chain of a
// plain loop; with one block per loop step the compiler merges the row and
// segment loads into 16-byte loads and pipelines the next block's column
// index (measured on gfx950 for 3-D p=3 blocks: 228 -> 185 us per SpMV).
#pragma once

#include <hip/hip_runtime.h>

namespace hfx {

// The block values may be stored in float (V = float): they are widened and
// the sums kept in double.
template <int BS, typename V = double>
__device__ __forceinline__ double bsr_row_dot(const V* __restrict__ a, const double* __restrict__ x) {
  double s = 0.0;
#pragma unroll
  for (int c = 0; c + 1 < BS; c += 2) s += (double)a[c] * x[c] + (double)a[c + 1] * x[c + 1];
  if constexpr (BS % 2 == 1) s += (double)a[BS - 1] * x[BS - 1];
  return s;
}

// sum over the blocks k in [k0, k1) (stored at pos[k], or at k when pos is
// null) of row r of the block times the x segment of column col[k]
template <int BS, typename V = double>
__device__ __forceinline__ double bsr_row_sum(int k0, int k1, int r, const int* __restrict__ pos,
                                              const int* __restrict__ col, const V* __restrict__ vals,
                                              const double* __restrict__ x) {
  double acc = 0.0;
  for (int k = k0; k < k1; ++k) {
    const V* a = vals + ((size_t)(pos ? pos[k] : k) * BS + r) * BS;
    const double* xv = x + (size_t)col[k] * BS;
#pragma unroll
    for (int c = 0; c + 1 < BS; c += 2) acc += (double)a[c] * xv[c] + (double)a[c + 1] * xv[c + 1];
    if constexpr (BS % 2 == 1) acc += (double)a[BS - 1] * xv[BS - 1];
  }
  return acc;
}

}  // namespace hfx
