// Row of a dense row-major BS x BS block times a BS-vector, accumulated into
// acc in column order (the same rounding as the plain loop).  Even block
// sizes read the row and the vector as 16-byte pairs: a block row starts at a
// multiple of BS doubles, so the pairs are aligned, and one thread's row
// takes BS/2 loads instead of BS.
#pragma once

#include <hip/hip_runtime.h>

namespace hfx {

template <int BS>
__device__ __forceinline__ double bsr_row_acc(double acc, const double* __restrict__ a, const double* __restrict__ x) {
  if constexpr (BS % 2 == 0) {
    const double2* a2 = reinterpret_cast<const double2*>(a);
    const double2* x2 = reinterpret_cast<const double2*>(x);
#pragma unroll
    for (int c = 0; c < BS / 2; ++c) {
      const double2 u = a2[c], v = x2[c];
      acc += u.x * v.x;
      acc += u.y * v.y;
    }
  } else {
#pragma unroll
    for (int c = 0; c < BS; ++c) acc += a[c] * x[c];
  }
  return acc;
}

// acc - row . x, subtracting in column order
template <int BS>
__device__ __forceinline__ double bsr_row_sub(double acc, const double* __restrict__ a, const double* __restrict__ x) {
  if constexpr (BS % 2 == 0) {
    const double2* a2 = reinterpret_cast<const double2*>(a);
    const double2* x2 = reinterpret_cast<const double2*>(x);
#pragma unroll
    for (int c = 0; c < BS / 2; ++c) {
      const double2 u = a2[c], v = x2[c];
      acc -= u.x * v.x;
      acc -= u.y * v.y;
    }
  } else {
#pragma unroll
    for (int c = 0; c < BS; ++c) acc -= a[c] * x[c];
  }
  return acc;
}

}  // namespace hfx
