// Gauss-Jordan elimination on an LDS matrix by one workgroup, shared by the
// general element kernel and the Burgers cell kernel.
#pragma once

#include <hip/hip_runtime.h>

namespace hfx {

// Gauss-Jordan elimination with partial pivoting, without row swaps, on an
// LDS (or workspace) matrix of NR rows and NC columns (row stride NC,
// NR <= 128).  Step k picks the unused row p of largest |A[p][k]|, the lowest
// row among equals (every wave computes the same pivot with a wave-wide
// argmax over rows lane and lane + 64, so no broadcast barrier is needed) and
// eliminates column k from every other row.  On return row perm[k] holds
// variable k: its right-hand columns (NR..NC-1) are divided by the pivot,
// i.e. X[k][c] = A[perm[k]][c].  One workgroup barrier per step.
template <int NR, int NC, int NT>
__device__ void gauss_jordan(double* A, int* perm) {
  static_assert(NR <= 128, "gauss_jordan: at most 128 rows");
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  unsigned long long used = 0ull, used_hi = 0ull;  // rows 0..63, 64..127
  for (int k = 0; k < NR; ++k) {
    double v = -1.0;
    if (lane < NR && !((used >> lane) & 1ull)) {
      v = fabs(A[lane * NC + k]);
      if (!(v >= 0.0)) v = 0.0;  // NaN: keep the argmax total and uniform
    }
    int idx = lane;
    if constexpr (NR > 64) {
      const int r = lane + 64;
      if (r < NR && !((used_hi >> lane) & 1ull)) {
        double w = fabs(A[r * NC + k]);
        if (!(w >= 0.0)) w = 0.0;
        if (w > v) { v = w; idx = r; }
      }
    }
    for (int off = 32; off > 0; off >>= 1) {
      const double ov = __shfl_xor(v, off, 64);
      const int oi = __shfl_xor(idx, off, 64);
      if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
    }
    const int p = idx;
    if (p < 64) used |= 1ull << p;
    else used_hi |= 1ull << (p - 64);
    if (tid == 0) perm[k] = p;
    const double pinv = 1.0 / A[p * NC + k];
    constexpr int TOT = NR * NC;
    for (int t = tid; t < TOT; t += NT) {
      const int r = t / NC, c = t % NC;
      if (r == p || c <= k) continue;
      A[r * NC + c] -= (A[r * NC + k] * pinv) * A[p * NC + c];
    }
    __syncthreads();
  }
  // normalise the right-hand columns of each pivot row
  constexpr int NRHS = NC - NR;
  for (int t = tid; t < NR * NRHS; t += NT) {
    const int k = t / NRHS, c = NR + t % NRHS;
    const int r = perm[k];
    A[r * NC + c] /= A[r * NC + k];
  }
  __syncthreads();
}

}  // namespace hfx
