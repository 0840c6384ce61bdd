// FP64 Krylov solvers for the trace system.  See krylov.h.
#include <cmath>
#include <vector>

#include "bsr_row.h"
#include "krylov.h"

namespace hfx {

namespace {

constexpr int RT = 256;  // reduction block size

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  return v;
}

// block-wide sum, result valid in thread 0
__device__ inline double block_sum(double v, double* red) {
  v = wave_sum(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) red[wid] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
  __syncthreads();
  return s;
}

template <int BS>
__global__ void spmv_kernel(int nb, const int* __restrict__ rowptr, const int* __restrict__ col,
                            const double* __restrict__ vals, const double* __restrict__ x,
                            const double* __restrict__ b, double* __restrict__ y) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nb * BS) return;
  const int F = (int)(t / BS), r = (int)(t % BS);
  const double acc = bsr_row_sum<BS>(rowptr[F], rowptr[F + 1], r, nullptr, col, vals, x);
  y[t] = b ? b[t] - acc : acc;
}

// thread-per-block Gauss-Jordan inverse of the diagonal blocks
template <int BS>
__global__ void invert_diag_kernel(int nb, const int* __restrict__ diag, const double* __restrict__ vals,
                                   double* __restrict__ dinv) {
  const int F = blockIdx.x * blockDim.x + threadIdx.x;
  if (F >= nb) return;
  double A[BS * BS], I[BS * BS];
  const double* src = vals + (size_t)diag[F] * BS * BS;
  for (int i = 0; i < BS * BS; ++i) { A[i] = src[i]; I[i] = 0.0; }
  for (int i = 0; i < BS; ++i) I[i * BS + i] = 1.0;
  for (int k = 0; k < BS; ++k) {
    int p = k;
    for (int r = k + 1; r < BS; ++r)
      if (fabs(A[r * BS + k]) > fabs(A[p * BS + k])) p = r;
    if (p != k)
      for (int c = 0; c < BS; ++c) {
        double t = A[k * BS + c]; A[k * BS + c] = A[p * BS + c]; A[p * BS + c] = t;
        t = I[k * BS + c]; I[k * BS + c] = I[p * BS + c]; I[p * BS + c] = t;
      }
    const double inv = 1.0 / A[k * BS + k];
    for (int c = 0; c < BS; ++c) { A[k * BS + c] *= inv; I[k * BS + c] *= inv; }
    for (int r = 0; r < BS; ++r) {
      if (r == k) continue;
      const double f = A[r * BS + k];
      if (f == 0.0) continue;
      for (int c = 0; c < BS; ++c) { A[r * BS + c] -= f * A[k * BS + c]; I[r * BS + c] -= f * I[k * BS + c]; }
    }
  }
  double* dst = dinv + (size_t)F * BS * BS;
  for (int i = 0; i < BS * BS; ++i) dst[i] = I[i];
}

__global__ void point_diag_kernel(int nb, int bs, const int* __restrict__ diag, const double* __restrict__ vals,
                                  double* __restrict__ dinv) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nb * bs) return;
  const int F = (int)(t / bs), r = (int)(t % bs);
  const double d = vals[((size_t)diag[F] * bs + r) * bs + r];
  dinv[t] = d != 0.0 ? 1.0 / d : 1.0;
}

template <int BS>
__global__ void block_diag_apply_kernel(int nb, const double* __restrict__ dinv, const double* __restrict__ x,
                                        double* __restrict__ y) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)nb * BS) return;
  const int F = (int)(t / BS), r = (int)(t % BS);
  y[t] = bsr_row_dot<BS>(dinv + ((size_t)F * BS + r) * BS, x + (size_t)F * BS);
}

__global__ void point_diag_apply_kernel(long long n, const double* __restrict__ d, const double* __restrict__ x,
                                        double* __restrict__ y) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = d[t] * x[t];
}

// partial[j*G + blk] = sum over this block's slice of V_j . w
__global__ void multidot_kernel(long long n, const double* __restrict__ V, long long ldv, const double* __restrict__ w,
                                double* __restrict__ partial) {
  __shared__ double red[RT / 64];
  const int j = blockIdx.y, G = gridDim.x;
  const double* v = V + (size_t)j * ldv;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * RT + threadIdx.x; i < n; i += (long long)G * RT) acc += v[i] * w[i];
  double s = block_sum(acc, red);
  if (threadIdx.x == 0) partial[(size_t)j * G + blockIdx.x] = s;
}

// out[j] = sum_b partial[j*G + b], fixed order
__global__ void reduce_rows_kernel(int G, const double* __restrict__ partial, double* __restrict__ out) {
  __shared__ double red[RT / 64];
  const int j = blockIdx.x;
  double acc = 0.0;
  for (int b = threadIdx.x; b < G; b += RT) acc += partial[(size_t)j * G + b];
  double s = block_sum(acc, red);
  if (threadIdx.x == 0) out[j] = s;
}

__global__ void update_norm_kernel(long long n, int k, const double* __restrict__ V, long long ldv,
                                   const double* __restrict__ h, double* __restrict__ w, double* __restrict__ partial) {
  __shared__ double red[RT / 64];
  __shared__ double hs[64];
  if (threadIdx.x < k) hs[threadIdx.x] = h[threadIdx.x];
  __syncthreads();
  const int G = gridDim.x;
  double acc = 0.0;
  for (long long i = (long long)blockIdx.x * RT + threadIdx.x; i < n; i += (long long)G * RT) {
    double wi = w[i];
    for (int j = 0; j < k; ++j) wi -= hs[j] * V[(size_t)j * ldv + i];
    w[i] = wi;
    acc += wi * wi;
  }
  double s = block_sum(acc, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void scale_by_inv_sqrt_kernel(long long n, const double* __restrict__ norm2, const double* __restrict__ w,
                                         double* __restrict__ v) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const double nv = sqrt(*norm2);
  if (t < n) v[t] = nv > 0.0 ? w[t] / nv : 0.0;
}

__global__ void scale_kernel(long long n, double a, const double* __restrict__ x, double* __restrict__ y) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) y[t] = a * x[t];
}

// x += sum_j y_j Z_j
__global__ void combine_kernel(long long n, int k, const double* __restrict__ Z, long long ldz,
                               const double* __restrict__ y, double* __restrict__ x) {
  __shared__ double ys[64];
  if (threadIdx.x < k) ys[threadIdx.x] = y[threadIdx.x];
  __syncthreads();
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  double acc = x[t];
  for (int j = 0; j < k; ++j) acc += ys[j] * Z[(size_t)j * ldz + t];
  x[t] = acc;
}


// ---------------------------------------------------------------- BiCGStab
//
// Scalars live on the device (no host round trip per iteration):
//   sc[0] rho = (rhat, r)     sc[1] alpha     sc[2] omega     sc[3] rho_prev
//   sc[4] (rhat, v)           sc[5] (t, s)    sc[6] (t, t)    sc[7] (r, r)
//   sc[8] breakdown flag (1 when a denominator vanished)
enum { SC_RHO = 0, SC_ALPHA, SC_OMEGA, SC_RHOP, SC_RV, SC_TS, SC_TT, SC_RR, SC_BRK, SC_N };

// partial[j * G + blk] = sum over the block's slice of a_j . b_j, j < nd
__global__ void dots_kernel(long long n, int nd, const double* __restrict__ a0, const double* __restrict__ b0,
                            const double* __restrict__ a1, const double* __restrict__ b1,
                            double* __restrict__ partial) {
  __shared__ double red[RT / 64];
  const int G = gridDim.x;
  double s0 = 0.0, s1 = 0.0;
  for (long long i = (long long)blockIdx.x * RT + threadIdx.x; i < n; i += (long long)G * RT) {
    s0 += a0[i] * b0[i];
    if (nd > 1) s1 += a1[i] * b1[i];
  }
  double t0 = block_sum(s0, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = t0;
  if (nd > 1) {
    __syncthreads();
    double t1 = block_sum(s1, red);
    if (threadIdx.x == 0) partial[G + blockIdx.x] = t1;
  }
}

// Reduce nd partial rows (fixed order) and update the scalars:
//   op 0: rho_prev = rho; rho = d0; rr = d1          (after the r update)
//   op 1: rv = d0; alpha = rho / rv                   (after v = A p^)
//   op 2: ts = d0; tt = d1; omega = ts / tt           (after t = A s^)
// One workgroup of ST threads; the partials are read ST at a time with
// several loads in flight per thread.
constexpr int ST = 1024;
__global__ __launch_bounds__(ST) void scalars_kernel(int G, int nd, int op, const double* __restrict__ partial,
                                                     double* __restrict__ sc) {
  __shared__ double red[ST / 64];
  double d[2] = {0.0, 0.0};
  for (int j = 0; j < nd; ++j) {
    double acc[4] = {0.0, 0.0, 0.0, 0.0};
    const double* pj = partial + (size_t)j * G;
    int b = threadIdx.x;
    for (; b + 3 * ST < G; b += 4 * ST) {
#pragma unroll
      for (int u = 0; u < 4; ++u) acc[u] += pj[b + u * ST];
    }
    for (; b < G; b += ST) acc[0] += pj[b];
    d[j] = block_sum((acc[0] + acc[1]) + (acc[2] + acc[3]), red);
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  if (op == 0) {
    sc[SC_RHOP] = sc[SC_RHO];
    sc[SC_RHO] = d[0];
    sc[SC_RR] = d[1];
    if (d[0] == 0.0) sc[SC_BRK] = 1.0;
  } else if (op == 1) {
    sc[SC_RV] = d[0];
    if (d[0] == 0.0 || !isfinite(d[0])) { sc[SC_BRK] = 1.0; sc[SC_ALPHA] = 0.0; }
    else sc[SC_ALPHA] = sc[SC_RHO] / d[0];
  } else {
    sc[SC_TS] = d[0];
    sc[SC_TT] = d[1];
    if (d[1] == 0.0 || !isfinite(d[1])) { sc[SC_BRK] = 1.0; sc[SC_OMEGA] = 0.0; }
    else sc[SC_OMEGA] = d[0] / d[1];
    if (sc[SC_OMEGA] == 0.0) sc[SC_BRK] = 1.0;
  }
}

// p = r + beta (p - omega v), beta = (rho / rho_prev) (alpha / omega);
// first = 1: p = r
__global__ void bicg_p_kernel(long long n, int first, const double* __restrict__ sc, const double* __restrict__ r,
                              double* __restrict__ p, const double* __restrict__ v) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  if (first) { p[t] = r[t]; return; }
  const double om = sc[SC_OMEGA], rp = sc[SC_RHOP];
  const double beta = (rp != 0.0 && om != 0.0) ? (sc[SC_RHO] / rp) * (sc[SC_ALPHA] / om) : 0.0;
  p[t] = r[t] + beta * (p[t] - om * v[t]);
}

// s = r - alpha v
__global__ void bicg_s_kernel(long long n, const double* __restrict__ sc, const double* __restrict__ r,
                              const double* __restrict__ v, double* __restrict__ s) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) s[t] = r[t] - sc[SC_ALPHA] * v[t];
}

// x += alpha p^ + omega s^ ; r = s - omega t
__global__ void bicg_xr_kernel(long long n, const double* __restrict__ sc, const double* __restrict__ ph,
                               const double* __restrict__ sh, const double* __restrict__ s,
                               const double* __restrict__ tv, double* __restrict__ x, double* __restrict__ r) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const double al = sc[SC_ALPHA], om = sc[SC_OMEGA];
  x[t] += al * ph[t] + om * sh[t];
  r[t] = s[t] - om * tv[t];
}

// y = A x (thread per row) with the block partials of w1 . y (and y . y when
// `both`) in partial[blk] (and partial[gridDim.x + blk]): the SpMV of a
// BiCGStab step and its dot products in one pass.
template <int BS>
__global__ __launch_bounds__(RT) void spmv_dot_kernel(int nb, const int* __restrict__ rowptr,
                                                      const int* __restrict__ col, const double* __restrict__ vals,
                                                      const double* __restrict__ x, double* __restrict__ y,
                                                      const double* __restrict__ w1, int both,
                                                      double* __restrict__ partial) {
  __shared__ double red[RT / 64];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double a1 = 0.0, a2 = 0.0;
  if (t < (long long)nb * BS) {
    const int F = (int)(t / BS), r = (int)(t % BS);
    const double acc = bsr_row_sum<BS>(rowptr[F], rowptr[F + 1], r, nullptr, col, vals, x);
    y[t] = acc;
    a1 = w1[t] * acc;
    a2 = acc * acc;
  }
  const double s1 = block_sum(a1, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = s1;
  if (both) {
    __syncthreads();
    const double s2 = block_sum(a2, red);
    if (threadIdx.x == 0) partial[gridDim.x + blockIdx.x] = s2;
  }
}

// x += alpha p^ + omega s^ ; r = s - omega t, with the block partials of
// rhat . r and r . r
__global__ __launch_bounds__(RT) void bicg_xr_dot_kernel(long long n, const double* __restrict__ sc,
                                                         const double* __restrict__ ph, const double* __restrict__ sh,
                                                         const double* __restrict__ s, const double* __restrict__ tv,
                                                         const double* __restrict__ rh, double* __restrict__ x,
                                                         double* __restrict__ r, double* __restrict__ partial) {
  __shared__ double red[RT / 64];
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  double a1 = 0.0, a2 = 0.0;
  if (t < n) {
    const double al = sc[SC_ALPHA], om = sc[SC_OMEGA];
    x[t] += al * ph[t] + om * sh[t];
    const double rv = s[t] - om * tv[t];
    r[t] = rv;
    a1 = rh[t] * rv;
    a2 = rv * rv;
  }
  const double s1 = block_sum(a1, red);
  if (threadIdx.x == 0) partial[blockIdx.x] = s1;
  __syncthreads();
  const double s2 = block_sum(a2, red);
  if (threadIdx.x == 0) partial[gridDim.x + blockIdx.x] = s2;
}

__global__ void set_scalars_kernel(double* __restrict__ sc) {
  if (threadIdx.x == 0) {
    for (int i = 0; i < SC_N; ++i) sc[i] = 0.0;
    sc[SC_ALPHA] = 1.0;
    sc[SC_OMEGA] = 1.0;
    sc[SC_RHO] = 1.0;
  }
}

}  // namespace

void bsr_spmv(const BsrDev& A, const double* x, double* y, const double* b, hipStream_t s) {
  const long long n = A.nrows();
  if (n == 0) return;
#define HFX_BS(B)                                                                                         \
  case B:                                                                                                 \
    hipLaunchKernelGGL((spmv_kernel<B>), dim3(ceil_div(n, 256)), dim3(256), 0, s, A.nb, A.rowptr.p, A.col.p, \
                       A.vals.p, x, b, y);                                                                \
    break;
  switch (A.bs) {
    HFX_BS(1) HFX_BS(2) HFX_BS(3) HFX_BS(4) HFX_BS(5) HFX_BS(6) HFX_BS(7) HFX_BS(8) HFX_BS(9) HFX_BS(10) HFX_BS(11) HFX_BS(12) HFX_BS(15) HFX_BS(18) HFX_BS(21) HFX_BS(30)
    default: throw std::runtime_error("spmv: unsupported block size");
  }
#undef HFX_BS
  HFX_HIP(hipGetLastError());
}

void bsr_invert_diag(const BsrDev& A, double* dinv, hipStream_t s) {
  if (A.nb == 0) return;
#define HFX_BS(B)                                                                                              \
  case B:                                                                                                      \
    hipLaunchKernelGGL((invert_diag_kernel<B>), dim3(ceil_div(A.nb, 64)), dim3(64), 0, s, A.nb, A.diag.p, A.vals.p, \
                       dinv);                                                                                  \
    break;
  switch (A.bs) {
    HFX_BS(1) HFX_BS(2) HFX_BS(3) HFX_BS(4) HFX_BS(5) HFX_BS(6) HFX_BS(7) HFX_BS(8) HFX_BS(9) HFX_BS(10) HFX_BS(11) HFX_BS(12) HFX_BS(15) HFX_BS(18) HFX_BS(21) HFX_BS(30)
    default: throw std::runtime_error("block jacobi: unsupported block size");
  }
#undef HFX_BS
  HFX_HIP(hipGetLastError());
}

void bsr_inv_point_diag(const BsrDev& A, double* dinv, hipStream_t s) {
  const long long n = A.nrows();
  if (n == 0) return;
  hipLaunchKernelGGL(point_diag_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, A.nb, A.bs, A.diag.p, A.vals.p, dinv);
  HFX_HIP(hipGetLastError());
}

void apply_block_diag(int nb, int bs, const double* dinv, const double* x, double* y, hipStream_t s) {
  const long long n = (long long)nb * bs;
  if (n == 0) return;
#define HFX_BS(B)                                                                                              \
  case B:                                                                                                      \
    hipLaunchKernelGGL((block_diag_apply_kernel<B>), dim3(ceil_div(n, 256)), dim3(256), 0, s, nb, dinv, x, y); \
    break;
  switch (bs) {
    HFX_BS(1) HFX_BS(2) HFX_BS(3) HFX_BS(4) HFX_BS(5) HFX_BS(6) HFX_BS(7) HFX_BS(8) HFX_BS(9) HFX_BS(10) HFX_BS(11) HFX_BS(12) HFX_BS(15) HFX_BS(18) HFX_BS(21) HFX_BS(30)
    default: throw std::runtime_error("block jacobi: unsupported block size");
  }
#undef HFX_BS
  HFX_HIP(hipGetLastError());
}

void apply_point_diag(long long n, const double* dinv, const double* x, double* y, hipStream_t s) {
  if (n == 0) return;
  hipLaunchKernelGGL(point_diag_apply_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, dinv, x, y);
  HFX_HIP(hipGetLastError());
}

// round a count of doubles up to 256 bytes (caller workspaces are carved in
// 256-byte aligned pieces)
static inline size_t al32(size_t k) { return (k + 31) & ~(size_t)31; }

int Reducer::grid_for(long long n) {
  long long g = (n + RT * 8 - 1) / (RT * 8);
  if (g < 1) g = 1;
  if (g > 1024) g = 1024;
  return (int)g;
}

void Reducer::init(long long n) {
  grid = grid_for(n);
  partial_buf.alloc((size_t)grid * 66);
  out_buf.alloc(132);
  partial = partial_buf.p;
  out = out_buf.p;
}

void Reducer::init(long long n, double* mem) {
  grid = grid_for(n);
  partial = mem;
  out = mem + (size_t)grid * 66;
}

void Reducer::multidot(long long n, int k, const double* V, long long ldv, const double* w, hipStream_t s) {
  if (k > 64) throw std::runtime_error("multidot: at most 64 vectors");
  hipLaunchKernelGGL(multidot_kernel, dim3(grid, k), dim3(RT), 0, s, n, V, ldv, w, partial);
  HFX_HIP(hipGetLastError());
  hipLaunchKernelGGL(reduce_rows_kernel, dim3(k), dim3(RT), 0, s, grid, partial, out);
  HFX_HIP(hipGetLastError());
}

void Reducer::update_norm(long long n, int k, const double* V, long long ldv, const double* h, double* w,
                          double* out_norm2, hipStream_t s) {
  if (k > 64) throw std::runtime_error("update: at most 64 vectors");
  hipLaunchKernelGGL(update_norm_kernel, dim3(grid), dim3(RT), 0, s, n, k, V, ldv, h, w, partial);
  HFX_HIP(hipGetLastError());
  hipLaunchKernelGGL(reduce_rows_kernel, dim3(1), dim3(RT), 0, s, grid, partial, out_norm2);
  HFX_HIP(hipGetLastError());
}

double Reducer::norm(long long n, const double* x, hipStream_t s) {
  multidot(n, 1, x, 0, x, s);
  double v = 0.0;
  HFX_HIP(hipMemcpyAsync(&v, out, sizeof(double), hipMemcpyDeviceToHost, s));
  HFX_HIP(hipStreamSynchronize(s));
  return std::sqrt(v);
}

void fgmres(const BsrDev& A, const double* b, double* x, const std::function<void(const double*, double*)>& precond,
            const KrylovOpts& o, KrylovInfo& info, hipStream_t s) {
  const long long n = A.nrows();
  const int m = o.restart;
  if (m < 1 || m > 63) throw std::runtime_error("gmres: restart must be in [1, 63]");
  info = KrylovInfo();
  Reducer red;
  red.init(n);
  DevBuf<double> V, Zb, w, r, hdev;
  V.alloc((size_t)(m + 1) * n);
  Zb.alloc((size_t)m * n);
  w.alloc(n);
  r.alloc(n);
  hdev.alloc(2 * 64 + 4);
  const double bnorm = red.norm(n, b, s);
  if (bnorm == 0.0) {
    HFX_HIP(hipMemsetAsync(x, 0, n * sizeof(double), s));
    HFX_HIP(hipStreamSynchronize(s));
    info.reason = CONVERGED_ATOL;
    return;
  }
  const double target = o.rtol * bnorm;
  std::vector<double> H((size_t)(m + 1) * m), cs(m), sn(m), g(m + 1), y(m), hh(2 * 64 + 2);
  int it = 0;
  while (true) {
    bsr_spmv(A, x, r.p, b, s);  // r = b - A x (true FP64 residual)
    const double beta = red.norm(n, r.p, s);
    info.true_relres = beta / bnorm;
    if (!std::isfinite(beta)) { info.reason = DIVERGED_NANORINF; break; }
    if (beta <= target) { info.reason = CONVERGED_RTOL; break; }
    if (it >= o.maxits) { info.reason = DIVERGED_ITS; break; }
    hipLaunchKernelGGL(scale_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, 1.0 / beta, r.p, V.p);
    HFX_HIP(hipGetLastError());
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int k = 0;
    bool done = false;
    for (; k < m && it < o.maxits; ++k) {
      double* vk = V.p + (size_t)k * n;
      double* zk = Zb.p + (size_t)k * n;
      if (precond) precond(vk, zk);
      else HFX_HIP(hipMemcpyAsync(zk, vk, n * sizeof(double), hipMemcpyDeviceToDevice, s));
      bsr_spmv(A, zk, w.p, nullptr, s);
      // classical Gram-Schmidt, twice
      red.multidot(n, k + 1, V.p, n, w.p, s);
      HFX_HIP(hipMemcpyAsync(hdev.p, red.out, (k + 1) * sizeof(double), hipMemcpyDeviceToDevice, s));
      red.update_norm(n, k + 1, V.p, n, hdev.p, w.p, hdev.p + 128, s);
      red.multidot(n, k + 1, V.p, n, w.p, s);
      HFX_HIP(hipMemcpyAsync(hdev.p + 64, red.out, (k + 1) * sizeof(double), hipMemcpyDeviceToDevice, s));
      red.update_norm(n, k + 1, V.p, n, hdev.p + 64, w.p, hdev.p + 128, s);
      hipLaunchKernelGGL(scale_by_inv_sqrt_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, hdev.p + 128, w.p,
                         V.p + (size_t)(k + 1) * n);
      HFX_HIP(hipGetLastError());
      HFX_HIP(hipMemcpyAsync(hh.data(), hdev.p, (2 * 64 + 1) * sizeof(double), hipMemcpyDeviceToHost, s));
      HFX_HIP(hipStreamSynchronize(s));
      ++it;
      for (int j = 0; j <= k; ++j) H[(size_t)j * m + k] = hh[j] + hh[64 + j];
      const double hn = std::sqrt(hh[128]);
      H[(size_t)(k + 1) * m + k] = hn;
      for (int j = 0; j < k; ++j) {
        const double a0 = H[(size_t)j * m + k], a1 = H[(size_t)(j + 1) * m + k];
        H[(size_t)j * m + k] = cs[j] * a0 + sn[j] * a1;
        H[(size_t)(j + 1) * m + k] = -sn[j] * a0 + cs[j] * a1;
      }
      const double a0 = H[(size_t)k * m + k], a1 = H[(size_t)(k + 1) * m + k];
      const double rr = std::hypot(a0, a1);
      cs[k] = rr != 0.0 ? a0 / rr : 1.0;
      sn[k] = rr != 0.0 ? a1 / rr : 0.0;
      H[(size_t)k * m + k] = rr;
      H[(size_t)(k + 1) * m + k] = 0.0;
      g[k + 1] = -sn[k] * g[k];
      g[k] = cs[k] * g[k];
      if (!std::isfinite(g[k + 1])) { done = true; ++k; break; }
      if (std::fabs(g[k + 1]) <= target || hn == 0.0) { done = true; ++k; break; }
    }
    // solve the k x k triangular system and update x += Z y
    for (int i = k - 1; i >= 0; --i) {
      double acc = g[i];
      for (int j = i + 1; j < k; ++j) acc -= H[(size_t)i * m + j] * y[j];
      y[i] = H[(size_t)i * m + i] != 0.0 ? acc / H[(size_t)i * m + i] : 0.0;
    }
    if (k > 0) {
      HFX_HIP(hipMemcpyAsync(hdev.p, y.data(), k * sizeof(double), hipMemcpyHostToDevice, s));
      hipLaunchKernelGGL(combine_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, s, n, k, Zb.p, n, hdev.p, x);
      HFX_HIP(hipGetLastError());
    }
    ++info.restarts;
    (void)done;
  }
  info.iters = it;
  HFX_HIP(hipStreamSynchronize(s));
}

}  // namespace hfx

namespace hfx {

size_t bicgstab_ws_doubles(long long n) {
  const size_t G = (size_t)Reducer::grid_for(n), nblk_rows = (size_t)ceil_div(n, RT);
  return al32(Reducer::doubles(n)) + 8 * al32((size_t)n) + al32(SC_N) + al32(2 * G) + al32(2 * nblk_rows);
}

void bicgstab(const BsrDev& A, const double* b, double* x, const std::function<void(const double*, double*)>& precond,
              const KrylovOpts& o, KrylovInfo& info, hipStream_t s, double* ws) {
  const long long n = A.nrows();
  info = KrylovInfo();
  if (ws && o.graph) throw std::runtime_error("bicgstab: a caller workspace excludes the graph option");
  Reducer red;
  const int G = Reducer::grid_for(n);
  const int nblk_rows = ceil_div(n, RT);
  // r, rhat, p, v, p^, s, s^, t, the device scalars, the dot partials and the
  // fused-kernel partials: owned buffers, or carved from the caller's memory
  DevBuf<double> own[11];
  double* vec[8];
  double *sc, *part, *fpart;
  if (ws) {
    red.init(n, ws);
    double* q = ws + al32(Reducer::doubles(n));
    for (int k = 0; k < 8; ++k, q += al32((size_t)n)) vec[k] = q;
    sc = q;
    q += al32(SC_N);
    part = q;
    q += al32((size_t)2 * G);
    fpart = q;
  } else {
    red.init(n);
    for (int k = 0; k < 8; ++k) {
      own[k].alloc(n);
      vec[k] = own[k].p;
    }
    own[8].alloc(SC_N);
    own[9].alloc((size_t)2 * G);
    own[10].alloc((size_t)2 * nblk_rows);
    sc = own[8].p;
    part = own[9].p;
    fpart = own[10].p;
  }
  double *r = vec[0], *rh = vec[1], *p = vec[2], *v = vec[3], *ph = vec[4], *sv = vec[5], *sh = vec[6], *tv = vec[7];
  const double bnorm = red.norm(n, b, s);
  if (bnorm == 0.0) {
    HFX_HIP(hipMemsetAsync(x, 0, n * sizeof(double), s));
    HFX_HIP(hipStreamSynchronize(s));
    info.reason = CONVERGED_ATOL;
    return;
  }
  const double target = o.rtol * bnorm;
  const int nblk = ceil_div(n, 256);
  auto apply_M = [&](const double* in, double* out) {
    if (precond) precond(in, out);
    else HFX_HIP(hipMemcpyAsync(out, in, n * sizeof(double), hipMemcpyDeviceToDevice, s));
  };
  auto dots = [&](int nd, const double* a0, const double* b0, const double* a1, const double* b1, int op) {
    hipLaunchKernelGGL(dots_kernel, dim3(G), dim3(RT), 0, s, n, nd, a0, b0, a1, b1, part);
    HFX_HIP(hipGetLastError());
    hipLaunchKernelGGL(scalars_kernel, dim3(1), dim3(ST), 0, s, G, nd, op, part, sc);
    HFX_HIP(hipGetLastError());
  };
  // fused path: SpMV / update kernels leave one partial per 256-row block
  auto spmv_dot = [&](const double* xin, double* yout, const double* w1, bool both) {
#define HFX_BS(B)                                                                                          \
  case B:                                                                                                  \
    hipLaunchKernelGGL((spmv_dot_kernel<B>), dim3(nblk_rows), dim3(RT), 0, s, A.nb, A.rowptr.p, A.col.p,   \
                       A.vals.p, xin, yout, w1, both ? 1 : 0, fpart);                                    \
    break;
    switch (A.bs) {
      HFX_BS(1) HFX_BS(2) HFX_BS(3) HFX_BS(4) HFX_BS(5) HFX_BS(6) HFX_BS(7) HFX_BS(8) HFX_BS(9) HFX_BS(10) HFX_BS(11) HFX_BS(12) HFX_BS(15) HFX_BS(18) HFX_BS(21) HFX_BS(30)
      default: throw std::runtime_error("bicgstab: unsupported block size");
    }
#undef HFX_BS
    HFX_HIP(hipGetLastError());
  };
  auto scalars = [&](int nd, int op) {
    hipLaunchKernelGGL(scalars_kernel, dim3(1), dim3(ST), 0, s, nblk_rows, nd, op, fpart, sc);
    HFX_HIP(hipGetLastError());
  };
  const int check_every = 8;
  // one steady iteration (after the first of a restart): p, M p, A p^ with
  // (rhat, v), alpha, s, M s, A s^ with (s, t) and (t, t), omega, x, r with
  // (rhat, r) and (r, r)
  auto steady = [&]() {
    hipLaunchKernelGGL(bicg_p_kernel, dim3(nblk), dim3(256), 0, s, n, 0, sc, r, p, v);
    HFX_HIP(hipGetLastError());
    apply_M(p, ph);
    spmv_dot(ph, v, rh, false);
    scalars(1, 1);
    hipLaunchKernelGGL(bicg_s_kernel, dim3(nblk), dim3(256), 0, s, n, sc, r, v, sv);
    HFX_HIP(hipGetLastError());
    apply_M(sv, sh);
    spmv_dot(sh, tv, sv, true);
    scalars(2, 2);
    hipLaunchKernelGGL(bicg_xr_dot_kernel, dim3(nblk), dim3(RT), 0, s, n, sc, ph, sh, sv, tv, rh, x, r,
                       fpart);
    HFX_HIP(hipGetLastError());
    scalars(2, 0);
  };
  // check_every steady iterations as one graph, captured at first use; it
  // holds kernels only (no memset or copy nodes), and every buffer it names
  // lives until the end of this call
  struct GraphHold {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    ~GraphHold() {
      if (e) (void)hipGraphExecDestroy(e);
      if (g) (void)hipGraphDestroy(g);
    }
  } gh;
  const bool use_graph = o.graph && static_cast<bool>(precond);
  auto graph_block = [&]() {
    if (!gh.e) {
      HFX_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
      try {
        for (int k = 0; k < check_every; ++k) steady();
      } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(s, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
      }
      HFX_HIP(hipStreamEndCapture(s, &gh.g));
      HFX_HIP(hipGraphInstantiate(&gh.e, gh.g, nullptr, nullptr, 0));
    }
    HFX_HIP(hipGraphLaunch(gh.e, s));
  };
  int it = 0;
  bool restart = true;
  while (true) {
    if (!restart && use_graph && it % check_every == 0 && it + check_every <= o.maxits) {
      graph_block();
      it += check_every;
      double hs[SC_N];
      HFX_HIP(hipMemcpyAsync(hs, sc, sizeof(hs), hipMemcpyDeviceToHost, s));
      HFX_HIP(hipStreamSynchronize(s));
      const double rr = std::sqrt(std::fabs(hs[SC_RR]));
      if (!std::isfinite(hs[SC_RR])) { info.reason = DIVERGED_NANORINF; break; }
      if (rr <= target || hs[SC_BRK] != 0.0 || it >= o.maxits) restart = true;
      continue;
    }
    if (restart) {
      // r = b - A x (true FP64 residual), rhat = r, rho = (r, r)
      bsr_spmv(A, x, r, b, s);
      const double rn = red.norm(n, r, s);
      info.true_relres = rn / bnorm;
      if (!std::isfinite(rn)) { info.reason = DIVERGED_NANORINF; break; }
      if (rn <= target) { info.reason = CONVERGED_RTOL; break; }
      if (it >= o.maxits) { info.reason = DIVERGED_ITS; break; }
      HFX_HIP(hipMemcpyAsync(rh, r, n * sizeof(double), hipMemcpyDeviceToDevice, s));
      hipLaunchKernelGGL(set_scalars_kernel, dim3(1), dim3(64), 0, s, sc);
      HFX_HIP(hipGetLastError());
      dots(2, rh, r, r, r, 0);
      ++info.restarts;
      restart = false;
      hipLaunchKernelGGL(bicg_p_kernel, dim3(nblk), dim3(256), 0, s, n, 1, sc, r, p, v);
      HFX_HIP(hipGetLastError());
      apply_M(p, ph);
      spmv_dot(ph, v, rh, false);          // v = A p^, (rhat, v)
      scalars(1, 1);
      hipLaunchKernelGGL(bicg_s_kernel, dim3(nblk), dim3(256), 0, s, n, sc, r, v, sv);
      HFX_HIP(hipGetLastError());
      apply_M(sv, sh);
      spmv_dot(sh, tv, sv, true);          // t = A s^, (s, t), (t, t)
      scalars(2, 2);
      hipLaunchKernelGGL(bicg_xr_dot_kernel, dim3(nblk), dim3(RT), 0, s, n, sc, ph, sh, sv, tv, rh, x,
                         r, fpart);
      HFX_HIP(hipGetLastError());
      scalars(2, 0);                             // (rhat, r), (r, r)
    } else {
      steady();
    }
    ++it;
    if (it % check_every == 0 || it >= o.maxits) {
      double hs[SC_N];
      HFX_HIP(hipMemcpyAsync(hs, sc, sizeof(hs), hipMemcpyDeviceToHost, s));
      HFX_HIP(hipStreamSynchronize(s));
      const double rr = std::sqrt(std::fabs(hs[SC_RR]));
      if (!std::isfinite(hs[SC_RR])) { info.reason = DIVERGED_NANORINF; break; }
      // converged recurrence, breakdown, or the iteration budget: restart from
      // (and test) the true residual; maxits bounds the number of restarts
      if (rr <= target || hs[SC_BRK] != 0.0 || it >= o.maxits) restart = true;
    }
  }
  info.iters = it;
  HFX_HIP(hipStreamSynchronize(s));
}

}  // namespace hfx
