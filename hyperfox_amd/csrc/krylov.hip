// FP64 Krylov solvers for the trace system.  See krylov.h.
#include <cmath>
#include <vector>

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
  double acc = 0.0;
  for (int k = rowptr[F]; k < rowptr[F + 1]; ++k) {
    const double* a = vals + ((size_t)k * BS + r) * BS;
    const double* xv = x + (size_t)col[k] * BS;
#pragma unroll
    for (int c = 0; c < BS; ++c) acc += a[c] * xv[c];
  }
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
  const double* a = dinv + ((size_t)F * BS + r) * BS;
  const double* xv = x + (size_t)F * BS;
  double acc = 0.0;
#pragma unroll
  for (int c = 0; c < BS; ++c) acc += a[c] * xv[c];
  y[t] = acc;
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
    HFX_BS(1) HFX_BS(2) HFX_BS(3) HFX_BS(4) HFX_BS(5) HFX_BS(6) HFX_BS(10)
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
    HFX_BS(1) HFX_BS(2) HFX_BS(3) HFX_BS(4) HFX_BS(5) HFX_BS(6) HFX_BS(10)
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
    HFX_BS(1) HFX_BS(2) HFX_BS(3) HFX_BS(4) HFX_BS(5) HFX_BS(6) HFX_BS(10)
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

void Reducer::init(long long n) {
  long long g = (n + RT * 8 - 1) / (RT * 8);
  if (g < 1) g = 1;
  if (g > 1024) g = 1024;
  grid = (int)g;
  partial.alloc((size_t)grid * 66);
  out.alloc(132);
}

void Reducer::multidot(long long n, int k, const double* V, long long ldv, const double* w, hipStream_t s) {
  if (k > 64) throw std::runtime_error("multidot: at most 64 vectors");
  hipLaunchKernelGGL(multidot_kernel, dim3(grid, k), dim3(RT), 0, s, n, V, ldv, w, partial.p);
  HFX_HIP(hipGetLastError());
  hipLaunchKernelGGL(reduce_rows_kernel, dim3(k), dim3(RT), 0, s, grid, partial.p, out.p);
  HFX_HIP(hipGetLastError());
}

void Reducer::update_norm(long long n, int k, const double* V, long long ldv, const double* h, double* w,
                          double* out_norm2, hipStream_t s) {
  if (k > 64) throw std::runtime_error("update: at most 64 vectors");
  hipLaunchKernelGGL(update_norm_kernel, dim3(grid), dim3(RT), 0, s, n, k, V, ldv, h, w, partial.p);
  HFX_HIP(hipGetLastError());
  hipLaunchKernelGGL(reduce_rows_kernel, dim3(1), dim3(RT), 0, s, grid, partial.p, out_norm2);
  HFX_HIP(hipGetLastError());
}

double Reducer::norm(long long n, const double* x, hipStream_t s) {
  multidot(n, 1, x, 0, x, s);
  double v = 0.0;
  HFX_HIP(hipMemcpyAsync(&v, out.p, sizeof(double), hipMemcpyDeviceToHost, s));
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
      HFX_HIP(hipMemcpyAsync(hdev.p, red.out.p, (k + 1) * sizeof(double), hipMemcpyDeviceToDevice, s));
      red.update_norm(n, k + 1, V.p, n, hdev.p, w.p, hdev.p + 128, s);
      red.multidot(n, k + 1, V.p, n, w.p, s);
      HFX_HIP(hipMemcpyAsync(hdev.p + 64, red.out.p, (k + 1) * sizeof(double), hipMemcpyDeviceToDevice, s));
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
