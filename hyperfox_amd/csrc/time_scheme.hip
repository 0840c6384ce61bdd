// Explicit Euler and Runge-Kutta support kernels (HBM-bound elementwise
// passes; one thread per value, grid-stride).
#include "time_scheme.h"

#include <algorithm>

#include "hfx_internal.h"

namespace hfx {
namespace {

__global__ __launch_bounds__(256) void ts_local_kernel(int nEl, int nN, int dim, int nFpc, int nNf,
                                                       const int* __restrict__ c2f, const int8_t* __restrict__ perm,
                                                       double c_old, const double* __restrict__ u_old,
                                                       const double* __restrict__ q_old,
                                                       const double* __restrict__ l_old, int nk, RkCoef ck,
                                                       const double* __restrict__ ku, const double* __restrict__ kq,
                                                       const double* __restrict__ kl, size_t nu, size_t nq, size_t nl,
                                                       double* __restrict__ w) {
  const int nQ = nN * dim, nloc = nN + nQ + nFpc * nNf;
  const long long total = (long long)nEl * nloc;
  for (long long id = blockIdx.x * (long long)blockDim.x + threadIdx.x; id < total;
       id += (long long)gridDim.x * blockDim.x) {
    const int e = (int)(id / nloc), t = (int)(id % nloc);
    const double* o;
    const double* k;
    size_t idx, stride;
    if (t < nN) {
      o = u_old, k = ku, idx = (size_t)e * nN + t, stride = nu;
    } else if (t < nN + nQ) {
      o = q_old, k = kq, idx = (size_t)e * nQ + (t - nN), stride = nq;
    } else {
      const int f = (t - nN - nQ) / nNf, b = (t - nN - nQ) % nNf;
      o = l_old, k = kl, stride = nl;
      idx = (size_t)c2f[(size_t)e * nFpc + f] * nNf + perm[((size_t)e * nFpc + f) * nNf + b];
    }
    double v = o ? c_old * o[idx] : 0.0;
    if (k)
      for (int j = 0; j < nk; ++j) v += ck.c[j] * k[(size_t)j * stride + idx];
    w[id] = v;
  }
}

__global__ __launch_bounds__(256) void rk_stage_kernel(size_t n, double* __restrict__ cur,
                                                       const double* __restrict__ old, double* __restrict__ k,
                                                       int stage, RkCoef a, double dt) {
  const double idt = 1.0 / dt;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double ks = idt * (cur[i] - old[i]);
    k[(size_t)stage * n + i] = ks;
    double buf = 0.0;
    for (int j = 0; j < stage; ++j) buf += a.c[j] * k[(size_t)j * n + i];
    buf += a.c[stage] * ks;
    cur[i] = old[i] + dt * buf;
  }
}

__global__ __launch_bounds__(256) void rk_solution_kernel(size_t n, double* __restrict__ cur,
                                                          const double* __restrict__ old,
                                                          const double* __restrict__ k, int nst, RkCoef b,
                                                          double dt) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    double v = old[i];
    for (int j = 0; j < nst; ++j) v += (b.c[j] * dt) * k[(size_t)j * n + i];
    cur[i] = v;
  }
}

int grid_for(long long n) { return (int)std::min<long long>(ceil_div(n, 256), 8192); }

}  // namespace

void launch_ts_local(int nEl, int nN, int dim, int nFpc, int nNf, const int* cell2face, const int8_t* face_perm,
                     double c_old, const double* u_old, const double* q_old, const double* l_old, int nk, RkCoef ck,
                     const double* ku, const double* kq, const double* kl, size_t nu, size_t nq, size_t nl, double* w,
                     hipStream_t s) {
  const long long total = (long long)nEl * (nN * (dim + 1) + nFpc * nNf);
  if (total <= 0) return;
  if (nk < 0 || nk > RK_MAX_STAGES) throw std::runtime_error("launch_ts_local: stage count out of range");
  hipLaunchKernelGGL(ts_local_kernel, dim3(grid_for(total)), dim3(256), 0, s, nEl, nN, dim, nFpc, nNf, cell2face,
                     face_perm, c_old, u_old, q_old, l_old, nk, ck, ku, kq, kl, nu, nq, nl, w);
  HFX_HIP(hipGetLastError());
}

void launch_rk_stage(size_t n, double* cur, const double* old, double* k, int stage, RkCoef a, double dt,
                     hipStream_t s) {
  if (n == 0) return;
  if (stage < 0 || stage >= RK_MAX_STAGES) throw std::runtime_error("launch_rk_stage: stage out of range");
  hipLaunchKernelGGL(rk_stage_kernel, dim3(grid_for((long long)n)), dim3(256), 0, s, n, cur, old, k, stage, a, dt);
  HFX_HIP(hipGetLastError());
}

void launch_rk_solution(size_t n, double* cur, const double* old, const double* k, int nst, RkCoef b, double dt,
                        hipStream_t s) {
  if (n == 0) return;
  if (nst < 1 || nst > RK_MAX_STAGES) throw std::runtime_error("launch_rk_solution: stage count out of range");
  hipLaunchKernelGGL(rk_solution_kernel, dim3(grid_for((long long)n)), dim3(256), 0, s, n, cur, old, k, nst, b, dt);
  HFX_HIP(hipGetLastError());
}

}  // namespace hfx
