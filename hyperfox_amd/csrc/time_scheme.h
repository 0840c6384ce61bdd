// Explicit Euler and Runge-Kutta support of the HDG path: the cell-local
// vectors that the element kernels apply the model's u rows to, and the
// stage bookkeeping of RungeKutta (RungeKutta.cpp:161-253) on device fields.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace hfx {

constexpr int RK_MAX_STAGES = 8;
struct RkCoef {
  double c[RK_MAX_STAGES];
};

// w[e][t] = c_old old_t + sum_{j < nk} ck.c[j] k_j,t over the cell-local
// vector t of [u (nN) | q (nN*dim, index i*dim+k) | lambda (nFpc*nNf, in the
// cell's face-node order)].  Fields are whole-mesh: u nEl*nN, q nEl*nN*dim,
// lambda nFaces*nNf; stage j of ku/kq/kl starts at j*nu / j*nq / j*nl.
// q_old, l_old (and kq, kl) may be null: zero.
void launch_ts_local(int nEl, int nN, int dim, int nFpc, int nNf, const int* cell2face, const int8_t* face_perm,
                     double c_old, const double* u_old, const double* q_old, const double* l_old, int nk, RkCoef ck,
                     const double* ku, const double* kq, const double* kl, size_t nu, size_t nq, size_t nl, double* w,
                     hipStream_t s);

// RungeKutta::computeStage on one field of n values (k: nst x n):
// k_stage = (cur - old) / dt, then cur = old + dt sum_{j <= stage} a.c[j] k_j.
void launch_rk_stage(size_t n, double* cur, const double* old, double* k, int stage, RkCoef a, double dt,
                     hipStream_t s);

// RungeKutta::computeSolution on one field: cur = old + sum_j (b.c[j] dt) k_j.
void launch_rk_solution(size_t n, double* cur, const double* old, const double* k, int nst, RkCoef b, double dt,
                        hipStream_t s);

}  // namespace hfx
