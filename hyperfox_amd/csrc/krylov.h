// FP64 Krylov solvers for the trace system on the GPU: BSR SpMV,
// point/block Jacobi, restarted (F)GMRES with classical Gram-Schmidt plus one
// re-orthogonalisation, convergence judged on the true FP64 residual.
//
// Stands in for PetscInterface::solve (src/resolution/PetscInterface.cpp:222-247)
// with KSPGMRES semantics (restart 30, zero initial guess).
#pragma once

#include <hip/hip_runtime.h>

#include <functional>

#include "hfx_internal.h"
#include "trace_system.h"

namespace hfx {

enum PcType { PC_NONE = 0, PC_JACOBI = 1, PC_BJACOBI = 2, PC_LEARNED = 3 };

// KSP converged reasons (a subset of PETSc's codes).
enum SolveReason { CONVERGED_RTOL = 2, CONVERGED_ATOL = 3, DIVERGED_ITS = -3, DIVERGED_BREAKDOWN = -5, DIVERGED_NANORINF = -9 };

struct KrylovOpts {
  double rtol = 1e-10;   // on the true residual ||b - A x|| / ||b||
  int maxits = 10000;
  int restart = 30;
  // BiCGStab: run the steady iterations between two residual checks as one
  // captured HIP graph (the same kernels in the same order, so the same
  // results), when a preconditioner is given.  Off by default: measured no
  // faster on gfx950 (the launches are already asynchronous), and each solve
  // pays the capture
  bool graph = false;
};

struct KrylovInfo {
  int iters = 0;
  int reason = 0;
  double true_relres = 0.0;
  int restarts = 0;
};

// y = A x  or, with b non-null, y = b - A x.
void bsr_spmv(const BsrDev& A, const double* x, double* y, const double* b, hipStream_t s);

// Inverse diagonal blocks (bs x bs, row-major), batched Gauss-Jordan.
void bsr_invert_diag(const BsrDev& A, double* dinv, hipStream_t s);
// Inverse point diagonal.
void bsr_inv_point_diag(const BsrDev& A, double* dinv, hipStream_t s);
void apply_block_diag(int nb, int bs, const double* dinv, const double* x, double* y, hipStream_t s);
void apply_point_diag(long long n, const double* dinv, const double* x, double* y, hipStream_t s);

// Deterministic FP64 dot / norm helpers (two-stage reductions).  The
// partials live in buffers of its own (init(n)) or in caller memory of
// Reducer::doubles(n) values (init(n, mem)).
struct Reducer {
  DevBuf<double> partial_buf, out_buf;
  double* partial = nullptr;  // grid * 66
  double* out = nullptr;      // 132
  int grid = 0;
  static int grid_for(long long n);
  static size_t doubles(long long n) { return (size_t)grid_for(n) * 66 + 132; }
  void init(long long n);
  void init(long long n, double* mem);
  // out[j] = V_j . w for j < k (V column-major, leading dim ldv); result on device in out.p
  void multidot(long long n, int k, const double* V, long long ldv, const double* w, hipStream_t s);
  // w -= sum_j h[j] V_j ; then out[k] = ||w||^2 partial-reduced (device h, k <= 64)
  void update_norm(long long n, int k, const double* V, long long ldv, const double* h, double* w, double* out_norm2,
                   hipStream_t s);
  double norm(long long n, const double* x, hipStream_t s);  // synchronous
};

// Right-preconditioned flexible GMRES(restart).  `precond(in, out)` applies
// the preconditioner (identity if empty).  x holds the initial guess on entry.
void fgmres(const BsrDev& A, const double* b, double* x, const std::function<void(const double*, double*)>& precond,
            const KrylovOpts& o, KrylovInfo& info, hipStream_t s);

// Right-preconditioned BiCGStab with device-resident scalars (no host round
// trip per iteration; the recurrence residual is read back every 8
// iterations).  Convergence is confirmed on the true FP64 residual; a
// recurrence that converges early or breaks down restarts from it.
//
// ws: null (the solver allocates its vectors), or caller device memory of
// bicgstab_ws_doubles(n) values, 256-byte aligned, which then holds every
// vector, scalar and partial sum of the solve (nothing is allocated; the
// graph option must be off).
void bicgstab(const BsrDev& A, const double* b, double* x, const std::function<void(const double*, double*)>& precond,
              const KrylovOpts& o, KrylovInfo& info, hipStream_t s, double* ws = nullptr);
size_t bicgstab_ws_doubles(long long n);

}  // namespace hfx
