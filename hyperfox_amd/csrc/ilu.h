// Block ILU(0) preconditioner on the GPU: the level-scheduled factorisation
// and triangular sweeps of the algorithm stated in ilu_plan.h (PETSc PCILU,
// zero fill, selected through PetscOpts::preconditionnerType,
// src/resolution/PetscOpts.h:16 and src/resolution/PetscInterface.cpp:70).
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "hfx_internal.h"
#include "ilu_plan.h"
#include "trace_system.h"

namespace hfx {

// Device factors and schedule for one pattern.
struct Ilu {
  int nb = 0, bs = 0, ordering = -1;
  IluPlan plan;
  DevBuf<int> rowptr, fwd_rows, bwd_rows, lptr, lpos, lcol, uptr, upos, ucol, dslot, pptr, pslot, ppos, err;
  DevBuf<signed char> kind;
  DevBuf<double> lu, dinv, tmp;
  // low-precision sweeps: float copies of the factors, made after each
  // factorisation (the sweeps widen them and sum in double)
  bool low = false;
  DevBuf<float> lu32, dinv32;
  // Analyse the pattern (host) and upload the schedule.
  void setup(int nb, int bs, const int* rowptr_host, const int* col_host, int ordering, hipStream_t s);
  // Numeric factorisation of A (same pattern); throws on a singular pivot block.
  void factor(const BsrDev& A, hipStream_t s);
  // y = (L U)^-1 x  (x and y distinct)
  void apply(const double* x, double* y, hipStream_t s);
};

}  // namespace hfx
