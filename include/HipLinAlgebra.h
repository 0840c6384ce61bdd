/* HipLinAlgebra: the reference's LinAlgebraInterface implemented over the
 * hfx_la_* C ABI of libhfx_hip.so (declared in hfx_hip.h).
 *
 * Header-only.  It is compiled inside a HyperFox build, next to
 * src/resolution/LinAlgebraInterface.h:23-162 (which it includes), and is a
 * drop-in for PetscInterface (src/resolution/PetscInterface.h) wherever a
 * solver takes a LinAlgebraInterface* (Solver::setLinSystem, src/solver/Solver.h:54).
 * Errors throw the reference's ErrorHandle (src/globals/ErrorHandle.h:18),
 * as PetscInterface does (PetscInterface.cpp:84-185), with class
 * "HipLinAlgebra", the member function's name and the library's message.
 */
#ifndef HFX_HIP_LINALGEBRA_H
#define HFX_HIP_LINALGEBRA_H

#include <string>
#include <vector>

#include "ErrorHandle.h"
#include "LinAlgebraInterface.h"
#include "hfx_hip.h"

namespace hfox {

/* PetscOpts' fields (src/resolution/PetscOpts.h:9-31) plus the GMRES restart. */
struct HipOpts {
  int solverType = HFX_LA_GMRES;
  int preconditionnerType = HFX_PC_JACOBI;
  double rtol = 1e-6;
  int maxits = 1000;
  int restart = 30;
  int device = 0;
};

class HipLinAlgebra : public LinAlgebraInterface {
 public:
  HipLinAlgebra() { check(__func__, hfx_la_create(opts.device, &h)); }
  explicit HipLinAlgebra(HipOpts o) : opts(o) { check(__func__, hfx_la_create(opts.device, &h)); }
  ~HipLinAlgebra() { hfx_la_free(h); }
  HipLinAlgebra(const HipLinAlgebra&) = delete;
  HipLinAlgebra& operator=(const HipLinAlgebra&) = delete;

  void initialize() override {
    check(__func__, hfx_la_initialize(h));
    initialized = 1;
  }
  void configure() override {
    check(__func__, hfx_la_configure(h, opts.solverType, opts.preconditionnerType, opts.rtol, opts.maxits, opts.restart));
    configured = 1;
  }
  void allocate(int ndofs, const std::vector<int>* diagSparsePattern = NULL,
                const std::vector<int>* offSparsePattern = NULL) override {
    check(__func__, hfx_la_allocate(h, ndofs, diagSparsePattern ? diagSparsePattern->data() : NULL,
                          offSparsePattern ? offSparsePattern->data() : NULL));
    nDOFs = ndofs;
    allocated = 1;
  }
  void addValMatrix(int i, int j, const double& val) override { check(__func__, hfx_la_add_vals(h, 1, &i, 1, &j, &val)); }
  void addValsMatrix(std::vector<int>& is, std::vector<int>& js, const double* vals) override {
    check(__func__, hfx_la_add_vals(h, (int)is.size(), is.data(), (int)js.size(), js.data(), vals));
  }
  void addValRHS(int i, const double& val) override { check(__func__, hfx_la_add_rhs(h, 1, &i, &val)); }
  void addValsRHS(std::vector<int>& is, const double* vals) override {
    check(__func__, hfx_la_add_rhs(h, (int)is.size(), is.data(), vals));
  }
  void setValMatrix(int i, int j, const double& val) override { check(__func__, hfx_la_set_vals(h, 1, &i, 1, &j, &val)); }
  void setValsMatrix(std::vector<int>& is, std::vector<int>& js, const double* vals) override {
    check(__func__, hfx_la_set_vals(h, (int)is.size(), is.data(), (int)js.size(), js.data(), vals));
  }
  void setValRHS(int i, const double& val) override { check(__func__, hfx_la_set_rhs(h, 1, &i, &val)); }
  void setValsRHS(std::vector<int>& is, const double* vals) override {
    check(__func__, hfx_la_set_rhs(h, (int)is.size(), is.data(), vals));
  }
  void zeroOutRows(std::vector<int>& is) override { check(__func__, hfx_la_zero_rows(h, (int)is.size(), is.data())); }
  void assemble() override {
    check(__func__, hfx_la_assemble(h));
    assembled = 1;
  }
  void assembleFlush() override {
    check(__func__, hfx_la_assemble_flush(h));
    assembled = 1;
  }
  void solve(std::vector<double>* solution) override {
    int n = 0;
    check(__func__, hfx_la_num_dofs(h, &n));
    solution->resize(n);
    check(__func__, hfx_la_solve(h, solution->data(), n, lastInfo, &lastRelRes));
  }
  void getSolutionOwnership(std::vector<int>* range) override {
    int lo = 0, hi = 0;
    check(__func__, hfx_la_ownership(h, &lo, &hi));
    range->resize(hi - lo);
    for (int i = lo; i < hi; ++i) (*range)[i - lo] = i;
  }
  void clearSystem() override { check(__func__, hfx_la_clear(h)); }
  void destroySystem() override {
    check(__func__, hfx_la_destroy_system(h));
    initialized = configured = allocated = assembled = 0;
  }

  /* iterations and converged reason of the last solve, and its true relative residual */
  int lastIterations() const { return lastInfo[0]; }
  int lastReason() const { return lastInfo[1]; }
  double lastRelativeResidual() const { return lastRelRes; }
  const HipOpts* getOptions() const { return &opts; }

 private:
  static void check(const char* fn, int status) {
    if (status != 0) throw ErrorHandle("HipLinAlgebra", fn, hfx_last_error());
  }
  HipOpts opts;
  hfx_la* h = NULL;
  int lastInfo[2] = {0, 0};
  double lastRelRes = 0.0;
};

}  // namespace hfox

#endif
