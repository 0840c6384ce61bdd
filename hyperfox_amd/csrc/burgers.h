// Burgers cell kernel: the HDGBurgersModel local system (m = dim unknowns
// per node) assembled and condensed onto the face traces on the GPU.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/hfx_hip.h"

namespace hfx {

// Arguments of one Burgers cell-kernel launch (device pointers).
//   tabs       packed reference tables (ESz layout)
//   tau        nFaces x nNf x sides x m^2 (tau_ab at b m + a, side s of the
//              face's cell face2cell[2F + s])
//   Dn         nNodes scalar diffusion coefficients, null: no HDGDiffusion
//   sol        nEl x nN x m, the linearisation state ("BufferSolution")
//   trace      nFaces x nNf x m, the linearisation trace ("Trace")
//   f_ip       nEl x nIP x m source values at the cell IPs, null: none
// Outputs, element-major and row-major, lambda in the cell's local
// face-node order ((f nNf + j) m + a): S (mL x mL), S0 (mL), U (mN x mL),
// Q (d mN x mL, row (i d + k) m + a), U0, Q0.
struct BurgersArgs {
  int nEl = 0;
  int tau_sides = 1;
  const double* tabs = nullptr;
  const int* faceNodes = nullptr;
  const double* coords = nullptr;
  const int* cells = nullptr;
  const int* cell2face = nullptr;
  const int* face2cell = nullptr;
  const int8_t* face_perm = nullptr;
  const double* Dn = nullptr;
  const double* tau = nullptr;
  const double* sol = nullptr;
  const double* trace = nullptr;
  const double* f_ip = nullptr;
  // model 1: HDGnGammaModel (2-D, m = 2): Dn is its D field (n block), G
  // (nNodes, null: absent) the Gamma block's coefficient, bfield the nodal
  // b vectors (nNodes x 2), sound_speed c_s.  Model 2: a scalar model at
  // m = dim DOFs per node (HDGLaplaceModel, HDGDiffusionSource): diffusion
  // Dn on every component, no convection; sol and trace are not used.
  int model = 0;
  // implicit Euler on the u rows (RungeKutta BEuler / Euler, the rows divided
  // by dt): with ts_idt = 1/dt > 0, Suu + (M / dt) (x) I and F_u + (M (x) I)
  // u_old / dt, u_old nEl x nN x m (the "OldSolution")
  double ts_idt = 0.0;
  const double* u_old = nullptr;
  const double* G = nullptr;
  const double* bfield = nullptr;
  double sound_speed = 1.0;
  double *S = nullptr, *S0 = nullptr, *U = nullptr, *Q = nullptr, *U0 = nullptr, *Q0 = nullptr;
  // 3-D p = 3 (working set above the LDS): a device workspace of
  // burgers_workspace_doubles() per workgroup, gws_slots of them
  double* gws = nullptr;
  int gws_slots = 0;
};

// True when (dim, order) has a Burgers cell kernel: 2-D orders 1..5 and
// 3-D orders 1..3 (m nN <= 64 rows for the workgroup Gauss-Jordan; 3-D p = 3
// keeps its working set in a device workspace).
bool burgers_supported(int dim, int order);
// Doubles of device workspace per workgroup (0: the cell fits the LDS).
size_t burgers_workspace_doubles(int dim, int order, int model);

// Launch over all a.nEl cells (one workgroup of 256 threads per cell,
// grid-stride); throws on error.
void launch_burgers(int dim, int order, const BurgersArgs& a, hipStream_t s);

// m-fold face permutations: out[(e nF + f) nNf m + j m + a] = perm[...j] m + a.
void launch_expand_perm(long long nblk, int nNf, int m, const int8_t* perm, int8_t* out, hipStream_t s);

// NonLinearWrapper residual partials and update (NonLinearWrapper.cpp:12-79):
// sums of (cur - prev)^2 and prev^2 into out2[0..1] (deterministic two-stage
// reduction), and cur = prev = (1 - damp) cur + damp prev.
void burgers_residual(long long n, const double* cur, const double* prev, double* partial, double* out2,
                      hipStream_t s);
void burgers_damp(long long n, double damp, double* cur, double* prev, hipStream_t s);

// HDGBohmModel rows of the 2-D (n, Gamma) problem (HDGBohmModel.cpp:97-268)
// on the boundary faces faces[nfaces]: B (nfaces x 2 nNf x 8 nNf, columns
// u (j 2 + a), q ((j 2 + k) 2 + a), lambda (j 2 + a)) and r (nfaces x 2 nNf)
// from the face-local fields, as HDGSolver::applyBoundaryConditions gathers
// them: Solution and Flux of the face's cell at the face nodes, Trace, Tau
// (side 0) and the exterior normals in face-node order, b, D and G at the
// face nodes; fphi, fdphi, fw the face element's tables at its cubature.
struct BohmArgs {
  int nNf = 0, nipf = 0, nN = 0, nFpc = 0, tau_sides = 1, kind = HFX_BOHM_HEAVISIDE;
  const int* faces = nullptr;
  const double* normals = nullptr;   // nfaces x nNf x 2
  const int* face2cell = nullptr;
  const int* cell2face = nullptr;
  const int* faceNodes = nullptr;
  const int8_t* invperm = nullptr;
  const int* cells = nullptr;
  const double* coords = nullptr;
  const double *fphi = nullptr, *fdphi = nullptr, *fw = nullptr;
  const double *sol = nullptr, *flux = nullptr, *trace = nullptr, *tau = nullptr;
  const double *D = nullptr, *G = nullptr, *bfield = nullptr;
  double sound_speed = 1.0;
  double *B = nullptr, *r = nullptr;
};
void launch_bohm(const BohmArgs& a, int nfaces, hipStream_t s);

}  // namespace hfx
