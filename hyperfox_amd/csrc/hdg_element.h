// Fused per-element HDG assembly + static condensation on the GPU.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hfx {

// Compile-time sizes of the order-P simplex in dimension D.
template <int D, int P>
struct ESz {
  static constexpr int NN = D == 2 ? (P + 1) * (P + 2) / 2 : (P + 1) * (P + 2) * (P + 3) / 6;
  static constexpr int NNF = D == 2 ? P + 1 : (P + 1) * (P + 2) / 2;
  static constexpr int NIP = D == 2 ? (P == 1 ? 3 : P == 2 ? 6 : P == 3 ? 12 : P == 4 ? 16 : P == 5 ? 25
                                      : P == 6 ? 33 : P == 7 ? 42 : P == 8 ? 55 : P == 9 ? 67 : 79)
                                    : (P == 1 ? 4 : P == 2 ? 14 : P == 3 ? 24 : P == 4 ? 46 : 81);
  static constexpr int NIPF = D == 2 ? P + 1 : (P == 1 ? 3 : P == 2 ? 6 : P == 3 ? 12 : P == 4 ? 16 : 25);
  static constexpr int NF = D + 1;
  static constexpr int NL = NF * NNF;
  static constexpr int NQ = NN * D;
  static constexpr int NLOC = NN * (D + 1) + NL;
  // packed reference tables: PHI[NIP][NN] DPHI[NIP][NN][D] FPHI[NIPF][NNF]
  // FDPHI[NIPF][NNF][D-1] W[NIP] FW[NIPF]
  static constexpr int T_PHI = 0;
  static constexpr int T_DPHI = T_PHI + NIP * NN;
  static constexpr int T_FPHI = T_DPHI + NIP * NN * D;
  static constexpr int T_FDPHI = T_FPHI + NIPF * NNF;
  static constexpr int T_W = T_FDPHI + NIPF * NNF * (D - 1);
  static constexpr int T_FW = T_W + NIP;
  static constexpr int TAB = T_FW + NIPF;
};

// Arguments of one element-kernel launch (all device pointers).
struct ElemArgs {
  int nEl = 0;
  int kind = 0;        // 0 Laplace, 1 diffusion-source, 2 convection-diffusion-reaction-source
  int useDiff = 1;     // diffusion operator present
  int ndD = 0;         // 0: identity, 1: scalar per node, D*D: column-major tensor per node
  int hasB = 0, hasF = 0, hasR = 0;
  int tau_sides = 1;
  int el_offset = 0;   // first element handled by this launch
  const double* tabs = nullptr;
  const int* faceNodes = nullptr;  // NF*NNF
  const double* coords = nullptr;
  const int* cells = nullptr;
  const int* cell2face = nullptr;
  const int* face2cell = nullptr;
  const int8_t* face_perm = nullptr;
  const double* Dn = nullptr;
  const double* bn = nullptr;
  const double* tau = nullptr;
  const double* f_ip = nullptr;
  const double* r_ip = nullptr;
  // condensed outputs (element-major, row-major blocks)
  double *S = nullptr, *S0 = nullptr, *U = nullptr, *Q = nullptr, *U0 = nullptr, *Q0 = nullptr;
  // optional: full local matrices (NLOC x NLOC) and RHS instead of condensation
  double *Kloc = nullptr, *Floc = nullptr;
  // optional (affine kernel): fused trace scatter instead of storing S.  The
  // off-diagonal face blocks go straight into the BSR values bvals at block
  // bpos[(e*NF + f)*NF + g] in face-node order (zero rows on Dirichlet faces
  // is_dir); the diagonal block of face F from its side s (0 when
  // face2cell[2F] == e) goes to dside[(2F + s)*NNF*NNF ...], summed by
  // launch_scatter_diag.
  double* bvals = nullptr;
  double* dside = nullptr;
  const int* bpos = nullptr;
  const uint8_t* is_dir = nullptr;
  // optional: implicit Euler on the u rows (Euler::apply, Euler.cpp:18-37):
  // with ts_dt > 0 the element matrix of the model plus M / dt in Suu and
  // M u_old / dt in F_u, u_old (nEl x nN, the "Solution" field) -- the
  // reference's dt-scaled rows divided by dt, which condenses to the same
  // blocks.  The local-matrix export scales the u rows back by dt.
  double ts_dt = 0.0, ts_idt = 0.0;
  const double* u_old = nullptr;
  // optional: general time transform of the u rows (explicit Euler,
  // Runge-Kutta stages; ts_gen = 1, ts_idt = 0): with the model's rows
  // [Suu | Suq | Sul], F_u and the mass M,
  //   Suu <- ts_auu Suu + ts_mu M,   [Suq | Sul] <- ts_arest [Suq | Sul],
  //   F_u <- ts_phi F_u + ts_mu M u_old - [Suu | Suq | Sul] w,
  // w = the cell's local vector [u | q | lambda] in ts_w (nEl x NLOC, lambda
  // in cell-local face order).  Each scheme's rows divided by a constant
  // (condensation does not see it); ts_dt is the factor that restores them
  // in the local-matrix export.
  int ts_gen = 0;
  double ts_auu = 1.0, ts_arest = 1.0, ts_mu = 0.0, ts_phi = 1.0;
  const double* ts_w = nullptr;
  // optional: explicit HDG modes (WEXPLICIT / SEXPLICIT, HDGSolver.cpp:346-353):
  // S = Sll and S0 = F_lambda - Slu sol_old - Slq flux_old, with sol_old
  // (nEl x nN, "Solution") and flux_old (nEl x nN x D, "Flux"); U, Q, U0, Q0
  // as in the implicit mode
  int hdg_explicit = 0;
  const double* sol_old = nullptr;
  const double* flux_old = nullptr;
  // orders whose element working set exceeds the LDS (3-D p = 4, 5): a
  // device workspace of element_workspace_doubles() doubles per workgroup,
  // gws_slots of them; the grid is capped at gws_slots
  double* gws = nullptr;
  int gws_slots = 0;
};

// Launch the element kernel for (dim, order) on `count` elements starting at
// a.el_offset; `nt` threads per element (64, 128 or 256).  Throws on error.
void launch_element(int dim, int order, const ElemArgs& a, int nt, int grid_cap, hipStream_t s);

// Dynamic LDS bytes the kernel needs for (dim, order): 0 for the orders
// that keep the element's working set in a device workspace instead.
size_t element_lds_bytes(int dim, int order);
// Doubles of device workspace per workgroup for (dim, order); 0 when the
// working set is in LDS.
size_t element_workspace_doubles(int dim, int order);

// Affine-cell kernel (hdg_affine.hip): one wavefront per cell, `waves`
// wavefronts per workgroup; scalar or identity diffusion only.
void launch_affine(int dim, int order, bool tensor_d, int waves, const ElemArgs& a, int grid_cap, hipStream_t s);
size_t affine_lds_bytes(int dim, int order, bool tensor_d, int waves);

// One work item of the batched affine kernels: cells [e0, e1) of problem
// `prob` (index into the ElemArgs table) of order `order` (2-D).
struct BatchItem {
  int prob = 0, order = 0, e0 = 0, e1 = 0;
};

// Batched affine-cell kernel over the cells of many 2-D problems of one
// order (1..5); `probs` and `items` are device arrays; grid = number of
// workgroups (persistent).
void launch_affine_batch(int order, int waves, const ElemArgs* probs, const BatchItem* items, int nitems, int grid,
                         hipStream_t s, bool attribute_only = false);
size_t affine_batch_lds_bytes(int order, int waves);
// Raise the batched kernel's dynamic-LDS limit (done outside graph capture).
void prepare_affine_batch(int order, int waves);

// Physical coordinates of the element IPs: out[(e*NIP + q)*D + k].
void launch_ip_coords(int dim, int order, int nEl, const double* tabs, const double* coords, const int* cells,
                      double* out, hipStream_t s);

}  // namespace hfx
