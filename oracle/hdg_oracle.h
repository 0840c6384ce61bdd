/* FP64 CPU oracle for the HDG element-by-element path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against, and the CPU baseline the benchmark times beside it.  Only
 * the tests, the smoke check and the benchmark's CPU-baseline leg may load it;
 * the product library never links or calls it.
 *
 * It restates, in plain C++ with no external libraries, the reference's:
 *   ReferenceElement / Cubature      (src/element/ReferenceElement.cpp, Cubature.cpp)
 *   Operator::calcJacobians & co     (src/operator/Operator.cpp:14-84)
 *   HDGBase / HDGDiffusion / HDGConvection / Convection / Mass / Source / Reaction
 *   HDG models                       (src/model/HDG*.cpp)
 *   HDGSolver::calcElementalMatrices / applyBoundaryConditions / assembleSystem / solve
 *                                    (src/solver/HDGSolver.cpp:176-779)
 *   PetscInterface KSP semantics     (GMRES(restart) + left Jacobi, x0 = 0)
 *
 * Parity pinning of each function is listed in the tests that exercise it
 * (tests/test_oracle_*.py).
 */
#ifndef HFX_HDG_ORACLE_H
#define HFX_HDG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Last error message of the calling thread ("" if none). */
const char* orc_last_error(void);

/* Sizes of the order-p simplex: out[0..6] = nN, nF, nNf, nIP, nIPf, n_local, lenL. */
int orc_sizes(int dim, int p, int* out);

/* Reference tables.  Any pointer may be NULL.
 * nodes nN*dim, face_nodes nF*nNf, ipx nIP*dim, ipw nIP, phi nIP*nN,
 * dphi nIP*nN*dim, fipx nIPf*(dim-1), fipw nIPf, fphi nIPf*nNf. */
int orc_refel(int dim, int p, double* nodes, int* face_nodes, double* ipx, double* ipw, double* phi,
              double* dphi, double* fipx, double* fipw, double* fphi);

/* Cubature rule of the given order on the dim-simplex: returns nIP, fills
 * pts (nIP*dim) and w (nIP) when non-NULL. */
int orc_cubature(int dim, int order, double* pts, double* w);

/* Model description shared by the local and global entry points.
 * kind: 0 HDGLaplaceModel, 1 HDGDiffusionSource, 2 HDGConvectionDiffusionReactionSource.
 * D: per element node, NULL (identity for kinds 0/1, absent for kind 2),
 *    ndD = 1 (scalar) or dim*dim (column-major tensor).
 * b: per element node dim values or NULL.
 * f_ip, r_ip: source / reaction at the element IPs (nIP) or NULL. */
typedef struct {
  int kind;
  int ndD;
} orc_model;

/* Local matrix (n x n, row-major) and RHS (n) of one element, with unknowns
 * ordered [u (nN) | q (nN*dim, (i*dim+k)) | lambda (nF*nNf, cell-local)].
 * x: nN*dim node coords; tau: nF*nNf in cell-local face-node order. */
int orc_local_system(int dim, int p, orc_model model, const double* x, const double* D, const double* b,
                     const double* tau, const double* f_ip, const double* r_ip, double* K, double* F);

/* Implicit Euler on the u rows of a local system, in place (Euler::apply,
 * isExplicit = false, Euler.cpp:18-37): u rows and F_u times dt, plus the
 * element mass M and M u_old (u_old: nN cell-wise nodal values). */
int orc_apply_euler(int dim, int p, double dt, const double* u_old, double* K, double* F);
/* Explicit Euler (Euler::apply, isExplicit = true, Euler.cpp:30-35): u rows
 * and F_u times dt, F_u += (M - Suu) u_old, then Suu = M. */
int orc_apply_euler_explicit(int dim, int p, double dt, const double* u_old, double* K, double* F);
/* One Runge-Kutta stage on the u rows (RungeKutta::apply,
 * RungeKutta.cpp:113-159) with auxiliary fields {Flux, Trace}: a = Butcher
 * row `stage` without its first column (nst entries), ut = cell-local
 * [OldSolution | OldFlux | OldTrace] (n), kst = nst cell-local stage vectors
 * (nst x n). */
int orc_apply_rk(int dim, int p, double dt, int nst, const double* a, int stage, const double* ut, const double* kst,
                 double* K, double* F);

/* Explicit HDG modes (HDGSolver.cpp:346-353) on one local system: S = Sll,
 * S0 = F_lambda - Slu sol - Slq flux (sol nN, flux nN*dim). */
int orc_explicit_blocks(int dim, int p, const double* K, const double* F, const double* sol, const double* flux,
                        double* S, double* S0);

/* Physical coordinates of the element IPs (nIP*dim). */
int orc_ip_coords(int dim, int p, const double* x, double* out);

/* Static condensation of one local system by Householder QR, as
 * HDGSolver.cpp:331-348.  Outputs are row-major:
 * S lenL*lenL, S0 lenL, U lenU*lenL, Q lenQ*lenL, U0 lenU, Q0 lenQ. */
int orc_condense(int dim, int p, const double* K, const double* F, double* S, double* S0, double* U, double* Q,
                 double* U0, double* Q0);

/* Whole-mesh pipeline: elemental matrices, Dirichlet BCs, CSR trace system.
 * Mesh arrays follow the reference's Mesh: cells nEl*nN (global node ids),
 * faces nFaces*nNf, cell2face nEl*nF, face2cell nFaces*2.
 * Node fields: D nNodes*ndD (or NULL), b nNodes*dim (or NULL).
 * tau: nFaces*nNf*tau_sides (tau_sides = 1, or 2 for a double-valued field,
 *      side 0 belonging to face2cell[0]).
 * f_ip, r_ip: nEl*nIP or NULL.  dirichlet: nFaces*nNf values (used on
 * boundary faces; NULL means 0).
 * Returns an opaque handle; free with orc_system_free. */
typedef struct orc_system orc_system;
orc_system* orc_system_build(int dim, int p, orc_model model, int nNodes, const double* coords, int nEl,
                             const int* cells, int nFaces, const int* faces, const int* cell2face,
                             const int* face2cell, const double* D, const double* b, const double* tau,
                             int tau_sides, const double* f_ip, const double* r_ip, const double* dirichlet,
                             int nthreads);
void orc_system_free(orc_system* s);
/* Implicit Euler with step dt (0: none) and "Solution" field u_old
 * (nEl*nN) on every element, then reassemble the trace system. */
int orc_system_set_euler(orc_system* s, double dt, const double* u_old, int nthreads);
/* HDGSolverOpts::type 0 IMPLICIT, 1 WEXPLICIT, 2 SEXPLICIT: the explicit modes
 * set S = Sll and S0 = F_lambda - Slu sol - Slq flux per element (sol nEl*nN,
 * flux nEl*nN*dim), then reassemble the trace system. */
int orc_system_set_hdg_type(orc_system* s, int type, const double* sol, const double* flux, int nthreads);
/* Time scheme on every element, then reassemble: kind 0 none, 1 implicit
 * Euler, 2 explicit Euler (u_old nEl*nN), 3 Runge-Kutta stage `stage` with
 * Butcher row a (nst), old_local (nEl x n) and stage_local (nst x nEl x n)
 * cell-local vectors as orc_apply_rk takes them. */
int orc_system_set_time_scheme(orc_system* s, int kind, double dt, int nst, const double* a, int stage,
                               const double* u_old, const double* old_local, const double* stage_local, int nthreads);
/* Boundary model of the boundary faces `faces` (nfaces; NULL: every
 * boundary face), then reassemble: 0 DirichletModel (Set/Set, identity rows,
 * rhs g), 1 IntegratedDirichletModel (Set/Set, the face mass M and rhs M g,
 * IntegratedDirichletModel.cpp:35-56). */
int orc_system_set_boundary_model(orc_system* s, int kind, const int* faces, int nfaces, int nthreads);
/* HDG-type boundary model (HDGSolver.cpp:489-518) on the boundary faces
 * `faces` (replacing any earlier HDG-type set): B nfaces x nNf x nNf(2+dim),
 * the model's local matrix [B_u | B_q (node-major, dim per node) | B_lambda]
 * in face-node order, r nfaces x nNf its rhs; matrix_set / rhs_set: Set (1)
 * or Add (0) assembly.  Then reassemble. */
int orc_system_set_boundary_hdg(orc_system* s, const int* faces, int nfaces, const double* B, const double* r,
                                int matrix_set, int rhs_set, int nthreads);
/* Face mass of IntegratedDirichletModel: xf nNf*dim face nodes (face-node
 * order), M nNf*nNf. */
int orc_face_mass(int dim, int p, const double* xf, double* M);
/* Cell-local vectors [u | q | lambda] (nEl x n) of the fields u (nEl*nN),
 * q (nEl*nN*dim) and lambda (nFaces*nNf), the trace in each cell's face-node
 * order (Solver::prepareLocalFieldMap, Solver.cpp:14-57). */
int orc_system_local_vectors(const orc_system* s, const double* u, const double* q, const double* lambda, double* out);
/* HDGBurgersModel local system (HDGBurgersModel.cpp:86-132), m = dim
 * unknowns per node, n = nN m (dim + 1) + nF nNf m rows in the order u
 * (i m + a), q ((i dim + k) m + a), lambda ((f nNf + j) m + a).  Dn: nN
 * scalar diffusion coefficients (NULL: no HDGDiffusion); tau: nF nNf m^2,
 * tau_ab of each cell-local face node at (b m + a); sol: nN m (the
 * BufferSolution); trace: nF nNf m in the cell's face-node order; f_ip:
 * nIP m source values (NULL: none).  K n x n row-major, F n. */
int orc_burgers_local(int dim, int p, const double* x, const double* Dn, const double* tau, const double* sol,
                      const double* trace, const double* f_ip, double* K, double* F);
/* HDGnGammaModel local system (HDGnGammaModel.cpp:105-170), 2-D, m = 2
 * (n, Gamma), in orc_burgers_local's order.  D, G: nN scalars on the n and
 * Gamma diagonal blocks (NULL: absent); b: nN x 2; sol: nN x 2 (the
 * BufferSolution, n != 0); tau as orc_burgers_local; cs: the sound speed;
 * f_ip: nIP x 2 source values (NULL: none). */
int orc_ngamma_local(int p, const double* x, const double* D, const double* G, const double* b, const double* sol,
                     const double* tau, double cs, const double* f_ip, double* K, double* F);
/* Static condensation (HDGSolver.cpp:331-348) of a local system with lu, lq
 * and ll rows of u, q and lambda. */
int orc_condense_sizes(int lu, int lq, int ll, const double* K, const double* F, double* S, double* S0, double* U,
                       double* Q, double* U0, double* Q0);
/* sizes: out[0]=ndofs, out[1]=nnz */
int orc_system_sizes(const orc_system* s, long long* out);
int orc_system_csr(const orc_system* s, long long* rowptr, int* cols, double* vals, double* rhs);
/* Per-element condensed blocks after BCs (row-major, element-major). */
int orc_system_condensed(const orc_system* s, double* S, double* S0, double* U, double* Q, double* U0, double* Q0);

/* Solve the trace system.  method 0: GMRES(restart) with left Jacobi, zero
 * initial guess, stop on the preconditioned residual <= rtol*||P^-1 b||
 * (PETSc's KSPGMRES + PCJACOBI defaults); method 1: dense LU (small systems).
 * info[0]=iterations, info[1]=converged(1)/not(0); relres[0] = true
 * ||b-Ax||/||b||. */
int orc_system_solve(orc_system* s, int method, double rtol, int maxits, int restart, double* lambda, int* info,
                     double* relres);

/* Element reconstruction u = U lambda_e + U0, q = Q lambda_e + Q0. */
int orc_system_reconstruct(const orc_system* s, const double* lambda, double* u, double* q);

/* Wall time (s) of one [elemental matrices + condensation + BC + scatter]
 * pass over the first nEl_sample elements (the CPU baseline). */
double orc_system_time_assembly(orc_system* s, int nEl_sample, int nthreads);

/* Dense-matrix GMRES test hook for the LinAlgebraInterface known-answer
 * systems: solves A x = b (A n x n row-major) with GMRES(restart)+Jacobi. */
int orc_dense_gmres(int n, const double* A, const double* b, double rtol, int maxits, int restart, double* x,
                    int* iters);

#ifdef __cplusplus
}
#endif
#endif
