/* C ABI of libhfx_hip.so: the MI355X (gfx950) HDG element-by-element hot path.
 *
 * Every entry point returns 0 on success and -1 on failure; the failure text is
 * available from hfx_last_error() in the calling thread.  Plain pointers and
 * sizes only.  All double arrays are FP64, all index arrays int32 unless noted,
 * all 2-D arrays row-major.  Host pointers are copied; the library owns its
 * device memory.  Calls on one handle are stream-ordered and return after the
 * work they describe has completed.
 *
 * The groups below replace, in the reference (julienfausty/HyperFox):
 *   mesh      src/mesh/Mesh.cpp:31-46, 183-537 (setMesh, computeFaces, orderings)
 *   refel     src/element/ReferenceElement.cpp:5-26, Cubature.cpp:52-59
 *   hdg       src/solver/HDGSolver.cpp:5-106 (allocate), 166-675 (assemble),
 *             677-779 (solve + reconstruction), with the HDG models of
 *             src/model/HDG*.cpp:* and operators of src/operator/HDG*.cpp:*
 *   la        src/resolution/LinAlgebraInterface.h:23-162 (the plug point),
 *             behaviour of src/resolution/PetscInterface.cpp:5-265
 */
#ifndef HFX_HIP_H
#define HFX_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

/* Last error message of the calling thread ("" if none). */
const char* hfx_last_error(void);

/* Number of visible GPUs (0 when no GPU).  Does not create a context. */
int hfx_device_count(int* n);

/* ------------------------------------------------------------ reference element */

/* out[0..6] = nN, nF (faces per cell), nNf, nIP, nIPf, n_local, lenL
 * (ReferenceElement::getNumNodes/.../Cubature::getNumIPs, HDGOperator.cpp:5-15). */
int hfx_refel_sizes(int dim, int order, int* out);

/* Host tables of the reference element (ReferenceElement getters,
 * ReferenceElement.h:29-119).  Any pointer may be NULL.  nodes nN*dim,
 * face_nodes nF*nNf, ipx nIP*dim, ipw nIP, phi nIP*nN, dphi nIP*nN*dim,
 * fipx nIPf*(dim-1), fipw nIPf, fphi nIPf*nNf. */
int hfx_refel_tables(int dim, int order, double* nodes, int* face_nodes, double* ipx, double* ipw, double* phi,
                     double* dphi, double* fipx, double* fipw, double* fphi);

/* Cubature on the dim-simplex exact to `order` (Cubature::Cubature,
 * Cubature.cpp:5-26): returns the point count (or -1), fills pts/w if non-NULL. */
int hfx_cubature(int dim, int order, double* pts, double* w);

/* ----------------------------------------------------------------------- mesh */

typedef struct hfx_mesh hfx_mesh;

/* Unit square (dim 2, n x n squares, 2 triangles each) or cube (dim 3, n^3
 * cubes, 6 tetrahedra each) with interior vertices jittered by
 * U(-jitter*h, jitter*h) from a splitmix64 stream seeded with `seed`;
 * high-order nodes by the affine map of tools/convertGmsh2H5HO.cpp:367-397. */
int hfx_mesh_structured(int dim, int order, int n, double jitter, unsigned long long seed, hfx_mesh** out);

/* Mesh from high-order cells (HDF5Io::load + Mesh::setMesh,
 * HDF5Io.cpp:13-62, Mesh.cpp:31-46): cells nCells*nN global node ids in the
 * reference node order, coords nNodes*dim. */
int hfx_mesh_from_cells(int dim, int order, int nNodes, const double* coords, int nCells, const int* cells,
                        hfx_mesh** out);

/* out[0..4] = nNodes, nCells, nFaces, nBoundaryFaces, trace DOFs. */
/* The same meshes with the topology (faces, cell2face, face2cell, boundary,
 * face-node permutations; the reference's MOAB computeFaces,
 * src/mesh/Mesh.cpp:183-537) built on GPU `device` by a key sort instead of
 * the host hash pass; identical arrays.  3-D meshes need fewer than 2^21
 * vertices. */
int hfx_mesh_structured_gpu(int device, int dim, int order, int n, double jitter, unsigned long long seed,
                            hfx_mesh** out);
int hfx_mesh_from_cells_gpu(int device, int dim, int order, int nNodes, const double* coords, int nCells,
                            const int* cells, hfx_mesh** out);

int hfx_mesh_sizes(const hfx_mesh* m, long long* out);

/* Copies of the topology (Mesh getters, Mesh.cpp:145-161, 292-295, 435-448,
 * 513-515).  Any pointer may be NULL.  faces nFaces*nNf (first cell's local
 * face-node order), cell2face nCells*nF, face2cell nFaces*2 (-1 = boundary),
 * boundary ascending face ids, face_perm nCells*nF*nNf int8 (face2CellMap of
 * HDGSolver.cpp:258-275). */
int hfx_mesh_get(const hfx_mesh* m, double* coords, int* cells, int* faces, int* cell2face, int* face2cell,
                 int* boundary, signed char* face_perm);

void hfx_mesh_free(hfx_mesh* m);

/* ------------------------------------------------------- HDG device bulk path */

typedef struct hfx_hdg hfx_hdg;

/* Model kinds: HDGLaplaceModel, HDGDiffusionSource,
 * HDGConvectionDiffusionReactionSource (src/model/HDG*.cpp). */
enum { HFX_MODEL_LAPLACE = 0, HFX_MODEL_DIFFUSION_SOURCE = 1, HFX_MODEL_CDR_SOURCE = 2 };

/* Trace preconditioners: none, point Jacobi (PCJACOBI, the reference's
 * default, PetscOpts.h:16), face-block Jacobi, the learned correction of
 * face-block Jacobi, and block ILU(0) (PCILU, zero fill, on the face-block
 * pattern) in a multicolour elimination order (rows of one colour share no
 * block: as many sweep levels as colours) or in the natural face order
 * (PETSc's default ordering; one sweep level per dependency level). */
enum {
  HFX_PC_NONE = 0,
  HFX_PC_JACOBI = 1,
  HFX_PC_BJACOBI = 2,
  HFX_PC_LEARNED = 3,
  HFX_PC_ILU0 = 4,
  HFX_PC_ILU0_NATURAL = 5,
  /* HFX_PC_ILU0 (multicolour) factorised in FP64 and applied from FP32
   * copies of the factors, summed in FP64: half the bytes per sweep.  The
   * solve still stops on the FP64 true residual; only the iteration count
   * can differ.  No counterpart in the reference (PCILU is FP64). */
  HFX_PC_ILU0_F32 = 6
};

/* Create a problem on GPU `device` for simplices of `dim` and `order`
 * (HDGSolver construction + initialize, Solver.cpp:5-12). */
int hfx_hdg_create(int device, int dim, int order, hfx_hdg** out);
void hfx_hdg_destroy(hfx_hdg* h);

/* Upload the mesh (Solver::setMesh + HDGSolver::allocate's pattern, HDGSolver.cpp:117-164). */
int hfx_hdg_set_mesh(hfx_hdg* h, const hfx_mesh* m);

/* Model kind (Solver::setModel). */
int hfx_hdg_set_model(hfx_hdg* h, int kind);

/* Field map (Solver::setFieldMap; field layouts checked as in HDGSolver.cpp:24-73):
 *   tau       nFaces*nNf*tau_sides ("Tau" face field; tau_sides 2 = double valued,
 *             side 0 belongs to face2cell[0])
 *   D         nNodes*ndD or NULL ("DiffusionTensor" node field, ndD = 1 or dim*dim
 *             column-major); NULL = identity for kinds 0/1, no diffusion for kind 2
 *   b         nNodes*dim or NULL ("Velocity" node field)
 *   dirichlet nFaces*nNf or NULL ("Dirichlet" face field, used on boundary faces). */
int hfx_hdg_set_fields(hfx_hdg* h, const double* tau, int tau_sides, const double* D, int ndD, const double* b,
                       const double* dirichlet);

/* Physical IP coordinates computed on the GPU, out nCells*nIP*dim, so the
 * caller can evaluate its source/reaction functions once per problem
 * (replaces the per-IP callbacks of Source.cpp:5-22, Reaction.cpp:5-20). */
int hfx_hdg_ip_coords(hfx_hdg* h, double* out);

/* Source and reaction values at the element IPs (nCells*nIP each, or NULL). */
int hfx_hdg_set_source(hfx_hdg* h, const double* f_ip, const double* r_ip);

/* Launch shape of the element kernel: threads per element of the general
 * kernel (64/128/256) and a cap on the grid (0 = default: one workgroup per
 * element for the general kernel; for the affine kernel as many workgroups
 * as are resident on the device at once, each looping over its cells). */
int hfx_hdg_set_launch(hfx_hdg* h, int threads, int grid_cap);

/* Element kernel selection: 0 automatic (the affine-cell kernel when every
 * cell is straight-sided and diffusion is scalar, else the general kernel
 * with per-IP Jacobians), 1 general, 2 affine. */
int hfx_hdg_set_kernel(hfx_hdg* h, int which);

/* out[0] = kernel in use (1 general, 2 affine), out[1] = threads per
 * workgroup, out[2] = mesh is affine, out[3] = dynamic LDS bytes. */
int hfx_hdg_kernel_info(hfx_hdg* h, int* out);

/* Elemental matrices + static condensation + Dirichlet rows + trace-matrix
 * scatter, all on the GPU (HDGSolver::assemble, HDGSolver.cpp:166-675). */
int hfx_hdg_assemble(hfx_hdg* h);

/* Assembly-only: local matrices (n_local^2, reference unknown order
 * [u | q (i*dim+k) | lambda]) and RHS of cells [first, first+count)
 * (FEModel::getLocalMatrix/getLocalRHS after Model::compute, HDGModel.cpp:35-51). */
int hfx_hdg_local_matrices(hfx_hdg* h, int first, int count, double* K, double* F);

/* Condensed blocks of every cell before the Dirichlet rows are applied,
 * row-major: S lenL^2, S0 lenL, U nN*lenL, Q (nN*dim)*lenL, U0 nN, Q0 nN*dim. */
int hfx_hdg_get_condensed(hfx_hdg* h, double* S, double* S0, double* U, double* Q, double* U0, double* Q0);

/* out[0..2] = trace DOFs, scalar nnz, block nnz. */
int hfx_hdg_trace_sizes(hfx_hdg* h, long long* out);

/* Trace system as CSR (rows sorted by column) and its RHS (after assemble). */
int hfx_hdg_get_trace_csr(hfx_hdg* h, long long* rowptr, int* cols, double* vals, double* rhs);

/* Time scheme on the u rows of every element matrix (HDGModel::compute,
 * HDGModel.cpp:39-47).  HFX_TS_BEULER is Euler with isExplicit = false
 * (Euler::apply, Euler.cpp:18-37): u rows scaled by dt, plus the element
 * mass M in Suu and M u_old in F_u, where u_old is the problem's "Solution"
 * field at assembly (the last solve's u, or hfx_hdg_set_solution).
 * HFX_TS_FEULER is Euler with isExplicit = true (Euler.cpp:30-35): u rows
 * scaled by dt, F_u += (M - Suu) u_old, Suu = M.  HFX_TS_RK is RungeKutta
 * (RungeKutta.cpp:113-159) with the auxiliary fields Flux and Trace, driven
 * by hfx_hdg_set_butcher and hfx_hdg_rk_*.  dt must be > 0, as the
 * reference's apply requires. */
enum { HFX_TS_NONE = 0, HFX_TS_BEULER = 1, HFX_TS_FEULER = 2, HFX_TS_RK = 3 };
int hfx_hdg_set_time_scheme(hfx_hdg* h, int scheme, double dt);

/* Butcher tables of RungeKutta::setUpDB (RungeKutta.cpp:197-285), in the
 * order of RKType (src/operator/RKType.h); HFX_RK_CUSTOM takes a table. */
enum {
  HFX_RK_FEULER = 0, HFX_RK_EMIDPOINT = 1, HFX_RK_HEUN = 2, HFX_RK_KUTTA3 = 3, HFX_RK_HEUN3 = 4,
  HFX_RK_SSPRK3 = 5, HFX_RK_RK4 = 6, HFX_RK_BEULER = 7, HFX_RK_IMIDPOINT = 8, HFX_RK_CRANKNICOLSON = 9,
  HFX_RK_KS2 = 10, HFX_RK_QZ2 = 11, HFX_RK_ALX2 = 12, HFX_RK_RK43 = 13, HFX_RK_CUSTOM = -1
};
/* RungeKutta::setButcherTable: a table of the database (type >= 0), or the
 * n x n row-major `table` (HFX_RK_CUSTOM), which must have zeros above its
 * diagonal, as setButcherTable(EMatrix) requires (RungeKutta.cpp:17-31).
 * Row i < n-1 is [c_i | a_i0 .. ], the last row [0 | b_0 .. ]; at most 8
 * stages.  Resets the stage counter. */
int hfx_hdg_set_butcher(hfx_hdg* h, int type, const double* table, int n);
/* The database table `type`: n[0] = its size, table (n x n row-major, room
 * for 25 values) may be NULL. */
int hfx_rk_table(int type, double* table, int* n);
/* Start a Runge-Kutta time step: OldSolution / OldFlux / OldTrace = the
 * current Solution u, Flux q and Trace lambda, every RKStage field zero,
 * stage counter 0 (the caller's step set-up in the reference's drivers,
 * e.g. tests/parallel/TestParHDGConvectionDiffusionReactionSource.cpp:281-292).
 * Each stage is then hfx_hdg_assemble, hfx_hdg_solve and
 * hfx_hdg_rk_compute_stage; hfx_hdg_rk_compute_solution ends the step. */
int hfx_hdg_rk_start_step(hfx_hdg* h);
/* RungeKutta::computeStage (RungeKutta.cpp:161-203) on u, q and lambda. */
int hfx_hdg_rk_compute_stage(hfx_hdg* h);
/* RungeKutta::computeSolution (RungeKutta.cpp:205-241) on u, q and lambda. */
int hfx_hdg_rk_compute_solution(hfx_hdg* h);
/* out[0] = stages of the table (RungeKutta::getNumStages), out[1] = stage counter. */
int hfx_hdg_rk_stages(hfx_hdg* h, int* out);
/* Boundary model of the boundary faces `faces` (n of them; NULL: every
 * boundary face), applied by hfx_hdg_assemble (HDGSolver::
 * applyBoundaryConditions, HDGSolver.cpp:361-529): HFX_BC_DIRICHLET is
 * DirichletModel (the default: identity rows, rhs g; DirichletModel.cpp:19-44),
 * HFX_BC_INTEGRATED_DIRICHLET is IntegratedDirichletModel (rows = the face
 * mass M of the face's nodes, rhs M g; IntegratedDirichletModel.cpp:35-56).
 * Both Set the face's rows and RHS.  A non-boundary face is an error. */
enum { HFX_BC_DIRICHLET = 0, HFX_BC_INTEGRATED_DIRICHLET = 1 };
int hfx_hdg_set_boundary_model(hfx_hdg* h, int kind, const int* faces, int n);
/* HDG-type boundary model (BoundaryModelType HDGType, applied by
 * HDGSolver::applyBoundaryConditions, HDGSolver.cpp:503-518) on the boundary
 * faces `faces` (n; n = 0 clears; replaces any earlier HDG-type set): B holds
 * n local matrices of nNf x nNf(2 + dim), [B_u | B_q (node-major, dim per
 * node) | B_lambda] in face-node order, as a boundary model's compute()
 * returns them, r n right-hand sides of nNf; matrix_set / rhs_set choose Set
 * (nonzero) or Add assembly.  Each face's rows become (or gain)
 * B_u U_F + B_q Q_F + B_lambda and r - B_u U0_F - B_q Q0_F. */
int hfx_hdg_set_boundary_hdg(hfx_hdg* h, const int* faces, int n, const double* B, const double* r, int matrix_set,
                             int rhs_set);
/* Set the "Trace" field lambda (nFaces x nNf, face-node order), e.g. an
 * initial condition; hfx_hdg_solve overwrites it. */
int hfx_hdg_set_trace(hfx_hdg* h, const double* lambda);
/* Set the "Solution" field u (nCells x nN, cell-wise nodal values), e.g. an
 * initial condition; hfx_hdg_solve overwrites it. */
int hfx_hdg_set_solution(hfx_hdg* h, const double* u);

/* HDGSolverOpts::type (src/solver/HDGSolverOpts.h): IMPLICIT (default),
 * WEXPLICIT or SEXPLICIT.  The explicit modes condense with S = Sll and
 * S0 = F_lambda - Slu u - Slq q, u and q being the problem's current
 * "Solution" and "Flux" fields (HDGSolver.cpp:346-353); WEXPLICIT solves the
 * resulting block-diagonal trace system with the Krylov method and
 * preconditioner of hfx_hdg_solve, SEXPLICIT by face-local direct solves
 * (HDGSolver.cpp:709-729; hfx_hdg_solve then reports one iteration and the
 * true residual). */
enum { HFX_HDG_IMPLICIT = 0, HFX_HDG_WEXPLICIT = 1, HFX_HDG_SEXPLICIT = 2 };
int hfx_hdg_set_solver_type(hfx_hdg* h, int type);
/* Set the "Flux" field q (nCells x nN x dim, index (i*dim + k)). */
int hfx_hdg_set_flux(hfx_hdg* h, const double* q);

/* Trace scatter of hfx_hdg_assemble on straight-sided cells: fused (on = 1,
 * the default: the element kernel writes each off-diagonal face block of the
 * trace matrix in place and one face-diagonal pass adds the two cells'
 * diagonal contributions, so S is never stored) or stored (on = 0: S is
 * stored and gathered, as in HDGSolver::assembleSystem,
 * HDGSolver.cpp:531-675).  Both give the same trace matrix bit for bit;
 * hfx_hdg_get_condensed stores S on demand. */
int hfx_hdg_set_fused_scatter(hfx_hdg* h, int on);

/* Krylov method of hfx_hdg_solve: restarted FGMRES (default; the reference's
 * KSPGMRES) or BiCGStab with device-resident scalars (KSPBCGS). */
enum { HFX_KRYLOV_FGMRES = 0, HFX_KRYLOV_BICGSTAB = 1 };
int hfx_hdg_set_krylov(hfx_hdg* h, int method);
/* BiCGStab with a preconditioner runs the steady iterations between two
 * residual checks (8) as one captured HIP graph (on = 1) or as separate
 * launches (on = 0, the default: graphs measured no faster on MI355X, and
 * each solve pays the capture); the kernels and their order are the same, so
 * are the results.  No counterpart in the reference. */
int hfx_hdg_set_solver_graphs(hfx_hdg* h, int on);

/* Trace solve (LinAlgebraInterface::solve, HDGSolver.cpp:699-708): FP64
 * Krylov method of hfx_hdg_set_krylov with preconditioner `pc` (restart
 * applies to FGMRES), zero initial guess, stopping when
 * the true residual ||b - A lambda||_2 / ||b||_2 <= rtol.  info[0] =
 * iterations, info[1] = converged reason (2 rtol, 3 zero rhs, -3 max its,
 * -9 nan); relres[0] = final true relative residual. */
int hfx_hdg_solve(hfx_hdg* h, int pc, double rtol, int maxits, int restart, int* info, double* relres);

/* The trace solve of hfx_hdg_solve with no device allocation.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  BiCGStab (HFX_KRYLOV_BICGSTAB) with `pc` HFX_PC_NONE,
 * HFX_PC_JACOBI or HFX_PC_BJACOBI, zero initial guess, stopping on the true
 * residual ||b - A lambda||_2 / ||b||_2 <= rtol, then the reconstruction of u
 * and q; the same kernels in the same order as hfx_hdg_solve with BiCGStab, so
 * the same iterations and fields bit for bit.  The workspace is a fixed
 * 64 MiB device array per GPU, held until the process ends (about
 * (8 + nNf) doubles per trace DOF with block Jacobi: up to ~0.6M trace DOFs);
 * a larger system is refused with an error.  Calls that overlap on one GPU
 * (several host threads) race on that array and must be serialised.
 * info[0] = iterations, info[1] = converged reason; relres[0] = final true
 * relative residual. */
int hfx_hdg_solve_noalloc(hfx_hdg* h, int pc, double rtol, int maxits, int* info, double* relres);

/* Set up preconditioner `pc` for the assembled trace matrix and apply it
 * once: y = M^-1 x (host vectors of trace DOFs).  For tests and callers that
 * drive their own iteration. */
int hfx_hdg_pc_apply(hfx_hdg* h, int pc, const double* x, double* y);

/* Block ILU(0) factors of the assembled trace matrix in `ordering` (0
 * multicolour, 1 natural): lu (nnzb x nNf^2, the trace CSR's block order:
 * earlier-column blocks L_ik, the diagonal A'_ii, later-column blocks
 * Dinv_i U_ij), dinv (nFaces x nNf^2, inverse pivot blocks); either may be
 * NULL.  info[0..2] = colours, forward levels, backward levels. */
int hfx_hdg_ilu_factors(hfx_hdg* h, int ordering, double* lu, double* dinv, int* info);

/* Host analysis of a block pattern for ILU(0) (no GPU): rank[nb] (position
 * in the elimination order), levf[nb], levb[nb] (sweep levels) and colour[nb]
 * (any may be NULL); info[0..3] = colours, forward levels, backward levels,
 * update pairs of the factorisation. */
int hfx_ilu_analyse(int nb, const int* rowptr, const int* col, int ordering, int* rank, int* levf, int* levb,
                    int* colour, long long* info);

/* Solution fields after solve (HDGSolver.cpp:741-775): u nCells*nN ("Solution"),
 * q nCells*nN*dim ("Flux"), lambda nFaces*nNf ("Trace").  Any may be NULL. */
int hfx_hdg_get_solution(hfx_hdg* h, double* u, double* q, double* lambda);

/* Timed repetitions with HIP events on the problem's stream:
 * what = 0 element kernel only, 1 element kernel + scatter (assembly+condense),
 * 2 trace solve with the last solve's options.  ms receives `reps` values. */
int hfx_hdg_time(hfx_hdg* h, int what, int warmup, int reps, double* ms);

/* Dynamic LDS bytes used by the general element kernel of (dim, order); 0
 * for the orders (3-D p = 4, 5) whose element working set does not fit the
 * LDS and lives in a device workspace instead (hfx_element_workspace_bytes).
 * The reference supports 2-D p <= 10 and 3-D p <= 5
 * (ReferenceElement.cpp:620-623), and so does the GPU path (2-D p >= 7 and
 * 3-D p >= 4 in the device workspace). */
int hfx_element_lds_bytes(int dim, int order, long long* bytes);
/* Device workspace bytes per workgroup of the general element kernel of
 * (dim, order) (0 when its working set is in LDS); a problem allocates it for
 * up to twice as many workgroups as the GPU has compute units. */
int hfx_element_workspace_bytes(int dim, int order, long long* bytes);

/* ------------------------------------------- linear-algebra plug point (la) */

/* A sparse FP64 system A x = b behind the reference's LinAlgebraInterface
 * (src/resolution/LinAlgebraInterface.h:23-162), with the behaviour of its
 * PETSc implementation (src/resolution/PetscInterface.cpp:5-265): the same
 * call order and misuse errors, ADD/INSERT semantics, row-major value blocks
 * (MatSetValues' default), zero initial guess.  Values are gathered on the
 * host and reduced into CSR at assembly; the solve runs on the GPU.  Only
 * hfx_la_solve touches the GPU. */
typedef struct hfx_la hfx_la;

/* Solver types of hfx_la_configure (PetscOpts::solverType, PetscOpts.h:14):
 * KSPGMRES and KSPBCGS. */
enum { HFX_LA_GMRES = 0, HFX_LA_BCGS = 1 };

/* Construction (PetscInterface::PetscInterface, PetscInterface.cpp:5-12):
 * a system that will solve on GPU `device`. */
int hfx_la_create(int device, hfx_la** out);
/* Destructor (PetscInterface.cpp:25-31). */
void hfx_la_free(hfx_la* h);

/* LinAlgebraInterface::initialize (PetscInterface.cpp:48-58). */
int hfx_la_initialize(hfx_la* h);

/* LinAlgebraInterface::configure with the PetscOpts fields (PetscInterface.cpp:60-86,
 * PetscOpts.h:9-31): solver_type HFX_LA_GMRES or HFX_LA_BCGS, pc HFX_PC_NONE, HFX_PC_JACOBI,
 * HFX_PC_ILU0_NATURAL (PCILU: point ILU(0) in the natural row order) or
 * HFX_PC_ILU0 (point ILU(0) in a multicolour order; a row without a diagonal
 * entry or a zero pivot fails the solve),
 * rtol (on the true residual ||b - A x||_2 / ||b||_2), maxits, GMRES restart
 * (1..63; PETSc's default is 30).  Fails before initialize. */
int hfx_la_configure(hfx_la* h, int solver_type, int pc, double rtol, int maxits, int restart);

/* LinAlgebraInterface::allocate (PetscInterface.cpp:89-114): ndofs rows;
 * diag_nnz/off_nnz per-row counts (both NULL or both set) only size the
 * value stash.  Fails before initialize + configure.  Resets values. */
int hfx_la_allocate(hfx_la* h, int ndofs, const int* diag_nnz, const int* off_nnz);

/* addValsMatrix / setValsMatrix (PetscInterface.cpp:116-169): the nr x nc
 * row-major block vals at rows x cols, ADD or INSERT.  ADD and INSERT cannot
 * be mixed before a flush.  Entries outside the pattern are created.  The
 * single-value forms are nr = nc = 1. */
int hfx_la_add_vals(hfx_la* h, int nr, const int* rows, int nc, const int* cols, const double* vals);
int hfx_la_set_vals(hfx_la* h, int nr, const int* rows, int nc, const int* cols, const double* vals);

/* addValsRHS / setValsRHS (PetscInterface.cpp:134-187). */
int hfx_la_add_rhs(hfx_la* h, int nr, const int* rows, const double* vals);
int hfx_la_set_rhs(hfx_la* h, int nr, const int* rows, const double* vals);

/* zeroOutRows (PetscInterface.cpp:189-196): pending values are flushed, then
 * the rows' stored values are set to zero (pattern kept, RHS untouched). */
int hfx_la_zero_rows(hfx_la* h, int nr, const int* rows);

/* assemble / assembleFlush (PetscInterface.cpp:198-220). */
int hfx_la_assemble(hfx_la* h);
int hfx_la_assemble_flush(hfx_la* h);

/* solve (PetscInterface.cpp:222-247): FP64 FGMRES on the GPU from a zero
 * guess; x receives ndofs values (n = length of x).  info[0] = iterations,
 * info[1] = converged reason (2 rtol, 3 zero rhs, -3 max its, -9 nan);
 * relres[0] = final true relative residual.  Not converging is not an error
 * (PETSc's KSPSolve default). */
int hfx_la_solve(hfx_la* h, double* x, int n, int* info, double* relres);

/* getSolutionOwnership (PetscInterface.cpp:249-255): [low, high) of the last
 * solve's rows (one GPU: [0, ndofs)). */
int hfx_la_ownership(hfx_la* h, int* low, int* high);

/* clearSystem (PetscInterface.cpp:257-265): values and RHS to zero, pattern kept. */
int hfx_la_clear(hfx_la* h);

/* destroySystem (PetscInterface.cpp:33-46): back to the constructed state. */
int hfx_la_destroy_system(hfx_la* h);

/* getNumDofs (LinAlgebraInterface.h:138). */
int hfx_la_num_dofs(hfx_la* h, int* n);

/* Read back assembled values, row-major nr x nc (absent entries read 0), and
 * RHS entries: the MatGetValues / VecGetValues reads of the reference's tests
 * (tests/unittests/resolution/TestPetscInterface.cpp:31-36). */
int hfx_la_get_values(hfx_la* h, int nr, const int* rows, int nc, const int* cols, double* out);
int hfx_la_get_rhs(hfx_la* h, int nr, const int* rows, double* out);
/* Stored entries of the assembled matrix, and the whole assembled system as
 * CSR (rowptr ndofs+1, columns sorted within a row) with its RHS; any
 * pointer may be NULL. */
int hfx_la_nnz(hfx_la* h, long long* nnz);
int hfx_la_get_csr(hfx_la* h, long long* rowptr, int* cols, double* vals, double* rhs);

/* ------------------------------------------------------- batched assembly */

/* Many independent 2-D problems (straight-sided cells, scalar diffusion,
 * orders 1..5, any mix) assembled and condensed by one persistent launch per
 * order present, the orders running concurrently, then each problem's
 * scatter.  No counterpart in the reference, which assembles one problem at a
 * time on the host (HDGSolver::assemble, HDGSolver.cpp:166-675); each
 * problem's results are exactly those of hfx_hdg_assemble.  The problems stay
 * owned by the caller and must outlive the batch. */
typedef struct hfx_batch hfx_batch;

int hfx_batch_create(int device, hfx_batch** out);
void hfx_batch_free(hfx_batch* b);
/* Add a problem whose mesh and fields are set. */
int hfx_batch_add(hfx_batch* b, hfx_hdg* h);
/* Number of problems and total cells. */
int hfx_batch_size(hfx_batch* b, int* n, long long* cells);
/* Launch shape of order `order` (1..5, or 0 for every order): wavefronts per
 * workgroup (1, 2 or 4) and cells per work item. */
int hfx_batch_set_config(hfx_batch* b, int order, int waves, int tile);
/* out[4*(p-1) + 0..3] = waves, tile, work items, workgroups of order p = 1..5
 * (zero items for an order not in the batch); out[20] bit 0 set when the
 * orders run concurrently on side streams, bit 1 when they do inside the
 * captured graph (21 ints). */
int hfx_batch_get_config(hfx_batch* b, int* out);
/* Autotune (waves, tile) of every order by timing each candidate launch
 * `reps` times on this batch, keeping the fastest per order, then choose
 * concurrent or one-after-another order launches, for eager launches and
 * (timed separately) inside the replayed graph; best_ms returns the time of
 * the tuned eager launch of all orders. */
int hfx_batch_tune(hfx_batch* b, int reps, double* best_ms);
/* Assemble every problem: the per-order launches + the scatters; use_graph
 * != 0 captures the sequence as a HIP graph on first use and replays it. */
int hfx_batch_assemble(hfx_batch* b, int use_graph);
/* HIP-event times (ms, `reps` values): what = 0 batched kernels only,
 * 1 batched kernels + scatters, 2 the same replayed as a graph, 3 the
 * problems one by one with their own kernels. */
int hfx_batch_time(hfx_batch* b, int what, int warmup, int reps, double* ms);

/* ------------------------------------------------ learned preconditioner (lpc) */

/* No counterpart in the reference.  A message-passing network over the
 * face-block graph of the trace matrix (three layers, width 64) emits one
 * correction block per matrix block, added to the face-block Jacobi inverse
 * in the diagonally normalised system; HFX_PC_LEARNED in hfx_hdg_solve then
 * applies M = S^-1/2 (inv(Ahat_FF) + C) S^-1/2 inside FGMRES, which still
 * stops on the true FP64 residual.  Training runs on the GPU (FP32 forward,
 * backward and Adam); inference runs on BF16 MFMA with FP32 accumulation. */
typedef struct hfx_lpc hfx_lpc;

/* A network for face blocks of `block_size` (= nNf) on GPU `device`,
 * parameters all zero (= block Jacobi). */
int hfx_lpc_create(int device, int block_size, hfx_lpc** out);
void hfx_lpc_free(hfx_lpc* h);
int hfx_lpc_num_params(hfx_lpc* h, long long* n);
/* He-uniform weights from a splitmix64 stream, zero biases and decoders
 * (the untrained network is exactly block Jacobi); resets Adam. */
int hfx_lpc_init_params(hfx_lpc* h, unsigned long long seed);
int hfx_lpc_set_params(hfx_lpc* h, const float* p, long long n);
int hfx_lpc_get_params(hfx_lpc* h, float* p, long long n);

/* Add an assembled and solved problem to the training set (its normalised
 * trace matrix and solution are copied on the GPU; the problem may then be
 * re-assembled with other fields).  Samples of one problem handle share a
 * graph; problems on different meshes may be mixed. */
int hfx_lpc_add_sample(hfx_lpc* h, hfx_hdg* prob);
int hfx_lpc_num_samples(hfx_lpc* h, int* n);
int hfx_lpc_clear_samples(hfx_lpc* h);
/* Graph of one sample: out[0..3] = nodes (faces), matrix blocks, edges
 * (off-diagonal blocks), block size. */
int hfx_lpc_graph_sizes(hfx_lpc* h, int sample, long long* out);

/* Loss and gradient (FP32 forward/backward on the GPU) of one sample with the
 * normalised solution probe plus n_random Rademacher probes drawn from
 * `seed`.  loss = sum over probes ||z - Mhat Ahat z||^2 / ||z||^2.  grad (may be
 * NULL) has hfx_lpc_num_params entries, in the parameter layout. */
int hfx_lpc_loss_grad(hfx_lpc* h, int sample, int n_random, unsigned long long seed, double* loss, float* grad);

/* nsteps Adam steps (lr, betas 0.9/0.999, eps 1e-8), step i on sample
 * order[i] with probe seed seed + i; losses[i] = loss before step i. */
int hfx_lpc_train(hfx_lpc* h, const int* order, int nsteps, float lr, int n_random, unsigned long long seed,
                  double* losses);

/* Network outputs of one sample: cn nodes*b*b (diagonal corrections), ce
 * edges*b*b (off-diagonal blocks in BSR order); bf16 != 0 uses the MFMA path. */
int hfx_lpc_outputs(hfx_lpc* h, int sample, int bf16, float* cn, float* ce);

/* Attach a network to a problem for HFX_PC_LEARNED (NULL detaches);
 * bf16 != 0 runs inference on BF16 MFMA, else FP32. */
int hfx_hdg_set_lpc(hfx_hdg* h, hfx_lpc* lpc, int bf16);
/* Host wall time of the last preconditioner set-up of hfx_hdg_solve, ms
 * (learned: normalise + network + M assembly; ILU(0): analysis on a new
 * pattern + factorisation; Jacobi: inverse diagonal). */
int hfx_hdg_pc_setup_ms(hfx_hdg* h, double* ms);

/* ------------------------------------------------------------- Burgers path
 * HDGBurgersModel (src/model/HDGBurgersModel.cpp:34-132) on HDGSolver's
 * IMPLICIT assembly with m = dim unknowns per node, the trace solve, and
 * NonLinearWrapper (src/solver/NonLinearWrapper.cpp:41-79).  Fields, as the
 * reference's Field values: Tau nFaces x nNf x sides x m^2 (tau_ab at
 * b m + a, column-major per node; side s of the face's cell face2cell[2F+s]),
 * DiffusionTensor nNodes scalars (NULL: no diffusion operator,
 * HDGBurgersModel.cpp:99-104), Dirichlet nFaces x nNf x m, source values
 * nCells x nIP x m at the cell IPs (NULL: no source; one Source per component,
 * HDGBurgersModel.cpp:118-128).  State: Solution and BufferSolution
 * nCells x nN x m, Flux nCells x nN x dim x m, Trace nFaces x nNf x m; the
 * model linearises about BufferSolution and Trace.  2-D orders 1-5 and 3-D
 * orders 1-3 (3-D p = 3 keeps each cell's working set in a device
 * workspace: it exceeds the LDS). */
typedef struct hfx_burgers hfx_burgers;
int hfx_burgers_create(int device, int dim, int order, hfx_burgers** out);
void hfx_burgers_destroy(hfx_burgers* h);
/* HDGnGammaModel (src/model/HDGnGammaModel.cpp:5-170): 2-D, m = 2 unknowns
 * (n, Gamma) per node, orders 1-5, on the same handle type: every
 * hfx_burgers_* call below applies, with hfx_ngamma_set_fields in place of
 * hfx_burgers_set_fields. */
int hfx_ngamma_create(int device, int order, hfx_burgers** out);
/* A scalar HDG model at m = dim DOFs per node (HDGModel::allocate with
 * nDOFsPerNode > 1, HDGOperator::multiplyDOFs, HDGOperator.cpp:17-28; the
 * reference's TestHDGLaplaceMulti.cpp): HDGLaplaceModel when the
 * DiffusionTensor passed to hfx_burgers_set_fields is NULL (D = I), the
 * diffusion-source model with that scalar nodal D otherwise, the source (m
 * values per IP) on every component, and the m x m Tau field coupling the
 * components on the faces.  Linear: hfx_burgers_assemble and
 * hfx_burgers_solve give the solution (the state it linearises about is not
 * read).  The same handle type and calls as hfx_burgers_create; 2-D orders
 * 1-5, 3-D orders 1-3. */
int hfx_multi_create(int device, int dim, int order, hfx_burgers** out);
/* HDGnGammaModel::setFieldMap (HDGnGammaModel.cpp:5-44): Tau, Dirichlet and
 * the source values as for Burgers; D (n block) and G (Gamma block) nNodes
 * scalars or NULL (absent); b nNodes x 2 (required); sound_speed
 * (nGammaParams::soundSpeed).  The BufferSolution's n must not vanish. */
int hfx_ngamma_set_fields(hfx_burgers* h, const double* tau, int tau_sides, const double* D, const double* G,
                          const double* b, double sound_speed, const double* dirichlet, const double* f_ip);
/* HDGSolver::setMesh + allocate (HDGSolver.cpp:5-164, block size nNf m);
 * the state fields start at zero. */
int hfx_burgers_set_mesh(hfx_burgers* h, const hfx_mesh* m);
/* HDGBurgersModel::setFieldMap (HDGBurgersModel.cpp:5-32) and the solver's
 * field checks (HDGSolver.cpp:30-78). */
int hfx_burgers_set_fields(hfx_burgers* h, const double* tau, int tau_sides, const double* D, const double* dirichlet,
                           const double* f_ip);
/* HFX_BC_DIRICHLET (DirichletModel) or HFX_BC_INTEGRATED_DIRICHLET
 * (IntegratedDirichletModel, the default; rows (M (x) I_m) and (M (x) I_m) g)
 * on every boundary face. */
int hfx_burgers_set_boundary_model(hfx_burgers* h, int kind);
/* The HDG-type boundary branch (HDGSolver.cpp:503-518) on n boundary faces,
 * with the model's local matrices B (n x bs x bs (2 + dim), bs = nNf m, in
 * face order: columns B_u (pj m + b), B_q ((pj dim + k) m + b), B_lambda
 * (pj m + b)) and rhs r (n x bs); rows Set or Add.  The other boundary faces
 * keep the Dirichlet model; a new call replaces the previous set (n = 0
 * clears it).  hfx_ngamma_set_boundary_bohm builds HDGBohmModel's rows on the GPU. */
int hfx_burgers_set_boundary_hdg(hfx_burgers* h, const int* faces, int n, const double* B, const double* r,
                                 int matrix_set, int rhs_set);
/* HDGBohmModel (src/model/HDGBohmModel.cpp:97-268), the Bohm-Chodura
 * boundary model of the nGamma problem (a handle of hfx_ngamma_create), on
 * the n boundary faces `faces` (n = 0 removes it).  Before every assembly the
 * GPU rebuilds each face's HDG-type rows B = [B_u | B_q | B_lambda] and r from
 * the current Solution, Flux and Trace, the model's b, D, G (both required)
 * and sound speed, the face's Tau (side 0) and the exterior normals
 * `normals` (n x nNf x 2 in face-node order: the model's ExteriorNormals
 * field), and applies them with Set assembly, as the model's AssemblyType
 * says (HDGBohmModel.cpp:89-92).  `transfer` picks the transfer function of
 * the reference's Bohm regression (tests/regression/HDG/TestHDGBohm.cpp:37-130):
 * HFX_BOHM_NEUMANN (0), HFX_BOHM_DIRICHLET (1), HFX_BOHM_HEAVISIDE (the step
 * with |y| tolerance 1e-5 and zero derivative) or HFX_BOHM_TANH (the tanh
 * smoothed step, width 1e-3, with its derivative); a std::function of the
 * reference cannot cross to the device.  Replaces an
 * hfx_burgers_set_boundary_hdg set; the other boundary faces keep the
 * Dirichlet model. */
enum { HFX_BOHM_NEUMANN = 0, HFX_BOHM_DIRICHLET = 1, HFX_BOHM_HEAVISIDE = 2, HFX_BOHM_TANH = 3 };
int hfx_ngamma_set_boundary_bohm(hfx_burgers* h, const int* faces, int n, const double* normals, int transfer);
/* The HDG-type rows of the last assembly: B (n x bs x bs (2 + dim)) and r
 * (n x bs) of the faces of hfx_burgers_set_boundary_hdg or
 * hfx_ngamma_set_boundary_bohm, in their order (either may be NULL). */
int hfx_burgers_boundary_rows(hfx_burgers* h, double* B, double* r);
int hfx_burgers_set_krylov(hfx_burgers* h, int krylov);
/* BiCGStab's steady iterations replayed as a captured HIP graph (same
 * kernels, same results; off by default), as hfx_hdg_set_solver_graphs. */
int hfx_burgers_set_solver_graphs(hfx_burgers* h, int on);
/* Implicit Euler (HFX_TS_BEULER: RungeKutta's BEuler, RungeKutta.cpp:113-159,
 * as in tests/regression/HDG/TestHDGBurgers.cpp) or none: the u rows gain
 * (M (x) I) / dt and F_u gains (M (x) I) OldSolution / dt.  OldSolution starts
 * at zero; hfx_burgers_start_step copies Solution into OldSolution and
 * BufferSolution (the driver's copies at the start of a step). */
int hfx_burgers_set_time_scheme(hfx_burgers* h, int scheme, double dt);
int hfx_burgers_start_step(hfx_burgers* h);
/* New Dirichlet values (nFaces x nNf x m; NULL: zero), e.g. for t_{n+1}. */
int hfx_burgers_set_dirichlet(hfx_burgers* h, const double* dirichlet);
/* Host copies in and out of the state fields (NULL: left as they are). */
int hfx_burgers_set_state(hfx_burgers* h, const double* sol, const double* buffer, const double* trace);
int hfx_burgers_get_state(hfx_burgers* h, double* sol, double* buffer, double* flux, double* trace);
/* out[0..4] = block rows, block size, stored blocks, u rows per cell,
 * trace entries per cell. */
int hfx_burgers_sizes(const hfx_burgers* h, long long* out);
/* HDGSolver::assemble (HDGSolver.cpp:166-675): the cell kernel (local
 * system + condensation) and the trace system with the boundary rows. */
int hfx_burgers_assemble(hfx_burgers* h);
/* Condensed blocks of every cell, row-major, the trace in each cell's
 * face-node order: S, S0, U, Q, U0, Q0 (any may be NULL). */
int hfx_burgers_condensed(hfx_burgers* h, double* S, double* S0, double* U, double* Q, double* U0, double* Q0);
/* The assembled BSR trace matrix (row-major blocks) and right-hand side. */
int hfx_burgers_trace_system(hfx_burgers* h, int* rowptr, int* col, double* vals, double* rhs);
/* HDGSolver::solve (HDGSolver.cpp:677-779): Krylov solve of the trace system
 * from a zero guess into Trace (pc HFX_PC_NONE, _JACOBI, _BJACOBI, _ILU0 or
 * _ILU0_F32), then Solution and Flux reconstructed. */
int hfx_burgers_solve(hfx_burgers* h, int pc, double rtol, int maxits, int restart, int* iters, int* reason,
                      double* relres);
/* NonLinearWrapper::solve: up to max_iters of assemble + solve; residual
 * ||Solution - BufferSolution|| / ||BufferSolution|| (the absolute norm when
 * BufferSolution is zero); below res_tol it stops, else Solution =
 * BufferSolution = (1 - damp) Solution + damp BufferSolution.  Fails when a
 * linear solve does not converge. */
int hfx_burgers_nonlinear_solve(hfx_burgers* h, int max_iters, double res_tol, double damp, int pc, double rtol,
                                int maxits, int restart, int* iters, double* residual, int* linear_iters);

#ifdef __cplusplus
}
#endif
#endif
