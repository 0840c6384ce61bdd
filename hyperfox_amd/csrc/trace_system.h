// Global trace system on the GPU: block-sparse (BSR) layout with one
// nNf x nNf dense block per (face, face) coupling, assembled by a
// deterministic gather from the condensed element blocks, and the
// element-wise reconstruction of u and q from the trace solution.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <vector>

#include "hfx_internal.h"

namespace hfx {

// Host-built BSR pattern of the trace matrix: the block row of face F couples
// with every face of every cell containing F (HDGSolver.cpp:117-164).
struct BsrPattern {
  int nb = 0, bs = 0;                 // block rows (= faces), block size (= nNf)
  std::vector<int> rowptr, col, rowf;  // nb+1, nnzb, nnzb (block row of each block)
  std::vector<int> diag;               // nb: index of the diagonal block
  // contribution of the condensed element blocks to block k:
  //   c[4k+0] = e0, c[4k+1] = a0*16 + b0, c[4k+2] = e1 (or -1), c[4k+3] = a1
  std::vector<int> contrib;
  // per face: the (cell, local face) pairs that hold it: fe[4F..4F+3] = e0 a0 e1 a1
  std::vector<int> face_cells;
};

// Build the pattern from cell2face/face2cell (nF faces per cell).
BsrPattern build_bsr_pattern(int nCells, int nFaces, int nFpc, int bs, const int* cell2face, const int* face2cell);

// Device BSR matrix (row-major dense blocks).
struct BsrDev {
  int nb = 0, bs = 0, nnzb = 0;
  DevBuf<int> rowptr, col, rowf, diag, contrib, face_cells;
  DevBuf<double> vals;
  long long nrows() const { return (long long)nb * bs; }
};

// vals/rhs from condensed S, S0 with Dirichlet rows set to identity/g on the
// faces flagged in is_dir (DirichletModel Set/Set, HDGSolver.cpp:493-502).
// invperm[(e*nFpc + lf)*bs + pos] = cell-local node j at face position pos.
void launch_scatter(const BsrDev& A, int nFpc, const double* S, const double* S0, const int8_t* invperm,
                    const uint8_t* is_dir, const double* g, double* rhs, hipStream_t s);

// Second half of the fused scatter (the element kernel has written every
// off-diagonal block): diagonal block of face F = dside[2F] (+ dside[2F+1]
// when F has a second cell), identity on Dirichlet faces; then the RHS as in
// launch_scatter.  Same values, bit for bit, as launch_scatter on stored S.
void launch_scatter_diag(const BsrDev& A, int nFpc, const double* dside, const double* S0, const int8_t* invperm,
                         const uint8_t* is_dir, const double* g, double* rhs, hipStream_t s);

// IntegratedDirichletModel on the boundary faces bfaces[nbf], after either
// scatter: their diagonal blocks become the face mass of the face's nodes
// (first cell's nodes through fnodes / face_perm, the node-level nF x nNf
// permutation; face tables fphi [nipf][nNf], fdphi [nipf][nNf][dim-1],
// fw [nipf]) and their RHS rows M g.  With m unknowns per node (block size
// nNf m) the block is M (x) I_m (the Mass operator with m DOFs per node).
void launch_integrated_dirichlet(const BsrDev& A, int nbf, const int* bfaces, int dim, int nN, int nipf, int nFpc,
                                 const int* fnodes, const int8_t* face_perm, const int* cells, const double* coords,
                                 const double* fphi, const double* fdphi, const double* fw, const double* g,
                                 double* rhs, hipStream_t s, int m = 1);

// HDG-type boundary model on the boundary faces hfaces[nhf] (not flagged
// Dirichlet, so their rows hold the cell's own S and S0), after either
// scatter: B (nhf x bs x bs(2+dim)) = [B_u | B_q | B_lambda] in face-node
// order and r (nhf x bs); the rows are set (matrix_set / rhs_set) or added.
// bpos: cell_block_positions of the pattern; face_perm / invperm: the
// node-level permutations.  With m unknowns per node (bs = nNf m) the
// columns are B_u (pj m + b), B_q ((pj dim + k) m + b), B_lambda (pj m + b).
void launch_hdg_boundary(const BsrDev& A, int nhf, const int* hfaces, int dim, int nN, int nFpc, const int* fnodes,
                         const int8_t* face_perm, const int8_t* invperm, const int* bpos, const double* U,
                         const double* Q, const double* U0, const double* Q0, const double* B, const double* r,
                         int matrix_set, int rhs_set, double* rhs, hipStream_t s, int m = 1);

// BSR block index of every (face f, face g) pair of every cell:
// out[(e*nFpc + f)*nFpc + g] (the diagonal block when f == g).
std::vector<int> cell_block_positions(const BsrPattern& P, int nCells, int nFpc, const int* cell2face);

// u = U lambda_e + U0, q = Q lambda_e + Q0 (HDGSolver.cpp:741-775).
void launch_reconstruct(int nEl, int nN, int nQ, int nFpc, int bs, const double* U, const double* U0, const double* Q,
                        const double* Q0, const int* cell2face, const int8_t* face_perm, const double* lambda, double* u,
                        double* q, hipStream_t s);

}  // namespace hfx
