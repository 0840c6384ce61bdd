// Host analysis for the block ILU(0) preconditioner of the BSR trace matrix
// (and of the scalar CSR of the LinAlgebraInterface path, block size 1).
//
// The reference picks its preconditioner through PETSc
// (PetscOpts::preconditionnerType, src/resolution/PetscOpts.h:16; PCSetType in
// src/resolution/PetscInterface.cpp:70).  PCILU with zero fill levels on a
// matrix made of dense b x b blocks is block ILU(0) on the block pattern,
// which is what this computes:
//   for each block row i in elimination order, for each block (i, k) whose
//   column comes earlier, by increasing position of k:
//     A'_ij -= A'_ik Uhat_kj  for every (i, j), (k, j) in the pattern with j
//                             later than k,
//     L_ik   = A'_ik Dinv_k,
//   then Dinv_i = (A'_ii)^-1 and Uhat_ij = Dinv_i A'_ij for j later than i.
// Applying M^-1: y_i = x_i - sum L_ik y_k (forward, in elimination order), then
// z_i = Dinv_i y_i - sum Uhat_ij z_j (backward).
//
// Both sweeps and the factorisation run level by level: a level is a set of
// block rows with no dependency among them (host analysis of the pattern).
// The elimination order is either the natural row order (PETSc's default) or
// a greedy multicolour order (rows sorted by colour; rows of one colour share
// no block), which cuts the number of levels to the number of colours.
#pragma once

#include <vector>

namespace hfx {

enum IluOrdering { ILU_MULTICOLOUR = 0, ILU_NATURAL = 1 };

// Host analysis of a block pattern (rowptr/col, any column order within a
// row) for ILU(0) in the given elimination order.  Requires every diagonal
// block.
struct IluPlan {
  int nb = 0, ncolours = 0, nlev_f = 0, nlev_b = 0, maxlen = 0;
  std::vector<int> rank;               // nb: position of each block row in the elimination order
  std::vector<int> colour;             // nb (multicolour order; zeros for natural)
  std::vector<int> levf, levb;         // nb: forward / backward level of each row
  std::vector<int> fwd_rows, fwd_off;  // rows grouped by forward level (nlev_f + 1 offsets)
  std::vector<int> bwd_rows, bwd_off;  // rows grouped by backward level
  std::vector<int> lptr, lpos, lcol;   // per row: earlier-column blocks by increasing rank of the column
  std::vector<int> uptr, upos, ucol;   // per row: later-column blocks
  std::vector<int> dslot;              // per row: diagonal block position - rowptr[i]
  std::vector<int> pptr, pslot, ppos;  // per earlier block (i, k): pairs (slot of j in row i, position of (k, j))
  std::vector<signed char> kind;       // per block: -1 earlier column, 0 diagonal, 1 later column
};

IluPlan ilu_analyse(int nb, const int* rowptr, const int* col, int ordering);

}  // namespace hfx
