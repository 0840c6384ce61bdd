"""FP64 numpy oracle of the block ILU(0) trace preconditioner.

TEST INFRASTRUCTURE ONLY: the checker for hyperfox_amd/csrc/ilu.hip and the
host analysis in hyperfox_amd/csrc/ilu_plan.cpp.  Only the tests load it.

The reference selects its preconditioner through PETSc
(PetscOpts::preconditionnerType, src/resolution/PetscOpts.h:16, applied by
PCSetType in src/resolution/PetscInterface.cpp:70); PCILU with zero levels of
fill is the incomplete LU factorisation restated here.  On a matrix whose
pattern is made of dense b x b blocks, point ILU(0) and block ILU(0) on the
block pattern coincide, so the factors are kept per block:

* elimination order: a rank per block row -- the natural order (rank = row,
  PETSc's default for sequential AIJ) or a greedy multicolour order (rows
  sorted by colour, then row; rows of one colour share no block);
* factors (textbook IKJ form, in rank order): for each row i and each block
  (i, k) with rank k < rank i, taken by increasing rank k,
      L_ik = A'_ik U_kk^-1,   A'_ij -= L_ik U_kj  for every (i, j) and (k, j)
  in the pattern with rank j > rank k; U_ij = A'_ij for rank j >= rank i;
* apply M^-1 x: forward y_i = x_i - sum_{rank k < rank i} L_ik y_k, then
  backward z_i = U_ii^-1 (y_i - sum_{rank j > rank i} U_ij z_j).

The GPU stores Dinv_i = U_ii^-1 and Uhat_ij = U_ii^-1 U_ij (so that the
backward sweep reads one array per row); `factor` returns those too.

Pinned in tests/test_ilu_oracle.py by the defining identity of ILU(0)
((L U)_ij = A_ij on every pattern block), by exactness where ILU(0) has no
dropped fill (dense pattern; block-tridiagonal matrix in natural order), by
permutation equivariance (ILU(0) in rank order = natural-order ILU(0) of the
permuted matrix), and by the colouring's properties.
"""
import numpy as np


def row_cols(rowptr, col, i):
    return col[rowptr[i]:rowptr[i + 1]]


def symmetric_adjacency(nb, rowptr, col):
    adj = [set() for _ in range(nb)]
    for i in range(nb):
        for j in row_cols(rowptr, col, i):
            j = int(j)
            if j != i:
                adj[i].add(j)
                adj[j].add(i)
    return adj


def greedy_colours(nb, rowptr, col):
    """Smallest colour not used by an already coloured neighbour, rows in
    natural order, on the symmetrised block graph."""
    adj = symmetric_adjacency(nb, rowptr, col)
    colour = -np.ones(nb, dtype=np.int64)
    for i in range(nb):
        used = {int(colour[j]) for j in adj[i] if colour[j] >= 0}
        c = 0
        while c in used:
            c += 1
        colour[i] = c
    return colour


def ranks(nb, rowptr, col, ordering):
    """rank[i] = position of block row i in the elimination order;
    ordering 'natural' or 'multicolour'."""
    if ordering == 'natural':
        return np.arange(nb, dtype=np.int64)
    colour = greedy_colours(nb, rowptr, col)
    order = np.lexsort((np.arange(nb), colour))
    rank = np.empty(nb, dtype=np.int64)
    rank[order] = np.arange(nb)
    return rank


def levels(nb, rowptr, col, rank):
    """Forward and backward dependency levels of the triangular sweeps."""
    order = np.argsort(rank)
    lf = np.zeros(nb, dtype=np.int64)
    for i in order:
        deps = [lf[k] + 1 for k in row_cols(rowptr, col, i) if rank[k] < rank[i]]
        lf[i] = max(deps) if deps else 0
    lb = np.zeros(nb, dtype=np.int64)
    for i in order[::-1]:
        deps = [lb[j] + 1 for j in row_cols(rowptr, col, i) if rank[j] > rank[i]]
        lb[i] = max(deps) if deps else 0
    return lf, lb


def factor(nb, bs, rowptr, col, vals, rank):
    """Block ILU(0).  vals: (nnzb, bs, bs) row-major blocks.  Returns
    (F, Dinv) where F holds L_ik on the strictly lower blocks (rank k <
    rank i), U_ii on the diagonal and Uhat_ij = U_ii^-1 U_ij on the strictly
    upper blocks, and Dinv[i] = U_ii^-1.  Raises on a singular pivot block."""
    A = np.array(vals, dtype=np.float64, copy=True).reshape(-1, bs, bs)
    pos = [dict() for _ in range(nb)]
    for i in range(nb):
        for p in range(rowptr[i], rowptr[i + 1]):
            pos[i][int(col[p])] = p
    Uinv = np.zeros((nb, bs, bs))
    for i in np.argsort(rank):
        lower = sorted((k for k in pos[i] if rank[k] < rank[i]), key=lambda k: rank[k])
        for k in lower:
            A[pos[i][k]] = A[pos[i][k]] @ Uinv[k]                     # L_ik
            for j, pkj in pos[k].items():
                if rank[j] > rank[k] and j in pos[i]:
                    A[pos[i][j]] -= A[pos[i][k]] @ A[pkj]              # U_kj is final
        if i not in pos[i]:
            raise ValueError('ILU(0): block row %d has no diagonal block' % i)
        d = A[pos[i][i]]
        if np.linalg.cond(d) > 1e15:
            raise ValueError('ILU(0): singular pivot block in row %d' % i)
        Uinv[i] = np.linalg.inv(d)
    F = A.copy()
    for i in range(nb):
        for j, p in pos[i].items():
            if rank[j] > rank[i]:
                F[p] = Uinv[i] @ A[p]
    return F, Uinv


def apply(nb, bs, rowptr, col, F, Dinv, rank, x):
    """z = (L U)^-1 x with the factors of `factor`."""
    x = np.asarray(x, dtype=np.float64).reshape(nb, bs)
    order = np.argsort(rank)
    y = np.zeros_like(x)
    for i in order:
        acc = x[i].copy()
        for p in range(rowptr[i], rowptr[i + 1]):
            k = col[p]
            if rank[k] < rank[i]:
                acc -= F[p] @ y[k]
        y[i] = acc
    z = np.zeros_like(x)
    for i in order[::-1]:
        acc = Dinv[i] @ y[i]
        for p in range(rowptr[i], rowptr[i + 1]):
            j = col[p]
            if rank[j] > rank[i]:
                acc -= F[p] @ z[j]
        z[i] = acc
    return z.reshape(-1)


def dense_factors(nb, bs, rowptr, col, F, Dinv, rank):
    """Dense L (unit block diagonal) and U of the factorisation, in the
    original row numbering: L U equals A on every pattern block."""
    n = nb * bs
    L = np.eye(n)
    U = np.zeros((n, n))
    for i in range(nb):
        Uii = np.linalg.inv(Dinv[i])
        for p in range(rowptr[i], rowptr[i + 1]):
            j = col[p]
            blk = (slice(i * bs, (i + 1) * bs), slice(j * bs, (j + 1) * bs))
            if rank[j] < rank[i]:
                L[blk] = F[p]
            elif j == i:
                U[blk] = Uii
            else:
                U[blk] = Uii @ F[p]
    return L, U


def dense(nb, bs, rowptr, col, vals):
    A = np.zeros((nb * bs, nb * bs))
    v = np.asarray(vals).reshape(-1, bs, bs)
    for i in range(nb):
        for p in range(rowptr[i], rowptr[i + 1]):
            j = col[p]
            A[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs] = v[p]
    return A
