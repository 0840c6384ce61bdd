"""Block ILU(0): the numpy oracle pinned by what the mathematics fixes, and
the library's host analysis (elimination order, colours, sweep levels)
against it.  No GPU.

Pins of the oracle's factorisation:
* (L U)_ij = A_ij on every pattern block -- the defining identity of ILU(0);
* exactness where no fill is dropped: a dense block pattern, and a block
  tridiagonal matrix in natural order (M^-1 = A^-1);
* the pivots of the 5-point Laplacian in natural order follow the textbook
  recurrence d_i = 4 - 1/d_(i-1) - 1/d_(i-m) (terms present where the
  neighbour exists);
* ILU(0) in a rank order equals natural-order ILU(0) of the permuted matrix.
"""
import numpy as np
import pytest

import hyperfox_amd as hf
import ilu_oracle as io


def csr_from_adj(adj):
    rowptr, col = [0], []
    for i, nb in enumerate(adj):
        cols = sorted(set(nb) | {i})
        col += cols
        rowptr.append(len(col))
    return np.array(rowptr), np.array(col)


def grid_pattern(m):
    adj = []
    for i in range(m * m):
        x, y = i % m, i // m
        nb = [j for j, ok in ((i - 1, x > 0), (i + 1, x < m - 1), (i - m, y > 0), (i + m, y < m - 1)) if ok]
        adj.append(nb)
    return csr_from_adj(adj)


def random_pattern(nb, deg, seed):
    rng = np.random.default_rng(seed)
    adj = [set() for _ in range(nb)]
    for i in range(nb):
        for j in rng.choice(nb, size=deg, replace=False):
            if j != i:
                adj[i].add(int(j))
                adj[int(j)].add(i)
    return csr_from_adj([sorted(a) for a in adj])


def chain_pattern(nb):
    return csr_from_adj([[j for j in (i - 1, i + 1) if 0 <= j < nb] for i in range(nb)])


def dense_pattern(nb):
    return csr_from_adj([list(range(nb)) for _ in range(nb)])


def random_values(rowptr, col, bs, seed, dom=4.0):
    rng = np.random.default_rng(seed)
    v = rng.standard_normal((col.size, bs, bs))
    for i in range(rowptr.size - 1):
        for p in range(rowptr[i], rowptr[i + 1]):
            if col[p] == i:
                v[p] += dom * (rowptr[i + 1] - rowptr[i]) * np.eye(bs)
    return v


def pattern_mask(nb, bs, rowptr, col):
    m = np.zeros((nb * bs, nb * bs), dtype=bool)
    for i in range(nb):
        for p in range(rowptr[i], rowptr[i + 1]):
            j = col[p]
            m[i * bs:(i + 1) * bs, j * bs:(j + 1) * bs] = True
    return m


@pytest.mark.parametrize('ordering', ['natural', 'multicolour'])
@pytest.mark.parametrize('bs', [1, 3, 4])
@pytest.mark.parametrize('kind', ['grid', 'random'])
def test_lu_equals_a_on_the_pattern(ordering, bs, kind):
    rowptr, col = grid_pattern(6) if kind == 'grid' else random_pattern(30, 3, 5)
    nb = rowptr.size - 1
    vals = random_values(rowptr, col, bs, 11)
    rank = io.ranks(nb, rowptr, col, ordering)
    F, Dinv = io.factor(nb, bs, rowptr, col, vals, rank)
    L, U = io.dense_factors(nb, bs, rowptr, col, F, Dinv, rank)
    A = io.dense(nb, bs, rowptr, col, vals)
    mask = pattern_mask(nb, bs, rowptr, col)
    LU = L @ U
    assert np.abs(LU - A)[mask].max() <= 1e-12 * np.abs(A).max()
    # fill outside the pattern was dropped, so the factorisation is incomplete
    assert np.abs(LU - A)[~mask].max() > 1e-6
    # the apply inverts L U
    x = np.random.default_rng(3).standard_normal(nb * bs)
    z = io.apply(nb, bs, rowptr, col, F, Dinv, rank, x)
    assert np.allclose(LU @ z, x, rtol=0, atol=1e-11 * np.abs(x).max())


@pytest.mark.parametrize('ordering', ['natural', 'multicolour'])
def test_dense_pattern_is_exact_lu(ordering):
    nb, bs = 6, 3
    rowptr, col = dense_pattern(nb)
    vals = random_values(rowptr, col, bs, 2)
    rank = io.ranks(nb, rowptr, col, ordering)
    F, Dinv = io.factor(nb, bs, rowptr, col, vals, rank)
    x = np.random.default_rng(4).standard_normal(nb * bs)
    z = io.apply(nb, bs, rowptr, col, F, Dinv, rank, x)
    A = io.dense(nb, bs, rowptr, col, vals)
    assert np.allclose(z, np.linalg.solve(A, x), rtol=1e-12, atol=1e-13)


def test_block_tridiagonal_natural_is_exact_lu():
    nb, bs = 12, 4
    rowptr, col = chain_pattern(nb)
    vals = random_values(rowptr, col, bs, 8, dom=1.0)
    rank = io.ranks(nb, rowptr, col, 'natural')
    F, Dinv = io.factor(nb, bs, rowptr, col, vals, rank)
    x = np.random.default_rng(5).standard_normal(nb * bs)
    z = io.apply(nb, bs, rowptr, col, F, Dinv, rank, x)
    A = io.dense(nb, bs, rowptr, col, vals)
    assert np.allclose(z, np.linalg.solve(A, x), rtol=1e-11, atol=1e-12)


def test_laplacian_pivots_follow_the_textbook_recurrence():
    m = 7
    rowptr, col = grid_pattern(m)
    nb = m * m
    vals = np.array([4.0 if col[p] == i else -1.0 for i in range(nb) for p in range(rowptr[i], rowptr[i + 1])])
    rank = io.ranks(nb, rowptr, col, 'natural')
    F, Dinv = io.factor(nb, 1, rowptr, col, vals.reshape(-1, 1, 1), rank)
    d = np.zeros(nb)
    for i in range(nb):
        d[i] = 4.0 - (1.0 / d[i - 1] if i % m else 0.0) - (1.0 / d[i - m] if i >= m else 0.0)
    assert np.allclose(1.0 / Dinv[:, 0, 0], d, rtol=1e-14)


def test_rank_order_equals_natural_order_of_the_permuted_matrix():
    rowptr, col = random_pattern(25, 3, 9)
    nb, bs = rowptr.size - 1, 2
    vals = random_values(rowptr, col, bs, 12)
    rank = io.ranks(nb, rowptr, col, 'multicolour')
    assert not np.array_equal(rank, np.arange(nb))
    F, Dinv = io.factor(nb, bs, rowptr, col, vals, rank)
    # permuted matrix: row rank[i] holds row i
    inv = np.argsort(rank)
    prow, pcol, pval = [0], [], []
    for r in range(nb):
        i = inv[r]
        blocks = sorted((rank[col[p]], p) for p in range(rowptr[i], rowptr[i + 1]))
        pcol += [c for c, _ in blocks]
        pval += [vals[p] for _, p in blocks]
        prow.append(len(pcol))
    prow, pcol, pval = np.array(prow), np.array(pcol), np.array(pval)
    Fp, Dp = io.factor(nb, bs, prow, pcol, pval, np.arange(nb))
    for r in range(nb):
        i = inv[r]
        assert np.allclose(Dp[r], Dinv[i], rtol=1e-13, atol=1e-15)
        for q in range(prow[r], prow[r + 1]):
            j = inv[pcol[q]]
            p = rowptr[i] + list(col[rowptr[i]:rowptr[i + 1]]).index(j)
            assert np.allclose(Fp[q], F[p], rtol=1e-12, atol=1e-14)


@pytest.mark.parametrize('kind', ['grid', 'random'])
def test_colouring_and_levels(kind):
    rowptr, col = grid_pattern(8) if kind == 'grid' else random_pattern(40, 3, 6)
    nb = rowptr.size - 1
    colour = io.greedy_colours(nb, rowptr, col)
    adj = io.symmetric_adjacency(nb, rowptr, col)
    assert all(colour[i] != colour[j] for i in range(nb) for j in adj[i])
    assert colour.max() + 1 <= max(len(a) for a in adj) + 1
    for ordering in ('natural', 'multicolour'):
        rank = io.ranks(nb, rowptr, col, ordering)
        lf, lb = io.levels(nb, rowptr, col, rank)
        for i in range(nb):
            for j in col[rowptr[i]:rowptr[i + 1]]:
                if rank[j] < rank[i]:
                    assert lf[i] > lf[j]
                if rank[j] > rank[i]:
                    assert lb[i] > lb[j]
        if ordering == 'multicolour':
            assert lf.max() + 1 <= colour.max() + 1 and lb.max() + 1 <= colour.max() + 1


@pytest.mark.parametrize('ordering', [0, 1])
@pytest.mark.parametrize('kind', ['grid', 'random', 'chain'])
def test_host_analysis_matches_oracle(ordering, kind):
    rowptr, col = {'grid': lambda: grid_pattern(9), 'random': lambda: random_pattern(60, 3, 7),
                   'chain': lambda: chain_pattern(20)}[kind]()
    nb = rowptr.size - 1
    # shuffle columns within rows: the analysis must not assume sorted rows
    rng = np.random.default_rng(1)
    col = col.copy()
    for i in range(nb):
        seg = col[rowptr[i]:rowptr[i + 1]]
        col[rowptr[i]:rowptr[i + 1]] = seg[rng.permutation(seg.size)]
    got = hf.ilu_analyse(rowptr, col, ordering)
    name = 'multicolour' if ordering == 0 else 'natural'
    rank = io.ranks(nb, rowptr, col, name)
    lf, lb = io.levels(nb, rowptr, col, rank)
    assert np.array_equal(got['rank'], rank)
    assert np.array_equal(got['levf'], lf) and np.array_equal(got['levb'], lb)
    assert got['nlev_f'] == lf.max() + 1 and got['nlev_b'] == lb.max() + 1
    if ordering == 0:
        assert np.array_equal(got['colour'], io.greedy_colours(nb, rowptr, col))
        assert got['ncolours'] == io.greedy_colours(nb, rowptr, col).max() + 1


def test_host_analysis_rejects_a_missing_diagonal_block():
    rowptr = np.array([0, 2, 3])
    col = np.array([0, 1, 0])
    with pytest.raises(hf.HfxError, match='diagonal'):
        hf.ilu_analyse(rowptr, col, 0)
