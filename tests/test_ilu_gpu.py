"""Block ILU(0) trace preconditioner on the GPU against its FP64 numpy oracle
(itself pinned in test_ilu_oracle.py):

* the factors (L_ik, Dinv_i, Dinv_i U_ij) of real trace matrices -- 2-D p=1,
  3, 5 and 3-D p=3 (face blocks 2, 4, 6, 10) -- in both elimination orders;
* one application of M^-1;
* FGMRES and BiCGStab with ILU(0) reach the same true FP64 residual and the
  same solution as block Jacobi, in fewer iterations;
* point ILU(0) behind the LinAlgebraInterface mirror (block size 1), on the
  same trace system.
"""
import numpy as np
import pytest

import hyperfox_amd as hf
import hyperfox_amd.problems as pr
import ilu_oracle as io

pytestmark = pytest.mark.gpu

CASES = [(2, 1, 8), (2, 3, 8), (2, 5, 5), (3, 3, 2)]


def assembled(dim, p, n):
    if dim == 2:
        w = pr.diffusion_grf_2d(n, p, 77)
    else:
        w = pr.convection_diffusion_3d(n, p)
    prob = hf.HDGProblem(dim, p)
    pr.setup_gpu(w, prob)
    prob.assemble()
    return w, prob


def block_system(prob):
    rowptr, cols, vals, rhs = prob.trace_csr()
    b = prob.sizes['nNf']
    nb = (rowptr.size - 1) // b
    brow, bcol, bval = [0], [], []
    for F in range(nb):
        r0, r1 = rowptr[F * b], rowptr[F * b + 1]
        nblk = (r1 - r0) // b
        bcol += list(cols[r0:r1:b] // b)
        rows = np.stack([vals[rowptr[F * b + r]:rowptr[F * b + r + 1]] for r in range(b)])
        bval.append(rows.reshape(b, nblk, b).transpose(1, 0, 2))
        brow.append(len(bcol))
    return nb, b, np.array(brow), np.array(bcol), np.concatenate(bval), rhs


@pytest.fixture(scope='module', params=CASES, ids=['%dd-p%d-n%d' % c for c in CASES])
def case(request):
    dim, p, n = request.param
    w, prob = assembled(dim, p, n)
    return dict(prob=prob, sys=block_system(prob))


@pytest.mark.parametrize('ordering', [0, 1], ids=['multicolour', 'natural'])
def test_factors_match_oracle(case, ordering):
    prob = case['prob']
    nb, b, rowptr, col, vals, _ = case['sys']
    lu, dinv, info = prob.ilu_factors(ordering)
    rank = io.ranks(nb, rowptr, col, 'multicolour' if ordering == 0 else 'natural')
    F, Dinv = io.factor(nb, b, rowptr, col, vals, rank)
    plan = hf.ilu_analyse(rowptr, col, ordering)
    assert info == (plan['ncolours'], plan['nlev_f'], plan['nlev_b'])
    e_lu = np.abs(lu - F).max() / np.abs(F).max()
    e_d = np.abs(dinv - Dinv).max() / np.abs(Dinv).max()
    print('nb %d b %d ordering %d: colours %d levels %d/%d, max rel err lu %.2e dinv %.2e' % (
        nb, b, ordering, info[0], info[1], info[2], e_lu, e_d))
    assert e_lu < 1e-11 and e_d < 1e-11


@pytest.mark.parametrize('pc', [hf.PC_ILU0, hf.PC_ILU0_NATURAL], ids=['multicolour', 'natural'])
def test_apply_matches_oracle(case, pc):
    prob = case['prob']
    nb, b, rowptr, col, vals, _ = case['sys']
    x = np.random.default_rng(12).standard_normal(nb * b)
    y = prob.pc_apply(pc, x)
    rank = io.ranks(nb, rowptr, col, 'multicolour' if pc == hf.PC_ILU0 else 'natural')
    F, Dinv = io.factor(nb, b, rowptr, col, vals, rank)
    z = io.apply(nb, b, rowptr, col, F, Dinv, rank, x)
    assert np.abs(y - z).max() <= 1e-11 * np.abs(z).max()


def test_float_factor_apply_is_the_oracle_apply_on_float_factors(case):
    """HFX_PC_ILU0_F32 reads the factors rounded to float and sums in double:
    it equals the FP64 oracle's apply on those rounded factors up to the
    summation order, and differs from the FP64 apply by the rounding only."""
    prob = case['prob']
    nb, b, rowptr, col, vals, _ = case['sys']
    x = np.random.default_rng(13).standard_normal(nb * b)
    y32 = prob.pc_apply(hf.PC_ILU0_F32, x)
    y64 = prob.pc_apply(hf.PC_ILU0, x)
    rank = io.ranks(nb, rowptr, col, 'multicolour')
    F, Dinv = io.factor(nb, b, rowptr, col, vals, rank)
    z = io.apply(nb, b, rowptr, col, F.astype(np.float32).astype(np.float64),
                 Dinv.astype(np.float32).astype(np.float64), rank, x)
    assert np.abs(y32 - z).max() <= 1e-11 * np.abs(z).max()
    d = np.abs(y32 - y64).max() / np.abs(y64).max()
    assert 0 < d < 1e-4, d


@pytest.mark.parametrize('krylov', [hf.KRYLOV_FGMRES, hf.KRYLOV_BICGSTAB], ids=['fgmres', 'bicgstab'])
def test_float_factor_solve_reaches_the_fp64_residual(krylov):
    w, prob = assembled(2, 3, 16)
    prob.set_krylov(krylov)
    it64, r64, _ = prob.solve(pc=hf.PC_ILU0, rtol=1e-12, maxits=20000)
    lam64 = prob.solution()[2].copy()
    it32, r32, rel32 = prob.solve(pc=hf.PC_ILU0_F32, rtol=1e-12, maxits=20000)
    lam32 = prob.solution()[2]
    print('krylov %d: ILU(0) %d its, FP32 factors %d its' % (krylov, it64, it32))
    assert r64 == 2 and r32 == 2 and rel32 <= 1e-12
    assert np.abs(lam32 - lam64).max() <= 1e-9 * np.abs(lam64).max()
    assert it32 <= 1.2 * it64 + 2


@pytest.mark.parametrize('krylov', [hf.KRYLOV_FGMRES, hf.KRYLOV_BICGSTAB], ids=['fgmres', 'bicgstab'])
@pytest.mark.parametrize('pc', [hf.PC_ILU0, hf.PC_ILU0_NATURAL], ids=['multicolour', 'natural'])
def test_ilu_solve_matches_block_jacobi(krylov, pc):
    w, prob = assembled(2, 3, 16)
    prob.set_krylov(krylov)
    it_b, r_b, rel_b = prob.solve(pc=hf.PC_BJACOBI, rtol=1e-12, maxits=20000)
    lam_b = prob.solution()[2].copy()
    it_i, r_i, rel_i = prob.solve(pc=pc, rtol=1e-12, maxits=20000)
    lam_i = prob.solution()[2]
    print('krylov %d pc %d: block Jacobi %d its, ILU(0) %d its' % (krylov, pc, it_b, it_i))
    assert r_b == 2 and r_i == 2 and rel_i <= 1e-12
    assert np.abs(lam_i - lam_b).max() <= 1e-9 * np.abs(lam_b).max()
    assert it_i < it_b


@pytest.mark.parametrize('pc', [hf.linalg.PC_ILU0, hf.linalg.PC_ILU0_NATURAL], ids=['multicolour', 'natural'])
def test_point_ilu_behind_the_linalg_interface(pc):
    w, prob = assembled(2, 2, 6)
    rowptr, cols, vals, rhs = prob.trace_csr()
    n = rowptr.size - 1
    la = hf.HipLinAlgebra(hf.HipOpts(preconditionnerType=pc, rtol=1e-12, maxits=5000))
    la.initialize()
    la.configure()
    la.allocate(n)
    for r in range(n):
        js = cols[rowptr[r]:rowptr[r + 1]]
        la.addValsMatrix([r], list(js), vals[rowptr[r]:rowptr[r + 1]])
    la.addValsRHS(list(range(n)), rhs)
    la.assemble()
    x = np.array(la.solve())
    A = np.zeros((n, n))
    for r in range(n):
        A[r, cols[rowptr[r]:rowptr[r + 1]]] = vals[rowptr[r]:rowptr[r + 1]]
    ref = np.linalg.solve(A, rhs)
    assert np.abs(x - ref).max() <= 1e-9 * np.abs(ref).max()
    its, reason, rel = la.last_info
    assert reason == 2 and rel <= 1e-12


@pytest.mark.parametrize('pc', [hf.PC_JACOBI, hf.PC_BJACOBI, hf.PC_ILU0, hf.PC_ILU0_F32],
                         ids=['jacobi', 'block-jacobi', 'ilu0', 'ilu0-f32'])
def test_bicgstab_graph_replay_gives_the_launch_by_launch_result(pc):
    """The captured graph replays the same kernels in the same order: the same
    iteration count and the same solution bit for bit, over several replays
    and a restart."""
    w, prob = assembled(2, 3, 16)
    prob.set_krylov(hf.KRYLOV_BICGSTAB)
    prob.set_solver_graphs(False)
    it0, r0, rel0 = prob.solve(pc=pc, rtol=1e-12, maxits=20000)
    lam0 = prob.solution()[2].copy()
    prob.set_solver_graphs(True)
    for rep in range(2):
        it1, r1, rel1 = prob.solve(pc=pc, rtol=1e-12, maxits=20000)
        lam1 = prob.solution()[2]
        assert (it1, r1) == (it0, r0) and rel1 == rel0, (rep, it0, it1)
        assert np.array_equal(lam1, lam0), rep
    assert r0 == 2 and it0 > 16
