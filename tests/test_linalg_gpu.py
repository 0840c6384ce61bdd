"""GPU solves through the LinAlgebraInterface mirror.

* the reference's stock systems (tests/unittests/resolution/TestLinAlgebraInterfaces.cpp:70-169):
  identity, lower-triangular all-ones and hinge matrices of sizes 1..100,
  rtol 1e-16, Jacobi, solutions within 1e-12 of the closed forms;
* the end of its call-ordering section (lines 57-67): solve after assembly,
  then clearSystem;
* an HDG trace system assembled through addValsMatrix (HDGSolver.cpp:580-625)
  solved on the GPU against the oracle's dense LU solution.
"""
import numpy as np
import pytest

import hyperfox_amd as hf
import oracle as orc
from hyperfox_amd import HipLinAlgebra, HipOpts

pytestmark = pytest.mark.gpu


def fresh(la, n):
    la.destroySystem()
    la.initialize()
    la.configure()
    la.allocate(n)


def test_ordering_then_solve_and_clear():
    la = HipLinAlgebra(HipOpts(rtol=1e-16))
    la.initialize()
    la.configure()
    la.allocate(1)
    la.setValsRHS([0], [1.0])
    la.setValRHS(0, 1.0)
    la.setValsMatrix([0], [0], [1.0])
    la.setValMatrix(0, 0, 1.0)
    la.assemble()
    sol = [1.0]
    la.solve(sol)
    assert sol == [1.0]
    assert la.getSolutionOwnership() == [0]
    la.clearSystem()
    assert la.solve() == [0.0]  # zero RHS: zero solution
    assert la.last_info[1] == 3


@pytest.mark.parametrize('solver', [0, 1])
def test_identity_solving(solver):
    la = HipLinAlgebra(HipOpts(solverType=solver, rtol=1e-16))
    for k in range(100):
        fresh(la, k + 1)
        for i in range(k + 1):
            la.setValMatrix(i, i, 1.0)
            la.setValRHS(i, float(i))
        la.assemble()
        sol = la.solve()
        assert np.allclose(sol, np.arange(k + 1.0), rtol=0, atol=1e-12), k


def test_triangular_matrix():
    la = HipLinAlgebra(HipOpts(rtol=1e-16))
    worst = 0.0
    for k in range(100):
        fresh(la, k + 1)
        for i in range(k + 1):
            for j in range(i + 1):
                la.setValMatrix(i, j, 1.0)
            la.setValRHS(i, 1.0)
        la.assemble()
        sol = np.array(la.solve())
        ana = np.zeros(k + 1)
        ana[0] = 1.0
        worst = max(worst, np.abs(sol - ana).max())
        assert np.allclose(sol, ana, rtol=0, atol=1e-12), (k, la.last_info)
    print('triangular: worst error %.2e' % worst)


@pytest.mark.parametrize('solver', [0, 1])
def test_hinge_matrix(solver):
    la = HipLinAlgebra(HipOpts(solverType=solver, rtol=1e-16))
    for k in range(100):
        fresh(la, k + 1)
        for i in range(k + 1):
            la.setValMatrix(i, i, 1.0)
            if i:
                la.setValMatrix(i, i - 1, 1.0)
            la.setValRHS(i, 1.0)
        la.assemble()
        sol = np.array(la.solve())
        ana = np.array([1.0 if i % 2 == 0 else 0.0 for i in range(k + 1)])
        assert np.allclose(sol, ana, rtol=0, atol=1e-12), (k, la.last_info)


@pytest.mark.parametrize('solver', [0, 1])
@pytest.mark.parametrize('dim,p,n', [(2, 3, 8), (3, 2, 3)])
def test_hdg_trace_system_through_the_interface(dim, p, n, solver):
    m = hf.Mesh.structured(dim, p, n, jitter=0.1, seed=11)
    a = m.arrays()
    s = m.sizes
    nNf, nF = s['nNf'], s['nF']
    tau = np.ones((m.n_faces, nNf))
    g = np.linspace(-1.0, 2.0, m.n_faces * nNf).reshape(m.n_faces, nNf)
    sysm = orc.System(dim, p, 0, a, tau, dirichlet=g)
    cond = sysm.condensed()
    la = HipLinAlgebra(HipOpts(solverType=solver, rtol=1e-13, maxits=20000))
    la.initialize()
    la.configure()
    la.allocate(m.n_trace)
    for e in range(m.n_cells):
        rows = np.array([a['cell2face'][e, f] * nNf + a['face_perm'][e, f, j] for f in range(nF) for j in range(nNf)])
        la.addValsMatrix(rows, rows, cond['S'][e])
        la.addValsRHS(rows, cond['S0'][e])
    la.assemble()
    lam = np.array(la.solve())
    its, reason, rel = la.last_info
    assert reason == 2 and rel <= 1e-13
    lam_ref, _, conv, _ = sysm.solve(method=1)
    assert conv
    err = np.abs(lam - lam_ref).max() / np.abs(lam_ref).max()
    print('dim %d p %d: %d iterations, rel err %.2e' % (dim, p, its, err))
    assert err < 1e-10
