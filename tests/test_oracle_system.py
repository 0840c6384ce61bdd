"""Whole-mesh oracle (CPU): the reference's solver-level known answers.

* constant Dirichlet data gives u = 3, q = 0, lambda = 3 exactly
  (tests/unittests/solver/TestHDGSolver.cpp:26-27, 86-100) on lightTri2.h5;
* HDG Laplace regression l2Err < 1e-2 on the reference's meshes
  (tests/regression/HDG/TestHDGLaplace.cpp:110-150);
* manufactured-solution convergence at rate ~p+1 on structured meshes;
* GMRES + Jacobi on the LinAlgebraInterface systems
  (tests/unittests/resolution/TestLinAlgebraInterfaces.cpp:70-169)."""
import os

import numpy as np
import pytest

import hyperfox_amd as hf
import oracle as orc
from hdg_helpers import l2_projection_error

from conftest import GOLDEN


def load_fixture(name):
    z = np.load(os.path.join(GOLDEN, 'meshes', name + '.npz'))
    return z['nodes'], z['cells']


def fixture_mesh(name, dim, p):
    nodes, cells = load_fixture(name)
    return hf.Mesh.from_cells(dim, p, nodes, cells)


@pytest.mark.parametrize('method', [0, 1])
def test_constant_dirichlet_light_tri2(method):
    m = fixture_mesh('lightTri2', 2, 2)
    a = m.arrays()
    nNf = m.sizes['nNf']
    tau = np.ones((m.n_faces, nNf))
    g = np.full((m.n_faces, nNf), 3.0)
    s = orc.System(2, 2, 0, a, tau, dirichlet=g)
    lam, its, conv, relres = s.solve(method=method, rtol=1e-12)
    u, q = s.reconstruct(lam)
    assert np.allclose(u, 3.0, atol=1e-12, rtol=0)
    assert np.allclose(q, 0.0, atol=1e-12, rtol=0)
    assert np.allclose(lam, 3.0, atol=1e-12, rtol=0)


def sin_exp(x):
    return np.sin(x[..., 0]) * np.exp(x[..., 1])


LAPLACE_CASES = [(2, h, p) for h in ('3e-1', '2e-1', '1e-1') for p in (1, 2, 3)] + [(3, '2e-1', p) for p in (1, 2, 3)]


@pytest.mark.parametrize('dim,h,p', LAPLACE_CASES)
def test_hdg_laplace_regression(dim, h, p):
    m = fixture_mesh('regression_dim-%d_h-%s_ord-%d' % (dim, h, p), dim, p)
    a = m.arrays()
    nNf = m.sizes['nNf']
    tau = np.ones((m.n_faces, nNf))
    g = np.zeros((m.n_faces, nNf))
    xb = a['coords'][a['faces'][a['boundary']]]
    g[a['boundary']] = sin_exp(xb)
    s = orc.System(dim, p, 0, a, tau, dirichlet=g)
    lam, its, conv, relres = s.solve(method=0, rtol=1e-14, maxits=20000)
    assert relres < 1e-10
    u, q = s.reconstruct(lam)
    ana = sin_exp(a['coords'][a['cells']])
    err = l2_projection_error(dim, p, a, u, ana)
    assert err < 1e-2, err


def test_manufactured_convergence_rates():
    from hyperfox_amd.problems import diffusion_source_2d
    import hyperfox_amd.problems as pr
    for p in (1, 2, 3):
        errs = []
        for n in (4, 8):
            w = diffusion_source_2d(n, p, jitter=0.1, seed=11)
            a = w.mesh.arrays()
            X = a['coords'][a['cells']]
            xip = np.stack([orc.ip_coords(2, p, X[e]) for e in range(X.shape[0])])
            s = orc.System(2, p, 1, a, w.tau, dirichlet=w.dirichlet, f_ip=pr.f_sin2d(xip))
            lam, its, conv, relres = s.solve(method=1)
            u, q = s.reconstruct(lam)
            errs.append(l2_projection_error(2, p, a, u, pr.u_sin2d(X)))
        rate = np.log2(errs[0] / errs[1])
        assert rate > p + 0.5, (p, errs, rate)


def test_gmres_matches_dense_solve():
    from hyperfox_amd.problems import convection_diffusion_3d
    w = convection_diffusion_3d(2, 2)
    a = w.mesh.arrays()
    X = a['coords'][a['cells']]
    xip = np.stack([orc.ip_coords(3, 2, X[e]) for e in range(X.shape[0])])
    s = orc.System(3, 2, 2, a, w.tau, tau_sides=2, D=w.D, ndD=1, b=w.b, dirichlet=w.dirichlet,
                   f_ip=w.source_at(xip))
    l1, _, _, r1 = s.solve(method=1)
    l0, its, conv, r0 = s.solve(method=0, rtol=1e-14, maxits=5000)
    assert conv and r1 < 1e-13 and r0 < 1e-11
    assert np.linalg.norm(l0 - l1) / np.linalg.norm(l1) < 1e-10


def _la_systems(k):
    n = k + 1
    I = np.eye(n)
    rhs_id = np.arange(n, dtype=float)
    L = np.tril(np.ones((n, n)))
    tri_sol = np.zeros(n)
    tri_sol[0] = 1.0
    H = np.eye(n) + np.diag(np.ones(n - 1), -1)
    hinge_sol = np.array([1.0 if i % 2 == 0 else 0.0 for i in range(n)])
    return [(I, rhs_id, rhs_id), (L, np.ones(n), tri_sol), (H, np.ones(n), hinge_sol)]


@pytest.mark.parametrize('k', [0, 1, 7, 30, 63, 99])
def test_oracle_gmres_known_answer_systems(k):
    for A, b, sol in _la_systems(k):
        x, its = orc.dense_gmres(A, b, rtol=1e-16, maxits=1000)
        assert np.allclose(x, sol, atol=1e-12, rtol=0)
