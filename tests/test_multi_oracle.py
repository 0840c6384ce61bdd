"""The oracle of the scalar models at m = dim DOFs per node
(oracle/burgers_oracle.py MultiSystem: HDGOperator::multiplyDOFs,
src/operator/HDGOperator.cpp:17-28, as tests/regression/HDG/
TestHDGLaplaceMulti.cpp runs HDGLaplaceModel), pinned on the CPU:

* with tau = t I every component is the scalar problem: the whole-mesh
  solution equals the scalar oracle's (itself pinned on the reference's
  fixtures) component by component, for HDGLaplaceModel and for the
  diffusion-source model with a nodal D and a source per component;
* a linear field per component (constant D, no source) is reproduced
  exactly for any (coupling, non-symmetric, two-sided) m x m tau, since
  u = lambda on the faces there;
* the same field under a coupling tau with Dirichlet data that are not
  linear differs from the decoupled solution (the tau does couple).
"""
import os
import sys

import numpy as np
import pytest

import hyperfox_amd as hf

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, '..', 'oracle'))
import oracle as orc  # noqa: E402
import burgers_oracle as bo  # noqa: E402


def mesh(dim, p, n=2, seed=3):
    m = hf.Mesh.structured(dim, p, n, 0.1 if dim == 2 else 0.05, seed)
    return m, m.arrays()


def face_values(a, fn):
    return fn(a['coords'][a['faces']])          # (nFaces, nNf, m)


@pytest.mark.parametrize('laplace', [True, False])
@pytest.mark.parametrize('dim,p', [(2, 1), (2, 3), (3, 1), (3, 2)])
def test_identity_tau_is_the_scalar_problem_per_component(dim, p, laplace):
    m, a = mesh(dim, p)
    s = m.sizes
    rng = np.random.default_rng(dim * 10 + p)
    t = 0.5 + rng.random((m.n_faces, s['nNf']))
    tau = np.einsum('fn,ab->fnab', t, np.eye(dim))[:, :, None]
    D = None if laplace else 0.5 + rng.random(m.n_nodes)
    g = rng.standard_normal((m.n_faces, s['nNf'], dim))
    f = None if laplace else rng.standard_normal((m.n_cells, s['nIP'], dim))
    ms = bo.MultiSystem(dim, p, a, tau, Dn=D, f_ip=f, dirichlet=g)
    ms.assemble_linear()
    lam = ms.solve()
    u, q = ms.reconstruct(lam)
    u = u.reshape(m.n_cells, s['nN'], dim)
    for c in range(dim):
        sc = orc.System(dim, p, 0 if laplace else 1, a, t, D=D, ndD=0 if laplace else 1,
                        f_ip=None if laplace else f[..., c], dirichlet=g[..., c])
        lo, _, _, _ = sc.solve(method=1)
        uo, _ = sc.reconstruct(lo)
        assert np.abs(u[..., c] - uo).max() <= 1e-10 * max(1.0, np.abs(uo).max())
        assert np.abs(lam.reshape(m.n_faces, s['nNf'], dim)[..., c] - lo.reshape(m.n_faces, s['nNf'])).max() < 1e-10


def coupling_tau(m, dim, rng):
    C = np.eye(dim) * 2.0 + 0.4 * rng.standard_normal((dim, dim))   # non-symmetric
    return np.einsum('fns,ab->fnsab', 1.0 + rng.random((m.n_faces, m.sizes['nNf'], 2)), C)


@pytest.mark.parametrize('dim,p', [(2, 1), (2, 2), (3, 1), (3, 2)])
def test_linear_fields_are_exact_for_any_tau(dim, p):
    m, a = mesh(dim, p)
    rng = np.random.default_rng(5 + p)
    G = rng.standard_normal((dim, dim))
    c0 = rng.standard_normal(dim)

    def ex(x):
        return x @ G.T + c0
    # a constant D, so that the linear field solves -D lap u = 0
    ms = bo.MultiSystem(dim, p, a, coupling_tau(m, dim, rng), Dn=np.full(m.n_nodes, 0.7),
                        dirichlet=face_values(a, ex))
    ms.assemble_linear()
    u, q = ms.reconstruct(ms.solve())
    ue = ex(a['coords'][a['cells']]).reshape(m.n_cells, -1)
    assert np.abs(u - ue).max() < 1e-11


def test_coupling_tau_couples():
    dim, p = 2, 2
    m, a = mesh(dim, p)
    rng = np.random.default_rng(9)
    g = face_values(a, lambda x: np.stack([np.sin(3 * x[..., 0]), np.cos(2 * x[..., 1])], -1))
    sols = []
    for tau in (coupling_tau(m, dim, rng), np.broadcast_to(np.eye(dim), (m.n_faces, m.sizes['nNf'], 1, dim, dim))):
        ms = bo.MultiSystem(dim, p, a, np.ascontiguousarray(tau), dirichlet=g)
        ms.assemble_linear()
        sols.append(ms.reconstruct(ms.solve())[0])
    assert np.abs(sols[0] - sols[1]).max() > 1e-4
