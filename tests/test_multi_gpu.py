"""Scalar models at m = dim DOFs per node on the GPU (hfx_multi_create:
HDGOperator::multiplyDOFs, src/operator/HDGOperator.cpp:17-28) against the
multi-DOF oracle (pinned in test_multi_oracle.py): condensed blocks and
solutions with a coupling two-sided m x m tau, a nodal D and a source per
component; HDGLaplaceModel on the reference's regression meshes as
tests/regression/HDG/TestHDGLaplaceMulti.cpp runs it (tau = I, Dirichlet
sin(x) e^y on every component, DirichletModel rows, its l2 error bar 1e-1);
linear fields exact under a coupling tau."""
import os
import sys

import numpy as np
import pytest

import hyperfox_amd as hf
from conftest import GOLDEN
from hdg_helpers import l2_projection_error

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, '..', 'oracle'))
import burgers_oracle as bo  # noqa: E402

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def workload(dim, p, seed=4, laplace=False):
    m = hf.Mesh.structured(dim, p, 2, 0.1 if dim == 2 else 0.05, seed)
    s = m.sizes
    rng = np.random.default_rng(seed + 7 * p + dim)
    C = 2.0 * np.eye(dim) + 0.3 * rng.standard_normal((dim, dim))
    tau = np.einsum('fns,ab->fnsab', 1.0 + rng.random((m.n_faces, s['nNf'], 2)), C)
    D = None if laplace else 0.5 + rng.random(m.n_nodes)
    g = rng.standard_normal((m.n_faces, s['nNf'], dim))
    f = None if laplace else rng.standard_normal((m.n_cells, s['nIP'], dim))
    return m, tau, D, g, f


def gpu(dim, p, m, tau, D, g, f):
    prob = hf.MultiProblem(dim, p)
    prob.set_mesh(m)
    prob.set_fields(tau, D=D, dirichlet=g, source=f)
    prob.set_boundary_model(hf.BC_DIRICHLET)
    return prob


@pytest.mark.parametrize('laplace', [False, True])
@pytest.mark.parametrize('dim,p', [(2, 1), (2, 2), (2, 3), (2, 4), (2, 5), (3, 1), (3, 2), (3, 3)])
def test_blocks_and_solution_match_oracle(dim, p, laplace):
    print('multi', dim, p, laplace, flush=True)
    m, tau, D, g, f = workload(dim, p, laplace=laplace)
    prob = gpu(dim, p, m, tau, D, g, f)
    prob.assemble()
    S, S0, U, Q, U0, Q0 = prob.condensed()
    ms = bo.MultiSystem(dim, p, m.arrays(), tau, Dn=D, f_ip=f, dirichlet=g)
    ms.assemble_linear()
    for name, x, y in zip(('S', 'S0', 'U', 'Q', 'U0', 'Q0'), (S, S0, U, Q, U0, Q0),
                          (ms.S, ms.S0, ms.U, ms.Q, ms.U0, ms.Q0)):
        assert rel(x, y) < 1e-10, name
    it, reason, rr = prob.solve(pc=hf.PC_BJACOBI, rtol=1e-13, maxits=5000)
    assert reason > 0 and rr <= 1e-13
    sol, _, flux, tr = prob.state()
    lam = ms.solve()
    u, q = ms.reconstruct(lam)
    nC = m.n_cells
    assert rel(tr.reshape(-1), lam) < 1e-9
    assert rel(sol.reshape(nC, -1), u) < 1e-9 and rel(flux.reshape(nC, -1), q) < 1e-9


@pytest.mark.parametrize('dim,h,p', [(2, '1e-1', 1), (2, '1e-1', 3), (3, '2e-1', 1), (3, '2e-1', 2)])
def test_laplace_multi_regression_meshes(dim, h, p):
    """TestHDGLaplaceMulti.cpp: HDGLaplaceModel with m = dim, tau = I,
    Dirichlet sin(x) e^y on each component; the reference checks the l2
    error (over every component) below 1e-1.  Also equal to the oracle."""
    print('laplace multi', dim, h, p, flush=True)
    z = np.load(os.path.join(GOLDEN, 'meshes', 'regression_dim-%d_h-%s_ord-%d.npz' % (dim, h, p)))
    m = hf.Mesh.from_cells(dim, p, z['nodes'], z['cells'])
    a = m.arrays()
    s = m.sizes
    sx = lambda x: np.sin(x[..., 0]) * np.exp(x[..., 1])
    g = np.zeros((m.n_faces, s['nNf'], dim))
    g[a['boundary']] = sx(a['coords'][a['faces'][a['boundary']]])[..., None]
    tau = np.broadcast_to(np.eye(dim), (m.n_faces, s['nNf'], 1, dim, dim)).copy()
    prob = gpu(dim, p, m, tau, None, g, None)
    prob.assemble()
    it, reason, rr = prob.solve(pc=hf.PC_BJACOBI, rtol=1e-13, maxits=20000)
    assert reason > 0
    sol = prob.state()[0]
    ex = sx(a['coords'][a['cells']])
    err2 = sum(l2_projection_error(dim, p, a, sol[..., c], ex) ** 2 for c in range(dim))
    assert np.sqrt(err2) < 1e-1
    ms = bo.MultiSystem(dim, p, a, tau, dirichlet=g)
    ms.assemble_linear()
    u, _ = ms.reconstruct(ms.solve())
    assert rel(sol.reshape(m.n_cells, -1), u) < 1e-9


@pytest.mark.parametrize('dim,p', [(2, 2), (3, 2)])
def test_linear_fields_exact_with_coupling_tau(dim, p):
    m, tau, _, _, _ = workload(dim, p)
    a = m.arrays()
    rng = np.random.default_rng(11)
    G, c0 = rng.standard_normal((dim, dim)), rng.standard_normal(dim)
    ex = lambda x: x @ G.T + c0
    prob = gpu(dim, p, m, tau, np.full(m.n_nodes, 0.7), ex(a['coords'][a['faces']]), None)
    prob.assemble()
    it, reason, rr = prob.solve(pc=hf.PC_BJACOBI, rtol=1e-13, maxits=5000)
    sol = prob.state()[0]
    assert np.abs(sol - ex(a['coords'][a['cells']])).max() < 1e-10


def test_multi_checks():
    with pytest.raises(hf.HfxError, match='2-D orders 1-5 and 3-D orders 1-3'):
        hf.MultiProblem(3, 4)
    m, tau, D, g, f = workload(2, 1)
    prob = hf.MultiProblem(2, 1)
    prob.set_mesh(m)
    with pytest.raises(hf.HfxError, match='not an nGamma problem'):
        hf.NGammaProblem.set_fields(prob, tau[:, :, :1], b=np.zeros((m.n_nodes, 2)))
