"""IntegratedDirichletModel on the GPU (src/model/IntegratedDirichletModel.cpp:
35-56 through HDGSolver::applyBoundaryConditions, HDGSolver.cpp:489-502)
against the oracle (pinned in test_boundary_models_oracle.py): the trace
system on straight and curved meshes with both element kernels, the solution
(that of DirichletModel), a per-face mix of the two models, and the checks."""
import numpy as np
import pytest

import hyperfox_amd as hf
import hyperfox_amd.problems as pr
from hdg_helpers import rel
from test_gpu_parity import oracle_system, run_gpu, small_workload

pytestmark = pytest.mark.gpu


def curved_workload(dim, p, seed=5):
    """small_workload's problem on a mesh whose high-order nodes are moved by
    a smooth field (vertices fixed): curved faces and cells."""
    w = small_workload(dim, p, 2, seed)
    a = w.mesh.arrays()
    x = a['coords'].copy()
    vert = np.unique(a['cells'][:, :dim + 1])
    ho = np.setdiff1d(np.arange(x.shape[0]), vert)
    for k in range(dim):
        x[ho, k] += 0.03 * np.sin(np.pi * x[ho, (k + 1) % dim]) * np.sin(np.pi * x[ho, k])
    m = hf.Mesh.from_cells(dim, p, x, a['cells'])
    assert m.n_faces == w.mesh.n_faces
    w.mesh = m
    return w


def boundary_faces(w):
    return np.nonzero(np.asarray(w.mesh.arrays()['face2cell'])[:, 1] < 0)[0]


@pytest.mark.parametrize('kernel', [0, 1])
@pytest.mark.parametrize('dim,p', [(2, 1), (2, 2), (2, 4), (3, 1), (3, 2), (3, 3)])
def test_trace_system_with_integrated_dirichlet_matches_oracle(dim, p, kernel):
    w = small_workload(dim, p, 2)
    prob, f = run_gpu(w)
    prob.set_kernel(kernel)
    prob.set_boundary_model(hf.BC_INTEGRATED_DIRICHLET)
    prob.assemble()
    rp, cols, vals, rhs = prob.trace_csr()
    s = oracle_system(w, f)
    s.set_boundary_model(1)
    rpo, colso, valso, rhso = s.csr()
    assert np.array_equal(rp, rpo) and np.array_equal(cols, colso)
    assert rel(vals, valso) < 1e-12, rel(vals, valso)
    assert rel(rhs, rhso) < 1e-12, rel(rhs, rhso)


@pytest.mark.parametrize('dim,p', [(2, 2), (2, 3), (3, 2)])
def test_curved_faces_match_oracle(dim, p):
    w = curved_workload(dim, p)
    prob, f = run_gpu(w)
    assert not prob.kernel_info()['mesh_affine']
    prob.set_boundary_model(hf.BC_INTEGRATED_DIRICHLET)
    prob.assemble()
    rp, cols, vals, rhs = prob.trace_csr()
    s = oracle_system(w, f)
    s.set_boundary_model(1)
    rpo, colso, valso, rhso = s.csr()
    assert np.array_equal(cols, colso)
    assert rel(vals, valso) < 1e-12, rel(vals, valso)
    assert rel(rhs, rhso) < 1e-12, rel(rhs, rhso)


@pytest.mark.parametrize('dim,p', [(2, 3), (3, 2)])
def test_solution_is_the_dirichlet_one(dim, p):
    w = small_workload(dim, p, 2)
    prob, f = run_gpu(w)
    prob.set_krylov(hf.KRYLOV_BICGSTAB)
    prob.assemble()
    assert prob.solve(pc=hf.PC_BJACOBI, rtol=1e-13, maxits=20000)[1] == 2
    u0, q0, l0 = prob.solution()
    bf = boundary_faces(w)
    prob.set_boundary_model(hf.BC_INTEGRATED_DIRICHLET, bf[::2])   # a mix of the two models
    prob.assemble()
    assert prob.solve(pc=hf.PC_BJACOBI, rtol=1e-13, maxits=20000)[1] == 2
    u1, q1, l1 = prob.solution()
    assert rel(l1, l0) < 1e-10 and rel(u1, u0) < 1e-10
    nNf = w.mesh.sizes['nNf']
    g = np.asarray(w.dirichlet).reshape(-1, nNf)
    assert np.allclose(l1.reshape(-1, nNf)[bf], g[bf], rtol=0, atol=1e-10 * np.abs(g).max())


def test_boundary_model_arguments_are_checked():
    w = small_workload(2, 2, 1)
    prob, f = run_gpu(w)
    interior = np.nonzero(np.asarray(w.mesh.arrays()['face2cell'])[:, 1] >= 0)[0]
    with pytest.raises(hf.HfxError, match='not a boundary face'):
        prob.set_boundary_model(hf.BC_INTEGRATED_DIRICHLET, interior[:1])
    with pytest.raises(hf.HfxError, match='unknown kind'):
        prob.set_boundary_model(5)
    prob.set_boundary_model(hf.BC_INTEGRATED_DIRICHLET)
    prob.set_boundary_model(hf.BC_DIRICHLET)          # back to identity rows everywhere
    prob.assemble()
    rp, cols, vals, rhs = prob.trace_csr()
    s = oracle_system(w, f)
    assert rel(vals, s.csr()[2]) < 1e-12


# ---------------------------------------------------------------- HDG type


@pytest.mark.parametrize('mode', ['set', 'add'])
@pytest.mark.parametrize('kernel', [0, 1])
@pytest.mark.parametrize('dim,p', [(2, 1), (2, 3), (2, 5), (3, 2), (3, 3)])
def test_hdg_type_boundary_matches_oracle(dim, p, kernel, mode):
    from test_boundary_models_oracle import hdg_B
    w = small_workload(dim, p, 2)
    prob, f = run_gpu(w)
    prob.set_kernel(kernel)
    bf = boundary_faces(w)[::2]
    B, r = hdg_B(dim, p, bf.size, 13)
    st = mode == 'set'
    prob.set_boundary_model(hf.BC_INTEGRATED_DIRICHLET, boundary_faces(w)[1::4])  # with a mix of the others
    prob.set_boundary_hdg(bf, B, r, matrix_set=st, rhs_set=st)
    prob.assemble()
    rp, cols, vals, rhs = prob.trace_csr()
    s = oracle_system(w, f)
    s.set_boundary_model(1, boundary_faces(w)[1::4])
    s.set_boundary_hdg(bf, B, r, matrix_set=st, rhs_set=st)
    rpo, colso, valso, rhso = s.csr()
    assert np.array_equal(cols, colso)
    assert rel(vals, valso) < 1e-12, rel(vals, valso)
    assert rel(rhs, rhso) < 1e-12, rel(rhs, rhso)


def test_hdg_type_arguments_are_checked_and_clearing_restores_dirichlet():
    from test_boundary_models_oracle import hdg_B
    w = small_workload(2, 2, 1)
    prob, f = run_gpu(w)
    bf = boundary_faces(w)
    B, r = hdg_B(2, 2, 2, 1)
    interior = np.nonzero(np.asarray(w.mesh.arrays()['face2cell'])[:, 1] >= 0)[0]
    with pytest.raises(hf.HfxError, match='not a boundary face'):
        prob.set_boundary_hdg([bf[0], interior[0]], B, r)
    with pytest.raises(hf.HfxError, match='twice'):
        prob.set_boundary_hdg([bf[0], bf[0]], B, r)
    prob.set_boundary_hdg(bf[:2], B, r)
    with pytest.raises(hf.HfxError, match='HDG-type'):
        prob.set_boundary_model(hf.BC_INTEGRATED_DIRICHLET, bf[:1])
    prob.set_boundary_hdg([], np.zeros((0, 3, 12)), np.zeros((0, 3)))
    prob.assemble()
    vals = prob.trace_csr()[2]
    assert rel(vals, oracle_system(w, f).csr()[2]) < 1e-12
