"""The nGamma path on the GPU (HDGnGammaModel, HDGnGammaModel.cpp:5-170, 2-D,
m = 2, on HDGSolver + NonLinearWrapper) against the nGamma oracle (pinned in
test_ngamma_oracle.py): the condensed blocks on straight and curved meshes
with two-sided matrix tau, D and G (each present or absent), b, a
linearisation state and a source; the trace system; the solve; the nonlinear
loop on the manufactured problem; the checks."""
import os
import sys

import numpy as np
import pytest

import hyperfox_amd as hf

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, '..', 'oracle'))
import burgers_oracle as bo  # noqa: E402
from test_ngamma_oracle import exact, manufactured  # noqa: E402

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def workload(p, seed=5, curved=False, D=True, G=True):
    rng = np.random.default_rng(seed)
    mesh = hf.Mesh.structured(2, p, 2, 0.15, seed)
    if curved:
        a = mesh.arrays()
        x = a['coords'].copy()
        ho = np.setdiff1d(np.arange(x.shape[0]), np.unique(a['cells'][:, :3]))
        for k in range(2):
            x[ho, k] += 0.03 * np.sin(np.pi * x[ho, (k + 1) % 2]) * np.sin(np.pi * x[ho, k])
        mesh = hf.Mesh.from_cells(2, p, x, a['cells'])
    s = mesh.sizes
    C = 2.0 * np.eye(2) + 0.2 * rng.standard_normal((2, 2))
    tau = np.einsum('fns,ab->fnsab', 1.0 + rng.random((mesh.n_faces, s['nNf'], 2)), C)
    nn = mesh.n_nodes
    fields = dict(tau=tau, D=0.5 + rng.random(nn) if D else None, G=0.5 + rng.random(nn) if G else None,
                  b=rng.standard_normal((nn, 2)), sound_speed=0.8,
                  dirichlet=rng.standard_normal((mesh.n_faces, s['nNf'], 2)),
                  source=rng.standard_normal((mesh.n_cells, s['nIP'], 2)))
    buf = np.stack([1.0 + rng.random((mesh.n_cells, s['nN'])), 0.5 * rng.standard_normal((mesh.n_cells, s['nN']))], 2)
    state = dict(buffer=buf, trace=0.5 * rng.standard_normal((mesh.n_faces, s['nNf'], 2)))
    return mesh, fields, state


def gpu_problem(p, mesh, fields, state):
    prob = hf.NGammaProblem(p)
    prob.set_mesh(mesh)
    prob.set_fields(**fields)
    prob.set_state(buffer=state['buffer'], trace=state['trace'])
    return prob


def oracle_problem(p, mesh, fields):
    ng = dict(G=fields['G'], b=fields['b'], cs=fields['sound_speed'])
    return bo.BurgersSystem(2, p, mesh.arrays(), fields['tau'], Dn=fields['D'], f_ip=fields['source'],
                            dirichlet=fields['dirichlet'], ngamma=ng)


@pytest.mark.parametrize('curved', [False, True])
@pytest.mark.parametrize('p', [1, 2, 3, 4, 5])
def test_condensed_blocks_match_oracle(p, curved):
    print('ngamma condensed', p, curved, flush=True)
    mesh, fields, state = workload(p, curved=curved)
    prob = gpu_problem(p, mesh, fields, state)
    prob.assemble()
    got = prob.condensed()
    orc = oracle_problem(p, mesh, fields)
    orc.assemble(state['buffer'].reshape(mesh.n_cells, -1), state['trace'])
    for name, g, o in zip(('S', 'S0', 'U', 'Q', 'U0', 'Q0'), got, (orc.S, orc.S0, orc.U, orc.Q, orc.U0, orc.Q0)):
        assert rel(g, o) < 1e-10, name


@pytest.mark.parametrize('D,G', [(False, True), (True, False), (False, False)])
def test_optional_diffusion_fields(D, G):
    print('ngamma optional fields', D, G, flush=True)
    mesh, fields, state = workload(2, D=D, G=G)
    prob = gpu_problem(2, mesh, fields, state)
    prob.assemble()
    rowptr, col, vals, rhs = prob.trace_system()
    bs = prob.bs
    A = np.zeros((prob.nb * bs, prob.nb * bs))
    for F in range(prob.nb):
        for k in range(rowptr[F], rowptr[F + 1]):
            A[F * bs:(F + 1) * bs, col[k] * bs:(col[k] + 1) * bs] = vals[k]
    orc = oracle_problem(2, mesh, fields)
    Ao, bo_ = orc.assemble(state['buffer'].reshape(mesh.n_cells, -1), state['trace'])
    assert rel(A, Ao.toarray()) < 1e-10 and rel(rhs, bo_) < 1e-10


@pytest.mark.parametrize('pc', [hf.PC_BJACOBI, hf.PC_ILU0])
def test_solve_matches_oracle(pc):
    print('ngamma solve', pc, flush=True)
    mesh, fields, state = workload(3)
    prob = gpu_problem(3, mesh, fields, state)
    prob.assemble()
    it, reason, rr = prob.solve(pc=pc, rtol=1e-13, maxits=5000)
    assert reason > 0 and rr <= 1e-13
    sol, buf, flux, tr = prob.state()
    orc = oracle_problem(3, mesh, fields)
    orc.assemble(state['buffer'].reshape(mesh.n_cells, -1), state['trace'])
    lam = orc.solve()
    u, q = orc.reconstruct(lam)
    assert rel(tr.reshape(-1), lam) < 1e-9 and rel(sol.reshape(mesh.n_cells, -1), u) < 1e-9
    assert rel(flux.reshape(mesh.n_cells, -1), q) < 1e-9


@pytest.mark.parametrize('p', [1, 2, 3])
def test_manufactured_problem_matches_oracle(p):
    print('ngamma manufactured', p, flush=True)
    mesh, a, sysm, f, tau, g, ng = manufactured(8, p)
    prob = hf.NGammaProblem(p)
    prob.set_mesh(mesh)
    prob.set_fields(tau, D=sysm.Dn, G=ng['G'], b=ng['b'], sound_speed=ng['cs'], dirichlet=g, source=f)
    init = np.tile([1.0, 0.0], (mesh.n_cells, sysm.s['nN']))
    prob.set_state(sol=init, buffer=init)
    its, res, lin = prob.nonlinear_solve(max_iters=20, res_tol=1e-10, pc=hf.PC_BJACOBI, rtol=1e-13)
    u, buf, q, lam, its_o, res_o = sysm.newton(init, init, np.zeros(sysm.ndofs), max_iters=20, tol=1e-10)
    sol, gbuf, flux, tr = prob.state()
    nC = mesh.n_cells
    assert its == its_o and res < 1e-10
    assert rel(sol.reshape(nC, -1), u) < 1e-8 and rel(flux.reshape(nC, -1), q) < 1e-8
    ex = exact(a['coords'][a['cells'].reshape(-1)]).reshape(nC, -1)
    err = np.linalg.norm(sol.reshape(nC, -1) - ex) / np.linalg.norm(ex)
    err_o = np.linalg.norm(u - ex) / np.linalg.norm(ex)
    assert abs(err - err_o) <= 1e-6 * err_o


def test_checks():
    with pytest.raises(hf.HfxError, match='2-D orders 1-5'):
        hf.NGammaProblem(6)
    mesh, fields, state = workload(1)
    prob = hf.NGammaProblem(1)
    prob.set_mesh(mesh)
    with pytest.raises(hf.HfxError, match='must provide a b vector field'):
        prob.set_fields(fields['tau'])
    with pytest.raises(hf.HfxError, match='takes hfx_ngamma_set_fields'):
        hf.BurgersProblem.set_fields(prob, fields['tau'])
    with pytest.raises(hf.HfxError, match='set the mesh and the fields first'):
        prob.assemble()
    burg = hf.BurgersProblem(2, 1)
    burg.set_mesh(mesh)
    with pytest.raises(hf.HfxError, match='not an nGamma problem'):
        hf.NGammaProblem.set_fields(burg, fields['tau'], b=fields['b'])
