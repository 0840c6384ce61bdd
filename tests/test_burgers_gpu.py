"""The Burgers path on the GPU (HDGBurgersModel, HDGBurgersModel.cpp:34-132,
with m = dim unknowns per node, on HDGSolver + NonLinearWrapper) against the
Burgers oracle (pinned in test_burgers_oracle.py): the condensed blocks of
every cell on straight and curved meshes with two-sided matrix tau, nodal
diffusion, a linearisation state and a source; the trace system with both
boundary models; the solve and reconstruction; the nonlinear loop on the
reference's stationary regression case, undamped and damped; the checks."""
import os
import sys

import numpy as np
import pytest

import hyperfox_amd as hf
from hyperfox_amd.burgers import BC_DIRICHLET, BC_INTEGRATED_DIRICHLET

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, '..', 'oracle'))
import burgers_oracle as bo  # noqa: E402
from test_burgers_oracle import cole_hopf, regression_case  # noqa: E402

pytestmark = pytest.mark.gpu

# 3-D p = 3: the cell's working set (279 KB) is kept in a device workspace
CASES = [(2, 1), (2, 2), (2, 3), (2, 4), (2, 5), (3, 1), (3, 2), (3, 3)]


def rel(a, b):
    return np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)


def workload(dim, p, seed=3, curved=False, n=2):
    """A jittered (or curved) structured mesh with random fields and state."""
    rng = np.random.default_rng(seed)
    mesh = hf.Mesh.structured(dim, p, n, 0.15, seed)
    if curved:
        a = mesh.arrays()
        x = a['coords'].copy()
        ho = np.setdiff1d(np.arange(x.shape[0]), np.unique(a['cells'][:, :dim + 1]))
        for k in range(dim):
            x[ho, k] += 0.03 * np.sin(np.pi * x[ho, (k + 1) % dim]) * np.sin(np.pi * x[ho, k])
        mesh = hf.Mesh.from_cells(dim, p, x, a['cells'])
    s, m = mesh.sizes, dim
    C = np.eye(m) + 0.2 * rng.standard_normal((m, m))
    tau = np.einsum('fns,ab->fnsab', 1.0 + rng.random((mesh.n_faces, s['nNf'], 2)), C)
    fields = dict(tau=tau, D=0.5 + rng.random(mesh.n_nodes),
                  dirichlet=rng.standard_normal((mesh.n_faces, s['nNf'], m)),
                  source=rng.standard_normal((mesh.n_cells, s['nIP'], m)))
    state = dict(buffer=0.5 * rng.standard_normal((mesh.n_cells, s['nN'], m)),
                 trace=0.5 * rng.standard_normal((mesh.n_faces, s['nNf'], m)))
    return mesh, fields, state


def gpu_problem(dim, p, mesh, fields, state, bc=BC_INTEGRATED_DIRICHLET):
    prob = hf.BurgersProblem(dim, p)
    prob.set_mesh(mesh)
    prob.set_fields(**fields)
    prob.set_boundary_model(bc)
    prob.set_state(buffer=state['buffer'], trace=state['trace'])
    return prob


def oracle_problem(dim, p, mesh, fields, bc='integrated'):
    return bo.BurgersSystem(dim, p, mesh.arrays(), fields['tau'], Dn=fields['D'],
                            f_ip=fields['source'], dirichlet=fields['dirichlet'], boundary=bc)


@pytest.mark.parametrize('curved', [False, True])
@pytest.mark.parametrize('dim,p', CASES)
def test_condensed_blocks_match_oracle(dim, p, curved):
    print('burgers condensed', dim, p, curved, flush=True)
    mesh, fields, state = workload(dim, p, curved=curved)
    prob = gpu_problem(dim, p, mesh, fields, state)
    prob.assemble()
    got = prob.condensed()
    orc = oracle_problem(dim, p, mesh, fields)
    orc.assemble(state['buffer'].reshape(mesh.n_cells, -1), state['trace'])
    for name, g, o in zip(('S', 'S0', 'U', 'Q', 'U0', 'Q0'), got, (orc.S, orc.S0, orc.U, orc.Q, orc.U0, orc.Q0)):
        assert rel(g, o) < 1e-10, name


@pytest.mark.parametrize('bc', [BC_INTEGRATED_DIRICHLET, BC_DIRICHLET])
@pytest.mark.parametrize('dim,p', [(2, 2), (2, 4), (3, 1), (3, 2)])
def test_trace_system_matches_oracle(dim, p, bc):
    print('burgers trace system', dim, p, bc, flush=True)
    mesh, fields, state = workload(dim, p)
    prob = gpu_problem(dim, p, mesh, fields, state, bc)
    prob.assemble()
    rowptr, col, vals, rhs = prob.trace_system()
    bs = prob.bs
    A = np.zeros((prob.nb * bs, prob.nb * bs))
    for F in range(prob.nb):
        for k in range(rowptr[F], rowptr[F + 1]):
            A[F * bs:(F + 1) * bs, col[k] * bs:(col[k] + 1) * bs] = vals[k]
    orc = oracle_problem(dim, p, mesh, fields, 'integrated' if bc == BC_INTEGRATED_DIRICHLET else 'dirichlet')
    Ao, bo_ = orc.assemble(state['buffer'].reshape(mesh.n_cells, -1), state['trace'])
    assert rel(A, Ao.toarray()) < 1e-10
    assert rel(rhs, bo_) < 1e-10


@pytest.mark.parametrize('pc', [hf.PC_BJACOBI, hf.PC_ILU0, hf.PC_JACOBI])
@pytest.mark.parametrize('dim,p', [(2, 3), (3, 1)])
def test_solve_and_reconstruction_match_oracle(dim, p, pc):
    print('burgers solve', dim, p, pc, flush=True)
    mesh, fields, state = workload(dim, p)
    prob = gpu_problem(dim, p, mesh, fields, state)
    prob.assemble()
    it, reason, rr = prob.solve(pc=pc, rtol=1e-13, maxits=5000)
    assert reason > 0 and rr <= 1e-13
    sol, buf, flux, tr = prob.state()
    orc = oracle_problem(dim, p, mesh, fields)
    orc.assemble(state['buffer'].reshape(mesh.n_cells, -1), state['trace'])
    lam = orc.solve()
    u, q = orc.reconstruct(lam)
    assert rel(tr.reshape(-1), lam) < 1e-9
    assert rel(sol.reshape(mesh.n_cells, -1), u) < 1e-9
    assert rel(flux.reshape(mesh.n_cells, -1), q) < 1e-9
    assert np.array_equal(buf, state['buffer'])


@pytest.mark.parametrize('damp', [0.0, 0.3])
@pytest.mark.parametrize('p', [1, 2, 3])
def test_stationary_regression_case_matches_oracle(p, damp):
    """TestHDGBurgersStat.cpp on regression_dim-2_h-2e-1: the reference's
    settings (zero initial guess, at most 10 iterations, tolerance 1e-6;
    GMRES + block Jacobi) with and without damping."""
    print('burgers regression', p, damp, flush=True)
    mesh_a, orc = regression_case('2e-1', p)
    mesh = hf.Mesh.from_cells(2, p, mesh_a['coords'], mesh_a['cells'])
    prob = hf.BurgersProblem(2, p)
    prob.set_mesh(mesh)
    prob.set_fields(orc.tau, D=orc.Dn, dirichlet=orc.g)
    # damping makes the convergence linear (factor about damp per step): more steps
    nmax = 10 if damp == 0.0 else 30
    its, res, lin = prob.nonlinear_solve(max_iters=nmax, res_tol=1e-6, damp=damp, pc=hf.PC_BJACOBI, rtol=1e-13)
    nC = mesh.n_cells
    z0 = np.zeros((nC, orc.s['nN'] * 2))
    u, buf, q, lam, its_o, res_o = orc.newton(z0, z0, np.zeros(orc.ndofs), max_iters=nmax, tol=1e-6, damp=damp)
    sol, gbuf, flux, tr = prob.state()
    assert its == its_o and res < 1e-6 and lin > 0
    assert abs(res - res_o) <= 1e-3 * res_o
    assert rel(sol.reshape(nC, -1), u) < 1e-8
    assert rel(flux.reshape(nC, -1), q) < 1e-8
    assert rel(tr.reshape(-1), lam) < 1e-8
    ana = cole_hopf(mesh_a['coords'][mesh_a['cells'].reshape(-1)]).reshape(nC, -1)
    err = np.sqrt(np.sum((ana - sol.reshape(nC, -1)) ** 2) / np.sum(ana ** 2))
    err_o = np.sqrt(np.sum((ana - u) ** 2) / np.sum(ana ** 2))
    assert abs(err - err_o) <= 1e-6 * err_o + 1e-12


@pytest.mark.parametrize('pc', [hf.PC_BJACOBI, hf.PC_JACOBI])
def test_3d_p3_solve_matches_oracle(pc):
    """3-D p = 3 (cell working set in a device workspace, block size 30):
    trace solve and reconstruction against the oracle."""
    print('burgers solve 3-D p=3', pc, flush=True)
    mesh, fields, state = workload(3, 3)
    prob = gpu_problem(3, 3, mesh, fields, state)
    prob.assemble()
    it, reason, rr = prob.solve(pc=pc, rtol=1e-13, maxits=5000)
    assert reason > 0 and rr <= 1e-13
    sol, buf, flux, tr = prob.state()
    orc = oracle_problem(3, 3, mesh, fields)
    orc.assemble(state['buffer'].reshape(mesh.n_cells, -1), state['trace'])
    lam = orc.solve()
    u, q = orc.reconstruct(lam)
    assert rel(tr.reshape(-1), lam) < 1e-9 and rel(sol.reshape(mesh.n_cells, -1), u) < 1e-9
    assert rel(flux.reshape(mesh.n_cells, -1), q) < 1e-9


@pytest.mark.parametrize('pc', [hf.PC_BJACOBI, hf.PC_JACOBI])
def test_3d_p2_solve_matches_oracle(pc):
    """Block size 18 (3-D p = 2, m = 3): the Jacobi preconditioners and the
    Krylov kernels at that size; ILU(0) refuses the block size at set-up."""
    print('burgers solve 3-D p=2', pc, flush=True)
    mesh, fields, state = workload(3, 2)
    prob = gpu_problem(3, 2, mesh, fields, state)
    prob.assemble()
    it, reason, rr = prob.solve(pc=pc, rtol=1e-13, maxits=5000)
    assert reason > 0 and rr <= 1e-13
    sol, buf, flux, tr = prob.state()
    orc = oracle_problem(3, 2, mesh, fields)
    orc.assemble(state['buffer'].reshape(mesh.n_cells, -1), state['trace'])
    lam = orc.solve()
    u, q = orc.reconstruct(lam)
    assert rel(tr.reshape(-1), lam) < 1e-9 and rel(sol.reshape(mesh.n_cells, -1), u) < 1e-9
    assert rel(flux.reshape(mesh.n_cells, -1), q) < 1e-9
    with pytest.raises(hf.HfxError, match='block size must be one of'):
        prob.solve(pc=hf.PC_ILU0)


def test_checks():
    with pytest.raises(hf.HfxError, match='2-D orders 1-5 and 3-D orders 1-3'):
        hf.BurgersProblem(3, 4)
    prob = hf.BurgersProblem(2, 1)
    with pytest.raises(hf.HfxError, match='set the mesh and the fields first'):
        prob.assemble()
    mesh, fields, state = workload(2, 1)
    prob.set_mesh(mesh)
    with pytest.raises(hf.HfxError, match='assemble first'):
        prob.solve()
    t1 = fields['tau'][:, :, :1]
    prob.set_fields(t1, D=None)  # one-sided tau, no diffusion operator: accepted
    with pytest.raises(hf.HfxError, match='unknown boundary model'):
        prob.set_boundary_model(7)
    with pytest.raises(hf.HfxError, match='unsupported preconditioner'):
        prob.assemble()
        prob.solve(pc=hf.PC_LEARNED)


def dense(prob):
    rowptr, col, vals, rhs = prob.trace_system()
    bs = prob.bs
    A = np.zeros((prob.nb * bs, prob.nb * bs))
    for F in range(prob.nb):
        for k in range(rowptr[F], rowptr[F + 1]):
            A[F * bs:(F + 1) * bs, col[k] * bs:(col[k] + 1) * bs] = vals[k]
    return A, rhs


@pytest.mark.parametrize('model', ['burgers', 'ngamma'])
@pytest.mark.parametrize('mset,rset', [(1, 1), (0, 0), (1, 0)])
def test_hdg_boundary_matches_oracle(model, mset, rset):
    """HDG-type boundary rows (HDGSolver.cpp:503-518) with m = 2 on half of
    the boundary faces, integrated Dirichlet on the rest."""
    print('multi-DOF HDG boundary', model, mset, rset, flush=True)
    p = 2
    if model == 'burgers':
        mesh, fields, state = workload(2, p)
        prob = gpu_problem(2, p, mesh, fields, state)
    else:
        from test_ngamma_gpu import workload as ng_workload, gpu_problem as ng_problem
        mesh, fields, state = ng_workload(p)
        prob = ng_problem(p, mesh, fields, state)
    a = mesh.arrays()
    rng = np.random.default_rng(9)
    bf = np.nonzero(a['face2cell'][:, 1] < 0)[0][::2]
    bs = prob.bs
    B = rng.standard_normal((bf.size, bs, bs * 4))
    r = rng.standard_normal((bf.size, bs))
    prob.set_boundary_hdg(bf, B, r, mset, rset)
    prob.assemble()
    A, rhs = dense(prob)
    hdg = dict(faces=bf, B=B, r=r, mset=bool(mset), rset=bool(rset))
    if model == 'burgers':
        orc = bo.BurgersSystem(2, p, a, fields['tau'], Dn=fields['D'], f_ip=fields['source'],
                               dirichlet=fields['dirichlet'], hdg=hdg)
    else:
        ng = dict(G=fields['G'], b=fields['b'], cs=fields['sound_speed'])
        orc = bo.BurgersSystem(2, p, a, fields['tau'], Dn=fields['D'], f_ip=fields['source'],
                               dirichlet=fields['dirichlet'], ngamma=ng, hdg=hdg)
    Ao, bo_ = orc.assemble(state['buffer'].reshape(mesh.n_cells, -1), state['trace'])
    assert rel(A, Ao.toarray()) < 1e-10 and rel(rhs, bo_) < 1e-10
    prob.set_boundary_hdg([], np.zeros((0, bs, bs * 4)), np.zeros((0, bs)))
    prob.assemble()
    A2, rhs2 = dense(prob)
    orc.hdg = None
    Ao2, bo2 = orc.assemble(state['buffer'].reshape(mesh.n_cells, -1), state['trace'])
    assert rel(A2, Ao2.toarray()) < 1e-10 and rel(rhs2, bo2) < 1e-10
    with pytest.raises(hf.HfxError, match='not a boundary face'):
        interior = np.nonzero(a['face2cell'][:, 1] >= 0)[0][:1]
        prob.set_boundary_hdg(interior, B[:1], r[:1])


@pytest.mark.parametrize('model,dim,p', [('burgers', 2, 2), ('burgers', 3, 1), ('ngamma', 2, 2)])
def test_implicit_euler_blocks_match_oracle(model, dim, p):
    """Implicit Euler (RungeKutta BEuler) on the u rows, with an OldSolution
    apart from the linearisation state."""
    print('multi-DOF implicit Euler', model, dim, p, flush=True)
    if model == 'burgers':
        mesh, fields, state = workload(dim, p)
        prob = gpu_problem(dim, p, mesh, fields, state)
        orc = oracle_problem(dim, p, mesh, fields)
    else:
        from test_ngamma_gpu import workload as ng_workload, gpu_problem as ng_problem, oracle_problem as ng_oracle
        mesh, fields, state = ng_workload(p)
        prob = ng_problem(p, mesh, fields, state)
        orc = ng_oracle(p, mesh, fields)
    rng = np.random.default_rng(21)
    u_old = state['buffer'] + 0.1 * rng.standard_normal(state['buffer'].shape)
    dt = 0.05
    prob.set_time_scheme(1, dt)
    prob.set_state(sol=u_old)
    prob.start_step()
    prob.set_state(buffer=state['buffer'])
    prob.assemble()
    got = prob.condensed()
    orc.dt, orc.u_old = dt, u_old.reshape(mesh.n_cells, -1)
    orc.assemble(state['buffer'].reshape(mesh.n_cells, -1), state['trace'])
    for name, g, o in zip(('S', 'S0', 'U', 'Q', 'U0', 'Q0'), got, (orc.S, orc.S0, orc.U, orc.Q, orc.U0, orc.Q0)):
        assert rel(g, o) < 1e-10, name
    with pytest.raises(hf.HfxError, match='implicit Euler \\(BEuler\\) only'):
        prob.set_time_scheme(3, dt)


def test_transient_burgers_steps_match_oracle():
    """Three implicit-Euler steps of the regression case (TestHDGBurgers.cpp's
    driver: copy Solution into BufferSolution and OldSolution, then the
    nonlinear loop), from zero towards the steady Cole-Hopf field."""
    print('transient burgers', flush=True)
    mesh_a, orc = regression_case('2e-1', 2)
    mesh = hf.Mesh.from_cells(2, 2, mesh_a['coords'], mesh_a['cells'])
    prob = hf.BurgersProblem(2, 2)
    prob.set_mesh(mesh)
    prob.set_fields(orc.tau, D=orc.Dn, dirichlet=orc.g)
    dt = 0.1
    prob.set_time_scheme(1, dt)
    nC = mesh.n_cells
    u = np.zeros((nC, orc.s['nN'] * 2))
    lam = np.zeros(orc.ndofs)
    orc.dt = dt
    for step in range(3):
        prob.start_step()
        its, res, lin = prob.nonlinear_solve(max_iters=12, res_tol=1e-8, pc=hf.PC_BJACOBI, rtol=1e-13)
        orc.u_old = u.copy()
        u, buf, q, lam, its_o, res_o = orc.newton(u, u, lam, max_iters=12, tol=1e-8)
        sol, gbuf, flux, tr = prob.state()
        assert its == its_o and res < 1e-8, step
        assert rel(sol.reshape(nC, -1), u) < 1e-8 and rel(tr.reshape(-1), lam) < 1e-8, step


def test_graph_replayed_solver_gives_the_same_fields():
    """BiCGStab's steady iterations replayed as a HIP graph: same kernels in the
    same order, so the same trace and solution bit for bit.  The mesh is
    large enough that the solve runs past the first residual check (8
    iterations) and restarts, so graphs are captured and replayed."""
    print('burgers solver graphs', flush=True)
    out = []
    for graphs in (False, True):
        mesh, fields, state = workload(2, 3, n=12)
        prob = gpu_problem(2, 3, mesh, fields, state)
        prob.set_krylov(1)
        prob.set_solver_graphs(graphs)
        prob.assemble()
        it, reason, rr = prob.solve(pc=hf.PC_ILU0, rtol=1e-12, maxits=2000)
        assert reason > 0
        out.append((it, prob.state()))
    assert out[0][0] == out[1][0] and out[0][0] > 16, out[0][0]
    for a, b in zip(out[0][1], out[1][1]):
        assert np.array_equal(a, b)
