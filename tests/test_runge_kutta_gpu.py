"""Runge-Kutta stages and explicit Euler on the u rows on the GPU
(RungeKutta::apply / computeStage / computeSolution,
src/operator/RungeKutta.cpp:113-253; Euler::apply with isExplicit = true,
src/operator/Euler.cpp:30-35) against the oracle (pinned in
test_runge_kutta_oracle.py): local matrices and condensed blocks of both
element kernels at a stage, whole time steps through the trace solve, and the
argument checks.

The GPU condenses each scheme's u rows divided by a constant (a_ss dt, or dt
when a_ss = 0), which leaves S, S0, U, Q, U0 and Q0 unchanged; the local
matrices it exports are scaled back to the reference's rows."""
import numpy as np
import pytest

import hyperfox_amd as hf
import oracle as orc
from hdg_helpers import rel
from test_gpu_parity import CASES, oracle_local, oracle_system, run_gpu, small_workload
from test_runge_kutta_oracle import oracle_rk_steps

pytestmark = pytest.mark.gpu

DT = 0.05
# stage 0 with a_00 = 0 (explicit), gamma, 1/2, and an explicit table
STAGE0 = ['SSPRK3', 'ALX2', 'RK43', 'CrankNicolson']


def random_fields(w, seed):
    rng = np.random.default_rng(seed)
    s, ne = w.mesh.sizes, w.mesh.n_cells
    return (rng.normal(size=(ne, s['nN'])), rng.normal(size=(ne, s['nN'] * w.dim)),
            rng.normal(size=(w.mesh.n_faces, s['nNf'])))


def set_fields(prob, u, q, lam):
    prob.set_solution(u)
    prob.set_flux(q)
    prob.set_trace(lam)


@pytest.mark.parametrize('kind', [1, 2])
@pytest.mark.parametrize('dim,p', [(2, 1), (2, 3), (2, 5), (3, 2), (3, 3)])
def test_local_matrices_with_explicit_euler_match_oracle(dim, p, kind):
    w = small_workload(dim, p, kind)
    prob, f = run_gpu(w)
    uo = random_fields(w, 3)[0]
    prob.set_solution(uo)
    prob.set_time_scheme(hf.TS_FEULER, DT)
    K, F = prob.local_matrices(0, w.mesh.n_cells)
    for e in range(w.mesh.n_cells):
        Ko, Fo = orc.apply_euler_explicit(dim, p, DT, uo[e], *oracle_local(w, f, e))
        assert rel(K[e], Ko) < 1e-12, (e, rel(K[e], Ko))
        assert rel(F[e], Fo) < 1e-12, (e, rel(F[e], Fo))


@pytest.mark.parametrize('name', STAGE0)
@pytest.mark.parametrize('dim,p', [(2, 2), (2, 4), (3, 2), (3, 3)])
def test_local_matrices_at_a_stage_match_oracle(dim, p, name):
    w = small_workload(dim, p, 2)
    prob, f = run_gpu(w)
    u, q, lam = random_fields(w, 4)
    set_fields(prob, u, q, lam)
    prob.set_time_scheme(hf.TS_RK, DT)
    prob.set_butcher(name)
    prob.rk_start_step()
    K, F = prob.local_matrices(0, w.mesh.n_cells)
    ut = oracle_system(w, f).local_vectors(u, q, lam)
    t = np.asarray(orc.RK_TABLES[name])
    kst = np.zeros((t.shape[0] - 1, ut.shape[1]))
    for e in range(w.mesh.n_cells):
        Ko, Fo = orc.apply_rk(dim, p, DT, t[0, 1:], 0, ut[e], kst, *oracle_local(w, f, e))
        assert rel(K[e], Ko) < 1e-12, (e, rel(K[e], Ko))
        assert rel(F[e], Fo) < 1e-11, (e, rel(F[e], Fo))


def check_condensed(g, c, e):
    for k in ('S', 'U', 'Q'):
        assert rel(g[k][e], c[k]) < 1e-11, (e, k, rel(g[k][e], c[k]))
    for k in ('S0', 'U0', 'Q0'):
        assert np.allclose(g[k][e], c[k], rtol=1e-10, atol=1e-11 * max(1, np.abs(c[k]).max())), (e, k)


@pytest.mark.parametrize('kernel', [1, 0])
@pytest.mark.parametrize('name', STAGE0 + ['FEulerScheme'])
@pytest.mark.parametrize('dim,p', CASES)
def test_condensed_blocks_at_a_stage_match_oracle(dim, p, name, kernel):
    w = small_workload(dim, p, 2)
    prob, f = run_gpu(w)
    prob.set_kernel(kernel)
    u, q, lam = random_fields(w, 6)
    set_fields(prob, u, q, lam)
    if name == 'FEulerScheme':
        prob.set_time_scheme(hf.TS_FEULER, DT)
    else:
        prob.set_time_scheme(hf.TS_RK, DT)
        prob.set_butcher(name)
        prob.rk_start_step()
        ut = oracle_system(w, f).local_vectors(u, q, lam)
        t = np.asarray(orc.RK_TABLES[name])
        kst = np.zeros((t.shape[0] - 1, ut.shape[1]))
    prob.assemble()
    g = prob.condensed()
    for e in range(w.mesh.n_cells):
        if name == 'FEulerScheme':
            Ko, Fo = orc.apply_euler_explicit(dim, p, DT, u[e], *oracle_local(w, f, e))
        else:
            Ko, Fo = orc.apply_rk(dim, p, DT, t[0, 1:], 0, ut[e], kst, *oracle_local(w, f, e))
        check_condensed(g, orc.condense(dim, p, Ko, Fo), e)


@pytest.mark.parametrize('name', ['FEuler', 'SSPRK3', 'RK4', 'CrankNicolson', 'ALX2', 'RK43', 'KS2'])
@pytest.mark.parametrize('dim,p', [(2, 3), (3, 2)])
def test_time_steps_match_oracle(dim, p, name):
    """Two whole steps (every stage: assemble, BiCGStab to a true residual of
    1e-13, computeStage; then computeSolution) against the oracle's steps with
    dense solves."""
    w = small_workload(dim, p, 2)
    prob, f = run_gpu(w)
    u, q, lam = random_fields(w, 9)
    set_fields(prob, u, q, lam)
    prob.set_time_scheme(hf.TS_RK, DT)
    prob.set_butcher(name)
    prob.set_krylov(hf.KRYLOV_BICGSTAB)
    s = oracle_system(w, f)
    fo = {'Solution': u, 'Flux': q, 'Trace': lam}
    for step in range(2):
        infos = prob.rk_step(pc=hf.PC_BJACOBI, rtol=1e-13, maxits=20000)
        assert all(r == 2 for _, r, _ in infos), infos
        assert prob.rk_stages() == (len(infos), 0)
        fo = oracle_rk_steps(s, name, DT, 1, fo['Solution'], fo['Flux'], fo['Trace'])
        ug, qg, lg = prob.solution()
        d = (rel(ug, fo['Solution']), rel(qg, fo['Flux']), rel(lg, fo['Trace'].reshape(-1)))
        print('%s step %d: stage its %s, rel diff u %.1e q %.1e lambda %.1e' % ((name, step, [i for i, _, _ in infos]) + d))
        assert max(d) < 1e-9, d


@pytest.mark.parametrize('dim,p', [(2, 2), (3, 2)])
def test_explicit_euler_steps_match_oracle(dim, p):
    w = small_workload(dim, p, 2)
    prob, f = run_gpu(w)
    u = random_fields(w, 11)[0]
    prob.set_solution(u)
    prob.set_time_scheme(hf.TS_FEULER, DT)
    prob.set_krylov(hf.KRYLOV_BICGSTAB)
    s = oracle_system(w, f)
    for step in range(2):
        prob.assemble()
        its, reason, _ = prob.solve(pc=hf.PC_BJACOBI, rtol=1e-13, maxits=20000)
        assert reason == 2
        s.set_time_scheme(2, DT, u_old=u)
        u = s.reconstruct(s.solve(method=1)[0])[0]
        assert rel(prob.solution()[0], u) < 1e-9, (step, rel(prob.solution()[0], u))


def test_heat_eigenfunction_decays_at_the_stability_function_on_gpu():
    """The oracle's closed-form pin run on the GPU path: ALX2 and RK43 steps
    of u_t = Laplace u from the discrete eigenfunction state."""
    from test_runge_kutta_oracle import heat, stability
    n, p, dt, steps = 8, 3, 0.02, 3
    s, u0, q0, l0 = heat(n, p)
    m = hf.Mesh.structured(2, p, n, 0.1, 11)
    for name in ('ALX2', 'RK43'):
        prob = hf.HDGProblem(2, p)
        prob.set_mesh(m)
        prob.set_model(1)
        prob.set_fields(np.ones((m.n_faces, m.sizes['nNf'])), 1, dirichlet=np.zeros((m.n_faces, m.sizes['nNf'])))
        prob.set_source(None, None)
        set_fields(prob, u0, q0, l0)
        prob.set_time_scheme(hf.TS_RK, dt)
        prob.set_butcher(name)
        prob.set_krylov(hf.KRYLOV_BICGSTAB)
        for _ in range(steps):
            prob.rk_step(pc=hf.PC_BJACOBI, rtol=1e-13, maxits=20000)
        amp = stability(name, -2 * np.pi ** 2 * dt) ** steps
        u = prob.solution()[0]
        err = np.abs(u - amp * u0).max() / (amp * np.abs(u0).max())
        print('%s: relative error %.2e' % (name, err))
        assert err < 2e-4


def test_time_scheme_arguments_are_checked():
    w = small_workload(2, 2, 1)
    prob, f = run_gpu(w)
    with pytest.raises(hf.HfxError, match='time step'):
        prob.set_time_scheme(hf.TS_RK, 0.0)
    prob.set_time_scheme(hf.TS_RK, 0.1)
    with pytest.raises(hf.HfxError, match='Butcher table'):
        prob.assemble()
    with pytest.raises(hf.HfxError, match='upper triangular'):
        prob.set_butcher([[1.0, 1.0], [0.0, 1.0]])
    with pytest.raises(hf.HfxError, match='square'):
        prob.set_butcher(np.zeros((2, 3)))
    prob.set_butcher([[0, 0, 0], [0.5, 0.5, 0], [0, 0, 1]])   # explicit midpoint as a custom table
    with pytest.raises(hf.HfxError, match='start the time step'):
        prob.assemble()
    prob.rk_start_step()
    with pytest.raises(hf.HfxError, match='all stages'):
        prob.rk_compute_solution()
    prob.rk_compute_stage()
    prob.rk_compute_stage()
    with pytest.raises(hf.HfxError, match='more stages'):
        prob.rk_compute_stage()
    with pytest.raises(hf.HfxError, match='compute the solution'):
        prob.assemble()
    prob.rk_compute_solution()
    assert prob.rk_stages() == (2, 0)


@pytest.mark.parametrize('hdg_type', [1, 2], ids=['wexplicit', 'sexplicit'])
@pytest.mark.parametrize('kernel', [1, 0])
@pytest.mark.parametrize('dim,p', [(2, 3), (3, 2)])
def test_runge_kutta_stage_with_the_explicit_hdg_modes(dim, p, kernel, hdg_type):
    """The reference runs its Runge-Kutta schemes with the explicit HDG modes
    (tests/parallel/TestParHDGConvectionDiffusionReactionSource.cpp:72-78,
    109, 226): the u-row transform and S = Sll, S0 from the Solution and
    Flux fields compose."""
    w = small_workload(dim, p, 2)
    prob, f = run_gpu(w)
    prob.set_kernel(kernel)
    u, q, lam = random_fields(w, 21)
    set_fields(prob, u, q, lam)
    prob.set_solver_type(hdg_type)
    prob.set_time_scheme(hf.TS_RK, DT)
    prob.set_butcher('ALX2')
    prob.rk_start_step()
    prob.assemble()
    g = prob.condensed()
    ut = oracle_system(w, f).local_vectors(u, q, lam)
    t = np.asarray(orc.RK_TABLES['ALX2'])
    kst = np.zeros((t.shape[0] - 1, ut.shape[1]))
    for e in range(w.mesh.n_cells):
        Ko, Fo = orc.apply_rk(dim, p, DT, t[0, 1:], 0, ut[e], kst, *oracle_local(w, f, e))
        c = orc.condense(dim, p, Ko, Fo)
        c['S'], c['S0'] = orc.explicit_blocks(dim, p, Ko, Fo, u[e], q[e])
        check_condensed(g, c, e)


@pytest.mark.parametrize('name', ['RK43', 'SSPRK3'])
@pytest.mark.parametrize('dim,p', [(2, 3), (3, 2)])
def test_runge_kutta_stage_on_curved_cells(dim, p, name):
    from test_boundary_models_gpu import curved_workload
    w = curved_workload(dim, p)
    prob, f = run_gpu(w)
    assert prob.kernel_info()['kernel'] == 'general'
    u, q, lam = random_fields(w, 23)
    set_fields(prob, u, q, lam)
    prob.set_time_scheme(hf.TS_RK, DT)
    prob.set_butcher(name)
    prob.rk_start_step()
    prob.assemble()
    g = prob.condensed()
    ut = oracle_system(w, f).local_vectors(u, q, lam)
    t = np.asarray(orc.RK_TABLES[name])
    kst = np.zeros((t.shape[0] - 1, ut.shape[1]))
    for e in range(w.mesh.n_cells):
        Ko, Fo = orc.apply_rk(dim, p, DT, t[0, 1:], 0, ut[e], kst, *oracle_local(w, f, e))
        check_condensed(g, orc.condense(dim, p, Ko, Fo), e)


def test_batches_take_implicit_euler_but_reject_the_general_transform():
    import hyperfox_amd.problems as pr
    w = pr.diffusion_source_2d(4, 2)          # straight cells, identity diffusion: batchable
    prob, f = run_gpu(w)
    b = hf.Batch()
    b.add(prob)
    prob.set_solution(random_fields(w, 3)[0])
    prob.set_time_scheme(hf.TS_BEULER, DT)
    b.assemble(use_graph=False)
    prob.set_time_scheme(hf.TS_RK, DT)
    prob.set_butcher('ALX2')
    prob.rk_start_step()
    with pytest.raises(hf.HfxError, match='implicit Euler only'):
        b.assemble(use_graph=False)
