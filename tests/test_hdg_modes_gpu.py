"""Explicit HDG modes (WEXPLICIT / SEXPLICIT, src/solver/HDGSolver.cpp:346-353,
709-729) on the GPU against the oracle (pinned in test_hdg_modes_oracle.py):
condensed S and S0 of both element kernels, the trace system, the two solve
paths, and the exact trace of a linear solution."""
import numpy as np
import pytest

import hyperfox_amd as hf
import oracle as orc
from hdg_helpers import rel
from test_gpu_parity import CASES_ALL, condense_bar, oracle_local, oracle_system, run_gpu, small_workload
from test_hdg_modes_oracle import linear_case

pytestmark = pytest.mark.gpu


def old_fields(w, seed=4):
    rng = np.random.default_rng(seed)
    nN = w.mesh.sizes['nN']
    return rng.normal(size=(w.mesh.n_cells, nN)), rng.normal(size=(w.mesh.n_cells, nN * w.dim))


@pytest.mark.parametrize('kernel', [1, 0])
@pytest.mark.parametrize('kind', [0, 2])
@pytest.mark.parametrize('dim,p', CASES_ALL)
def test_explicit_condensed_blocks_match_oracle(dim, p, kind, kernel):
    w = small_workload(dim, p, kind)
    prob, f = run_gpu(w)
    prob.set_kernel(kernel)
    sol, flux = old_fields(w)
    prob.set_solution(sol)
    prob.set_flux(flux)
    prob.set_solver_type(hf.HDG_WEXPLICIT)
    prob.assemble()
    g = prob.condensed()
    bar = condense_bar(w, f, 1e-11)
    for e in range(w.mesh.n_cells):
        Ko, Fo = oracle_local(w, f, e)
        S, S0 = orc.explicit_blocks(dim, p, Ko, Fo, sol[e], flux[e])
        c = orc.condense(dim, p, Ko, Fo)
        assert rel(g['S'][e], S) < 1e-12, (e, rel(g['S'][e], S))
        assert rel(g['S0'][e], S0) < 1e-11, (e, rel(g['S0'][e], S0))
        for k in ('U', 'Q'):
            assert rel(g[k][e], c[k]) < bar, (e, k)


@pytest.mark.parametrize('kind', [1, 2])
@pytest.mark.parametrize('dim,p', [(2, 2), (2, 5), (3, 1), (3, 3)])
def test_explicit_trace_system_and_solves_match_oracle(dim, p, kind):
    w = small_workload(dim, p, kind)
    prob, f = run_gpu(w)
    sol, flux = old_fields(w)
    prob.set_solution(sol)
    prob.set_flux(flux)
    s = oracle_system(w, f)
    s.set_hdg_type(1, sol, flux)
    rpo, colso, valso, rhso = s.csr()
    lam_o = s.solve(method=1)[0]
    for mode in (hf.HDG_WEXPLICIT, hf.HDG_SEXPLICIT):
        prob.set_solution(sol)
        prob.set_flux(flux)
        prob.set_solver_type(mode)
        prob.assemble()
        rp, cols, vals, rhs = prob.trace_csr()
        assert np.array_equal(rp, rpo) and np.array_equal(cols, colso)
        assert rel(vals, valso) < 1e-12 and rel(rhs, rhso) < 1e-12
        its, reason, relres = prob.solve(pc=hf.PC_BJACOBI, rtol=1e-12, maxits=100)
        assert reason == 2 and relres <= 1e-12, (mode, relres)
        lam = prob.solution()[2]
        assert np.abs(lam - lam_o).max() <= 1e-10 * np.abs(lam_o).max(), mode
        if mode == hf.HDG_SEXPLICIT:
            assert its == 1


@pytest.mark.parametrize('kind', [1, 2])
@pytest.mark.parametrize('dim,p', [(2, 3), (3, 2)])
def test_explicit_trace_of_an_exact_linear_solution(dim, p, kind):
    c = linear_case(dim, p, kind)
    m, fl = c['mesh'], c['fields']
    prob = hf.HDGProblem(dim, p)
    prob.set_mesh(m)
    prob.set_model(kind)
    prob.set_fields(c['tau'], 1, fl.get('D'), fl.get('ndD', 0), fl.get('b'), c['lam'])
    prob.set_source(None, None)
    prob.set_solution(c['sol'])
    prob.set_flux(c['flux'])
    prob.set_solver_type(hf.HDG_SEXPLICIT)
    prob.assemble()
    prob.solve(pc=hf.PC_BJACOBI, rtol=1e-12, maxits=10)
    lam, lam_exact = prob.solution()[2], c['lam'].reshape(-1)
    assert np.abs(lam - lam_exact).max() <= 1e-11 * max(1.0, np.abs(lam_exact).max())
