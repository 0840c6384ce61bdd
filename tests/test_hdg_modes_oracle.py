"""Explicit HDG modes (HDGSolverOpts::type WEXPLICIT / SEXPLICIT,
src/solver/HDGSolver.cpp:346-353 and 709-729) in the CPU oracle, pinned by
what the mathematics fixes.  No GPU.

With S = Sll and S0 = F_lambda - Slu u - Slq q, the trace equation of a face
is the flux balance with u and q given.  When (u, q) is an exact solution
that the space holds (u linear, q = grad u constant, smooth coefficients),
the balance holds for lambda = u on every face, so the explicit trace solve
must return the face values of u -- whatever tau, D and b are.  And the
trace matrix of the explicit modes couples no two faces (block diagonal).
"""
import numpy as np
import pytest

import hyperfox_amd as hf
import hyperfox_amd.problems as pr
import oracle as orc


def linear_case(dim, p, kind, seed=3):
    n = 3 if dim == 2 else 2
    m = hf.Mesh.structured(dim, p, n, 0.15 if dim == 2 else 0.05, seed)
    a = m.arrays()
    rng = np.random.default_rng(seed)
    c0, g = rng.normal(), rng.normal(size=dim)
    x = a['coords']
    nNf = m.sizes['nNf']
    u_nodes = c0 + x @ g
    faces = a['faces']
    dirichlet = u_nodes[faces]                        # exact traces (used on the boundary)
    tau = rng.uniform(0.5, 2.0, (m.n_faces, nNf))
    kw = {}
    if kind == 2:
        b = np.stack([np.sin(x[:, 1]), np.cos(x[:, 0])] + ([np.sin(x[:, 0] + x[:, 1])] if dim == 3 else []), 1)
        kw = dict(D=0.3 + 0.1 * np.cos(x[:, 0]), ndD=1, b=b)
    s = orc.System(dim, p, kind, a, tau, tau_sides=1, dirichlet=dirichlet, **kw)
    sol = u_nodes[a['cells']]                         # cell-wise nodal u
    flux = np.tile(g, (m.n_cells, m.sizes['nN'], 1)).reshape(m.n_cells, -1)
    return dict(mesh=m, oracle=s, sol=sol, flux=flux, lam=dirichlet, tau=tau, fields=kw)


@pytest.mark.parametrize('kind', [1, 2])
@pytest.mark.parametrize('dim,p', [(2, 1), (2, 3), (3, 1), (3, 2)])
def test_explicit_trace_of_an_exact_linear_solution(dim, p, kind):
    c = linear_case(dim, p, kind)
    s, sol, flux, lam_exact = c['oracle'], c['sol'], c['flux'], c['lam']
    for mode in (1, 2):
        s.set_hdg_type(mode, sol, flux)
        lam = s.solve(method=1)[0]
        err = np.abs(lam - lam_exact.reshape(-1)).max()
        assert err <= 1e-11 * max(1.0, np.abs(lam_exact).max()), (mode, err)


@pytest.mark.parametrize('dim,p', [(2, 2), (3, 2)])
def test_explicit_trace_matrix_is_block_diagonal(dim, p):
    c = linear_case(dim, p, 2)
    m, s, sol, flux = c['mesh'], c['oracle'], c['sol'], c['flux']
    rng = np.random.default_rng(1)
    s.set_hdg_type(1, rng.normal(size=sol.shape), rng.normal(size=flux.shape))
    rowptr, cols, vals, rhs = s.csr()
    nNf = m.sizes['nNf']
    rows = np.repeat(np.arange(rowptr.size - 1), np.diff(rowptr))
    off = rows // nNf != cols // nNf
    assert np.all(vals[off] == 0.0) and np.abs(vals[~off]).max() > 0
    # the implicit mode of the same system couples neighbouring faces
    s.set_hdg_type(0)
    assert np.abs(s.csr()[2][off]).max() > 1e-3
