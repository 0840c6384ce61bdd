"""Runge-Kutta stages and explicit Euler on the u rows in the CPU oracle
(RungeKutta::apply / computeStage / computeSolution,
src/operator/RungeKutta.cpp:113-253, Butcher tables :197-285; Euler::apply
with isExplicit = true, src/operator/Euler.cpp:30-35), pinned by what the
mathematics fixes.  No GPU.

* Every Butcher table of the database satisfies c = A 1 and the order
  conditions of its order (a mistyped entry breaks one of them).
* The reference's own scalar test (tests/regression/RungeKutta/
  TestRungeKutta.cpp:19-91: u rows = -M, so u' = u) made exact: one step of
  every table multiplies u by the stability function
  R(z) = 1 + z b^T (I - z A)^-1 1 at z = dt.
* Fixed points that involve the q and trace columns: when F_u = K_u uj at
  the stage predictor uj = ut + dt sum_j a_sj k_j, the stage solution is ut;
  explicit Euler keeps a consistent (u, q, lambda) fixed.  A dropped
  auxiliary column or a transposed coefficient breaks them.
* Heat equation u_t = Laplace u, u = 0 on the boundary, through the whole
  HDG system, from the discrete state that approximates the eigenfunction
  sin(pi x) sin(pi y): after k steps (u, q, lambda) = R(-2 pi^2 dt)^k times
  the initial state up to the space error, for implicit tables whose R
  differ by far more than that error.
"""
import numpy as np
import pytest

import hyperfox_amd as hf
import oracle as orc

ORDER = {'FEuler': 1, 'EMidpoint': 2, 'Heun': 2, 'Kutta3': 3, 'Heun3': 3, 'SSPRK3': 3, 'RK4': 4,
         'BEuler': 1, 'IMidpoint': 2, 'CrankNicolson': 2, 'KS2': 1, 'QZ2': 2, 'ALX2': 2, 'RK43': 3}
EXPLICIT = ['FEuler', 'EMidpoint', 'Heun', 'Kutta3', 'Heun3', 'SSPRK3', 'RK4']


def split(name):
    t = np.asarray(orc.RK_TABLES[name], dtype=np.float64)
    return t[:-1, 0], t[:-1, 1:], t[-1, 1:]


def stability(name, z):
    c, A, b = split(name)
    one = np.ones(len(b))
    return 1.0 + z * b @ np.linalg.solve(np.eye(len(b)) - z * A, one)


def test_table_names_follow_rktype():
    assert orc.RK_NAMES == list(ORDER)


@pytest.mark.parametrize('name', list(ORDER))
def test_tables_satisfy_their_order_conditions(name):
    c, A, b = split(name)
    assert np.allclose(A.sum(axis=1), c, rtol=0, atol=1e-15)       # c = A 1
    assert np.allclose(np.triu(A, 1), 0)                          # diagonally implicit at most
    if name in EXPLICIT:
        assert np.allclose(np.diag(A), 0)
    else:
        assert np.any(np.diag(A) != 0)
    assert np.isclose(b.sum(), 1.0, rtol=0, atol=1e-15)
    o = ORDER[name]
    if name == 'KS2':
        # the reference's KS2 entries give b.c = 2: order 1 as tabulated
        assert np.isclose(b @ c, 2.0)
    if o >= 2:
        assert np.isclose(b @ c, 0.5, rtol=0, atol=1e-15)
    else:
        assert not np.isclose(b @ c, 0.5) or name == 'FEuler'
    if o >= 3:
        assert np.isclose(b @ c ** 2, 1 / 3, rtol=0, atol=1e-15)
        assert np.isclose(b @ A @ c, 1 / 6, rtol=0, atol=1e-15)
    elif o == 2:
        assert not (np.isclose(b @ c ** 2, 1 / 3) and np.isclose(b @ A @ c, 1 / 6))
    if o >= 4:
        assert np.isclose(b @ c ** 3, 1 / 4) and np.isclose(b @ (c * (A @ c)), 1 / 8)
        assert np.isclose(b @ A @ c ** 2, 1 / 12) and np.isclose(b @ A @ A @ c, 1 / 24)


def one_cell(dim, p, seed=0):
    rng = np.random.default_rng(seed)
    s = orc.sizes(dim, p)
    m = hf.Mesh.structured(dim, p, 1, 0.0, 1)
    a = m.arrays()
    x = a['coords'][a['cells'][0]]
    tau = rng.uniform(0.5, 2, s['nF'] * s['nNf'])
    b = rng.normal(size=(s['nN'], dim))
    K, F = orc.local_system(dim, p, 2, x, D=np.full(s['nN'], 0.7), ndD=1, b=b, tau=tau,
                            f_ip=rng.normal(size=s['nIP']))
    return s, K, F, rng


def rk_local_step(name, dim, p, K, F, u0, dt):
    """One step of the reference's scalar test on one cell: stage solves of
    the u rows alone (their other columns are zero), the stage bookkeeping of
    RungeKutta on {Solution, Flux, Trace} with zero auxiliary fields."""
    s = orc.sizes(dim, p)
    n, nN = s['n_local'], s['nN']
    rk = orc.RungeKutta(name, dt)
    fields = {'Solution': u0.copy(), 'Flux': np.zeros(nN * dim), 'Trace': np.zeros(n - nN * (dim + 1))}
    rk.start_step(fields)
    for st in range(rk.nst):
        ut = np.concatenate([rk.old[k] for k in ('Solution', 'Flux', 'Trace')])
        kst = np.stack([np.concatenate([rk.k[k][j] for k in ('Solution', 'Flux', 'Trace')]) for j in range(rk.nst)])
        K2, F2 = orc.apply_rk(dim, p, dt, rk.a_row(st), st, ut, kst, K, F)
        assert np.all(K2[:nN, nN:] == 0.0)
        fields['Solution'][:] = np.linalg.solve(K2[:nN, :nN], F2[:nN])
        rk.compute_stage(fields)
    rk.compute_solution(fields)
    return fields['Solution']


@pytest.mark.parametrize('name', list(ORDER))
@pytest.mark.parametrize('dt', [0.2, 0.05])
def test_scalar_ode_step_is_the_stability_function(name, dt):
    dim, p = 2, 2
    s, K, F, rng = one_cell(dim, p)
    nN = s['nN']
    M = K[nN::dim, nN::dim][:nN, :nN]           # the q and trace rows stay those of a real cell
    K[:nN] = 0.0
    K[:nN, :nN] = -M                             # stiffness = -M: u' = u (TestRungeKutta.cpp:72)
    F = np.zeros_like(F)
    u = rng.uniform(0.5, 1.5, nN)
    R = stability(name, dt)
    for _ in range(3):
        u2 = rk_local_step(name, dim, p, K, F, u, dt)
        assert np.allclose(u2, R * u, rtol=1e-13, atol=0)
        u = u2
    # the reference's regression bound (error against exp(t) < 1) holds a fortiori
    assert abs(R ** 3 - np.exp(3 * dt)) < 1.0


@pytest.mark.parametrize('dim,p', [(2, 3), (3, 2)])
@pytest.mark.parametrize('name', ['SSPRK3', 'RK4', 'ALX2', 'RK43', 'KS2'])
def test_stage_solution_is_ut_when_the_predictor_is_steady(dim, p, name):
    s, K, F, rng = one_cell(dim, p, seed=5)
    n, nN = s['n_local'], s['nN']
    table = np.asarray(orc.RK_TABLES[name])
    nst, dt = table.shape[0] - 1, 0.3
    ut = rng.normal(size=n)
    kst = rng.normal(size=(nst, n))
    for st in range(nst):
        a = table[st, 1:]
        uj = ut + dt * (a[:st] @ kst[:st]) if st else ut.copy()
        F1 = F.copy()
        F1[:nN] = K[:nN] @ uj                    # the model's u rows vanish at the predictor
        K2, F2 = orc.apply_rk(dim, p, dt, a, st, ut, kst, K, F1)
        assert np.allclose(K2[:nN] @ ut, F2[:nN], rtol=0, atol=1e-12 * np.abs(F2[:nN]).max())
        assert np.array_equal(K2[nN:], K[nN:]) and np.array_equal(F2[nN:], F1[nN:])
        if a[st] != 0:  # u rows: a_ss dt K + [M 0 0]
            M = K[nN::dim, nN::dim][:nN, :nN]
            assert np.allclose(K2[:nN, nN:], a[st] * dt * K[:nN, nN:], rtol=1e-14, atol=1e-300)
            assert np.allclose(K2[:nN, :nN], a[st] * dt * K[:nN, :nN] + M, rtol=1e-12, atol=1e-15)
        # a stage vector the row does not weigh (j >= st) changes nothing
        k3 = kst.copy()
        k3[st:] = rng.normal(size=k3[st:].shape)
        K3, F3 = orc.apply_rk(dim, p, dt, a, st, ut, k3, K, F1)
        assert np.array_equal(K3, K2) and np.array_equal(F3, F2)


@pytest.mark.parametrize('dim,p', [(2, 2), (3, 2)])
def test_explicit_euler_rows_and_fixed_point(dim, p):
    s, K, F, rng = one_cell(dim, p, seed=7)
    n, nN = s['n_local'], s['nN']
    dt = 0.05
    x = rng.normal(size=n)                       # a consistent (u, q, lambda)
    F1 = F.copy()
    F1[:nN] = K[:nN] @ x
    K2, F2 = orc.apply_euler_explicit(dim, p, dt, x[:nN], K, F1)
    M = K[nN::dim, nN::dim][:nN, :nN]
    assert np.array_equal(K2[:nN, :nN], M)
    assert np.allclose(K2[:nN, nN:], dt * K[:nN, nN:], rtol=1e-15, atol=1e-300)
    assert np.array_equal(K2[nN:], K[nN:]) and np.array_equal(F2[nN:], F1[nN:])
    assert np.allclose(K2[:nN] @ x, F2[:nN], rtol=0, atol=1e-12 * np.abs(F2[:nN]).max())
    # scalar test: u' = u gives u (1 + dt), as FEuler does
    K[:nN] = 0.0
    K[:nN, :nN] = -M
    u = rng.uniform(0.5, 1.5, nN)
    K3, F3 = orc.apply_euler_explicit(dim, p, dt, u, K, np.zeros(n))
    assert np.allclose(np.linalg.solve(K3[:nN, :nN], F3[:nN]), (1 + dt) * u, rtol=1e-13)
    assert np.allclose(rk_local_step('FEuler', dim, p, K, np.zeros(n), u, dt), (1 + dt) * u, rtol=1e-13)
    with pytest.raises(orc.OracleError, match='time step'):
        orc.apply_euler_explicit(dim, p, 0.0, u, K, F)


def test_custom_tables_and_stage_counts_are_checked():
    with pytest.raises(orc.OracleError, match='square'):
        orc.RungeKutta(np.zeros((2, 3)), 0.1)
    with pytest.raises(orc.OracleError, match='upper triangular'):
        orc.RungeKutta([[1, 1], [0, 1]], 0.1)   # setButcherTable(EMatrix) takes explicit tables only
    rk = orc.RungeKutta([[0, 0, 0], [0.5, 0.5, 0], [0, 0, 1]], 0.1)
    f = {'Solution': np.ones(3)}
    rk.start_step(f)
    with pytest.raises(orc.OracleError, match='all stages'):
        rk.compute_solution(f)
    rk.compute_stage(f)
    rk.compute_stage(f)
    with pytest.raises(orc.OracleError, match='more stages'):
        rk.compute_stage(f)
    with pytest.raises(orc.OracleError, match='time step'):
        orc.apply_rk(2, 1, 0.0, [1.0], 0, np.zeros(15), np.zeros((1, 15)), np.eye(15), np.zeros(15))


def heat(n, p):
    """Heat system (no source) and a discrete initial state: the HDG solution
    (u, q, lambda) of -Laplace u = 2 pi^2 sin(pi x) sin(pi y), which is close
    to an eigenvector of the discrete operator, so that every stage (the
    explicit ones read q and lambda too) sees a consistent state."""
    m = hf.Mesh.structured(2, p, n, 0.1, 11)
    a = m.arrays()
    nNf = m.sizes['nNf']
    tau, g = np.ones((m.n_faces, nNf)), np.zeros((m.n_faces, nNf))
    xip = np.stack([orc.ip_coords(2, p, a['coords'][c]) for c in a['cells']])
    f_ip = 2 * np.pi ** 2 * np.sin(np.pi * xip[..., 0]) * np.sin(np.pi * xip[..., 1])
    st = orc.System(2, p, 1, a, tau, tau_sides=1, f_ip=f_ip, dirichlet=g)
    l0 = st.solve(method=1)[0]
    u0, q0 = st.reconstruct(l0)
    s = orc.System(2, p, 1, a, tau, tau_sides=1, dirichlet=g)
    return s, u0, q0, l0.reshape(m.n_faces, nNf)


def oracle_rk_steps(s, name, dt, steps, u, q, lam):
    rk = orc.RungeKutta(name, dt)
    fields = {'Solution': u.copy(), 'Flux': q.copy(), 'Trace': lam.copy()}
    for _ in range(steps):
        rk.start_step(fields)
        old = s.local_vectors(rk.old['Solution'], rk.old['Flux'], rk.old['Trace'])
        for st in range(rk.nst):
            stl = np.stack([s.local_vectors(rk.k['Solution'][j], rk.k['Flux'][j], rk.k['Trace'][j])
                            for j in range(rk.nst)])
            s.set_time_scheme(3, dt, rk.a_row(st), st, old_local=old, stage_local=stl)
            lam_s = s.solve(method=1)[0]
            u_s, q_s = s.reconstruct(lam_s)
            fields['Solution'][:], fields['Flux'][:], fields['Trace'][:] = u_s, q_s, lam_s.reshape(lam.shape)
            rk.compute_stage(fields)
        rk.compute_solution(fields)
    return fields


@pytest.mark.parametrize('name,other', [('BEuler', 'CrankNicolson'), ('CrankNicolson', 'BEuler'),
                                        ('ALX2', 'BEuler'), ('RK43', 'BEuler'), ('IMidpoint', 'ALX2')])
def test_heat_eigenfunction_decays_at_the_stability_function(name, other):
    dt, steps, tol = 0.02, 3, 2e-4
    z = -2 * np.pi ** 2 * dt
    amp = stability(name, z) ** steps
    assert abs(stability(other, z) ** steps - amp) > 5 * tol * amp  # the test tells the two apart
    s, u0, q0, l0 = heat(8, 3)
    f = oracle_rk_steps(s, name, dt, steps, u0, q0, l0)
    err = np.abs(f['Solution'] - amp * u0).max() / (amp * np.abs(u0).max())
    print('%s: relative error %.2e' % (name, err))
    assert err < tol
    # the flux decays with the same factor (q = grad u)
    assert np.abs(f['Flux'] - amp * q0).max() < 10 * tol * amp * np.abs(q0).max()
    assert np.abs(f['Trace'] - amp * l0).max() < 10 * tol * amp * np.abs(l0).max()
