"""The Burgers oracle (HDGBurgersModel, HDGBurgersModel.cpp:86-132, under
HDGSolver + NonLinearWrapper) pinned on the CPU:

* with no solution, no trace and tau = t I, the m = dim system is the scalar
  (m = 1) oracle's HDGBase + HDGDiffusion system, itself pinned on the
  reference's fixtures, Kronecker-multiplied by I_m (HDGDiffusion and the
  q blocks of HDGBase act per component);
* a tau matrix C enters as kron(tau part of the scalar system, C), which fixes
  the orientation of the reference's column-major m x m tau (HDGBase.cpp:
  18-32, 96-121);
* HDGUNabU's right-hand side op [u/2; lambda/2] (HDGUNabU.cpp:180-189) with a
  linear u and its own trace is the strong form: the u rows integrate
  phi_i (u . grad) u_a over the cell and the lambda rows psi_j u_a (u . n)
  over each face (integration by parts of the bulk weak form), evaluated
  here directly from the field, not from any operator;
* the u-u block of HDGUNabU is the exact Jacobian of its bulk quadratic form
  -int u_a div(phi_i u) (central differences of that form, computed here);
* one Source per component integrates f_a phi_i;
* the nonlinear loop on the reference's stationary regression case
  (tests/regression/HDG/TestHDGBurgersStat.cpp: Cole-Hopf solution with
  A = 0.5, x0 = (1, 0), D = 1, tau = 2 D / h I, integrated Dirichlet data,
  zero initial guess, 10 iterations at most, tolerance 1e-6) converges to the
  analytical field, closer on the finer mesh.
"""
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, '..', 'oracle'))
import oracle as orc  # noqa: E402
import burgers_oracle as bo  # noqa: E402

RNG = np.random.default_rng(7)


def cell(dim, p, rng):
    """A random affine cell's nodes (the reference nodes mapped affinely)."""
    nodes = orc.refel(dim, p)['nodes']
    A = np.eye(dim) + 0.3 * rng.standard_normal((dim, dim))
    if np.linalg.det(A) < 0:
        A[:, 0] *= -1
    return nodes @ A.T + rng.standard_normal(dim)


def sizes_m(dim, p):
    s = orc.sizes(dim, p)
    m = dim
    return s, m, s['nN'] * m, s['nN'] * dim * m, s['nF'] * s['nNf'] * m


@pytest.mark.parametrize('dim,p', [(2, 1), (2, 3), (3, 1), (3, 2)])
def test_no_flow_is_the_scalar_system_per_component(dim, p):
    s, m, lu, lq, ll = sizes_m(dim, p)
    x = cell(dim, p, RNG)
    D = 0.5 + RNG.random(s['nN'])
    t = 0.5 + RNG.random(s['nF'] * s['nNf'])
    K1, F1 = orc.local_system(dim, p, 1, x, D=D, ndD=1, tau=t)
    tau = np.einsum('j,ab->jba', t, np.eye(m)).reshape(-1)
    K, F = bo.burgers_local(dim, p, x, D, tau, np.zeros(lu), np.zeros(ll))
    assert np.allclose(K, np.kron(K1, np.eye(m)), rtol=0, atol=1e-12 * np.abs(K1).max())
    assert not F.any() and not F1.any()


@pytest.mark.parametrize('dim,p', [(2, 2), (3, 1)])
def test_tau_matrix_orientation(dim, p):
    s, m, lu, lq, ll = sizes_m(dim, p)
    x = cell(dim, p, RNG)
    t = 0.5 + RNG.random(s['nF'] * s['nNf'])
    C = RNG.standard_normal((m, m))
    K1t, _ = orc.local_system(dim, p, 0, x, tau=t)
    K10, _ = orc.local_system(dim, p, 0, x, tau=np.zeros_like(t))
    tau = np.einsum('j,ab->jba', t, C).reshape(-1)  # per node: C column-major
    Kt, _ = bo.burgers_local(dim, p, x, None, tau, np.zeros(lu), np.zeros(ll))
    K0, _ = bo.burgers_local(dim, p, x, None, np.zeros_like(tau), np.zeros(lu), np.zeros(ll))
    assert np.allclose(Kt - K0, np.kron(K1t - K10, C), rtol=0, atol=1e-12)
    assert not np.allclose(Kt - K0, np.kron(K1t - K10, C.T), rtol=0, atol=1e-6)


def geometry(dim, p, x):
    """Physical IPs, dV, face IPs, dS and outward normals of an affine cell,
    from the cell's vertices and the reference tables."""
    R = orc.refel(dim, p)
    J = np.einsum('qik,ic->qkc', R['dphi'], x)[0]  # d x / d xi (constant)
    dV = R['ipw'] * abs(np.linalg.det(J))
    xq = R['phi'] @ x
    fm = R['face_nodes']
    faces = []
    for f in range(dim + 1):
        xf = x[fm[f]]
        opp = [i for i in range(dim + 1) if i not in fm[f][:dim]]
        verts = xf[:dim]
        if dim == 2:
            tvec = verts[1] - verts[0]
            area = np.linalg.norm(tvec)
            n = np.array([tvec[1], -tvec[0]]) / area
        else:
            c = np.cross(verts[1] - verts[0], verts[2] - verts[0])
            area = 0.5 * np.linalg.norm(c)
            n = c / np.linalg.norm(c)
        if np.dot(x[opp[0]] - verts[0], n) > 0:
            n = -n
        dS = R['fipw'] * area / R['fipw'].sum()
        faces.append((R['fphi'] @ xf, dS, n))
    return R, xq, dV, faces


@pytest.mark.parametrize('dim,p', [(2, 2), (2, 3), (3, 2)])
def test_convection_rhs_is_the_strong_form(dim, p):
    s, m, lu, lq, ll = sizes_m(dim, p)
    x = cell(dim, p, RNG)
    c0 = RNG.standard_normal(dim)
    G = RNG.standard_normal((dim, dim))  # u(x) = c0 + G x, grad u_a = G[a]
    uf = lambda X: c0 + X @ G.T  # noqa: E731
    R, xq, dV, faces = geometry(dim, p, x)
    sol = uf(x).reshape(-1)
    fm = R['face_nodes']
    trace = np.concatenate([uf(x[fm[f]]).reshape(-1) for f in range(dim + 1)])
    K, F = bo.burgers_local(dim, p, x, None, np.zeros(s['nF'] * s['nNf'] * m * m), sol, trace)
    uq = uf(xq)
    conv = uq @ G.T  # (u . grad) u_a at the IPs
    Fu = np.einsum('q,qi,qa->ia', dV, R['phi'], conv).reshape(-1)
    assert np.allclose(F[:lu], Fu, rtol=0, atol=1e-12 * np.abs(Fu).max())
    Fl = []
    for f, (xfq, dS, n) in enumerate(faces):
        u = uf(xfq)
        Fl.append(np.einsum('q,qj,qa,q->ja', dS, R['fphi'], u, u @ n).reshape(-1))
    Fl = np.concatenate(Fl)
    assert np.allclose(F[lu + lq:], Fl, rtol=0, atol=1e-12 * np.abs(Fl).max())
    assert not F[lu:lu + lq].any()


@pytest.mark.parametrize('dim,p', [(2, 1), (2, 3), (3, 2)])
def test_convection_block_is_the_jacobian_of_its_form(dim, p):
    s, m, lu, lq, ll = sizes_m(dim, p)
    x = cell(dim, p, RNG)
    R = orc.refel(dim, p)
    J = np.einsum('qik,ic->qkc', R['dphi'], x)
    dV = R['ipw'] * np.abs(np.linalg.det(J))
    gphi = np.einsum('qck,qik->qic', np.linalg.inv(J), R['dphi'])  # physical grad phi_i

    def form(u):  # -int u_a div(phi_i u), u nodal (nN x m)
        uq = R['phi'] @ u
        div = np.einsum('qic,ic->q', gphi, u)
        ug = np.einsum('qc,qic->qi', uq, gphi)
        return -np.einsum('q,qa,qi->ia', dV, uq, div[:, None] * R['phi'] + ug).reshape(-1)

    u = RNG.standard_normal((s['nN'], m))
    tau0 = np.zeros(s['nF'] * s['nNf'] * m * m)
    K, _ = bo.burgers_local(dim, p, x, None, tau0, u.reshape(-1), np.zeros(ll))
    K0, _ = bo.burgers_local(dim, p, x, None, tau0, np.zeros(lu), np.zeros(ll))
    h = 1e-3
    Jfd = np.zeros((lu, lu))
    for c in range(lu):
        e = np.zeros(lu)
        e[c] = h
        Jfd[:, c] = (form((u.reshape(-1) + e).reshape(u.shape)) - form((u.reshape(-1) - e).reshape(u.shape))) / (2 * h)
    Kuu = (K - K0)[:lu, :lu]
    assert np.allclose(Kuu, Jfd, rtol=0, atol=1e-9 * np.abs(Jfd).max())


@pytest.mark.parametrize('dim,p', [(2, 2), (3, 1)])
def test_source_per_component(dim, p):
    s, m, lu, lq, ll = sizes_m(dim, p)
    x = cell(dim, p, RNG)
    c = RNG.standard_normal(m)
    f = np.tile(c, (s['nIP'], 1))
    tau0 = np.zeros(s['nF'] * s['nNf'] * m * m)
    _, F = bo.burgers_local(dim, p, x, None, tau0, np.zeros(lu), np.zeros(ll), f)
    R = orc.refel(dim, p)
    J = np.einsum('qik,ic->qkc', R['dphi'], x)[0]
    vol = R['ipw'].sum() * abs(np.linalg.det(J))
    assert np.allclose(F[:lu].reshape(-1, m).sum(0), c * vol, rtol=1e-12, atol=0)
    assert not F[lu:].any()


# ------------------------------------------------ the reference's regression case
def cole_hopf(X, D=1.0, A=0.5, x0=(1.0, 0.0)):
    """analyticalBurgersStat (TestHDGBurgersStat.cpp): u = -2 D grad(pot)/pot."""
    x, y = X[:, 0] - x0[0], X[:, 1] - x0[1]
    ep, em = np.exp(A * x), np.exp(-A * x)
    pot = 0.001 * A * np.exp((1 + x0[0]) * A) * (1 + X[:, 0]) + (ep + em) * np.cos(A * y)
    gx = 0.001 * A * np.exp((1 + x0[0]) * A) + A * (ep - em) * np.cos(A * y)
    gy = -A * (ep + em) * np.sin(A * y)
    return -2.0 * D * np.stack([gx, gy], 1) / pot[:, None]


def regression_case(h, p):
    import hyperfox_amd as hf
    z = np.load(os.path.join(HERE, 'golden', 'meshes', 'regression_dim-2_h-%s_ord-%d.npz' % (h, p)))
    mesh = hf.Mesh.from_cells(2, p, z['nodes'], z['cells']).arrays()
    s = orc.sizes(2, p)
    D = 1.0
    tau = np.zeros((mesh['faces'].shape[0], s['nNf'], 2, 2, 2))
    tau[:] = np.eye(2) * (D / float(h)) * 2.0
    g = cole_hopf(mesh['coords'][mesh['faces'].reshape(-1)]).reshape(-1, s['nNf'], 2)
    sysm = bo.BurgersSystem(2, p, mesh, tau, Dn=np.full(mesh['coords'].shape[0], D), dirichlet=g)
    return mesh, sysm


@pytest.mark.parametrize('p', [1, 2])
def test_stationary_regression_case_converges(p):
    errs = []
    for h in ('2e-1', '1e-1'):
        mesh, sysm = regression_case(h, p)
        nC = mesh['cells'].shape[0]
        z0 = np.zeros((nC, sysm.s['nN'] * 2))
        sol, buf, flux, lam, its, res = sysm.newton(z0, z0, np.zeros(sysm.ndofs), max_iters=10, tol=1e-6)
        assert res < 1e-6 and its <= 10
        ana = cole_hopf(mesh['coords'][mesh['cells'].reshape(-1)]).reshape(nC, -1)
        errs.append(np.sqrt(np.sum((ana - sol) ** 2) / np.sum(ana ** 2)))
    # order p + 1 in h: measured 8.3e-4 -> 2.3e-4 (p = 1), 3.2e-5 -> 5.2e-6 (p = 2), 4 iterations each
    assert errs[0] < [2e-3, 1e-4][p - 1] and errs[1] < errs[0] / 3


# ------------------------------------------------ HDG-type boundary rows, m > 1
def hdg_case(p=2, seed=4):
    import hyperfox_amd as hf
    rng = np.random.default_rng(seed)
    mesh = hf.Mesh.structured(2, p, 2, 0.1, seed)
    a = mesh.arrays()
    s = orc.sizes(2, p)
    tau = np.einsum('fns,ab->fnsab', 1.0 + rng.random((mesh.n_faces, s['nNf'], 2)), 2.0 * np.eye(2))
    g = rng.standard_normal((mesh.n_faces, s['nNf'], 2))
    f = rng.standard_normal((mesh.n_cells, s['nIP'], 2))
    bf = np.nonzero(a['face2cell'][:, 1] < 0)[0]
    return mesh, a, s, tau, g, f, bf, rng


def test_hdg_boundary_identity_is_dirichlet():
    """B = [0 | 0 | I], r = g reproduces DirichletModel's rows (HDGSolver.cpp:493-518)."""
    mesh, a, s, tau, g, f, bf, rng = hdg_case()
    bs = s['nNf'] * 2
    B = np.zeros((bf.size, bs, bs * 4))
    B[:, :, 3 * bs:] = np.eye(bs)
    hdg = dict(faces=bf, B=B, r=g[bf].reshape(bf.size, bs), mset=True, rset=True)
    z = np.zeros((mesh.n_cells, s['nN'] * 2))
    A1, b1 = bo.BurgersSystem(2, 2, a, tau, Dn=np.ones(mesh.n_nodes), f_ip=f, dirichlet=g,
                              boundary='dirichlet').assemble(z, np.zeros(mesh.n_faces * bs))
    A2, b2 = bo.BurgersSystem(2, 2, a, tau, Dn=np.ones(mesh.n_nodes), f_ip=f, dirichlet=g,
                              hdg=hdg).assemble(z, np.zeros(mesh.n_faces * bs))
    assert np.array_equal(A1.toarray(), A2.toarray()) and np.array_equal(b1, b2)


@pytest.mark.parametrize('mset', [True, False])
def test_hdg_boundary_rows_hold_after_the_solve(mset):
    """With random B (Set) the solved fields satisfy B_u u_F + B_q q_F +
    B_lambda lambda_F = r on every such face; with Add the rows are the cell's
    own plus those terms."""
    mesh, a, s, tau, g, f, bf, rng = hdg_case()
    bs = s['nNf'] * 2
    B = rng.standard_normal((bf.size, bs, bs * 4))
    B[:, :, 3 * bs:] += 4.0 * np.eye(bs)
    r = rng.standard_normal((bf.size, bs))
    st = rng.standard_normal((mesh.n_cells, s['nN'] * 2)) * 0.3
    tr = rng.standard_normal(mesh.n_faces * bs) * 0.3
    sysm = bo.BurgersSystem(2, 2, a, tau, Dn=np.ones(mesh.n_nodes), f_ip=f, dirichlet=g,
                            hdg=dict(faces=bf, B=B, r=r, mset=mset, rset=mset))
    A, b = sysm.assemble(st, tr)
    lam = sysm.solve()
    u, q = sysm.reconstruct(lam)
    fmap = orc.refel(2, 2)['face_nodes']
    if not mset:
        plain = bo.BurgersSystem(2, 2, a, tau, Dn=np.ones(mesh.n_nodes), f_ip=f, dirichlet=g)
        plain.bfaces = np.array([], dtype=int)
        A0, b0 = plain.assemble(st, tr)
    for ib, F in enumerate(bf):
        e = a['face2cell'][F, 0]
        lf = list(a['cell2face'][e]).index(F)
        node_at = {int(sysm.pos[e, lf, j]): int(fmap[lf][j]) for j in range(s['nNf'])}
        uF = np.array([u[e][node_at[pj] * 2 + c] for pj in range(s['nNf']) for c in range(2)])
        qF = np.array([q[e][(node_at[pj] * 2 + k) * 2 + c] for pj in range(s['nNf']) for k in range(2) for c in range(2)])
        lF = lam[F * bs:(F + 1) * bs]
        lhs = B[ib][:, :bs] @ uF + B[ib][:, bs:3 * bs] @ qF + B[ib][:, 3 * bs:] @ lF
        if mset:
            assert np.allclose(lhs, r[ib], rtol=0, atol=1e-9 * max(1.0, np.abs(r[ib]).max()))
        else:
            own = A0[F * bs:(F + 1) * bs, :] @ lam - b0[F * bs:(F + 1) * bs]
            assert np.allclose(lhs + own, r[ib], rtol=0, atol=1e-9 * max(1.0, np.abs(r[ib]).max()))


@pytest.mark.parametrize('dim,p', [(2, 2), (3, 1)])
def test_implicit_euler_is_the_scalar_scheme_per_component(dim, p):
    """With no flow the m-unknown implicit-Euler system is the scalar oracle's
    apply_euler (pinned on the heat-equation decay) Kronecker-multiplied by I_m."""
    s, m, lu, lq, ll = sizes_m(dim, p)
    x = cell(dim, p, RNG)
    D = 0.5 + RNG.random(s['nN'])
    t = 0.5 + RNG.random(s['nF'] * s['nNf'])
    u_old = RNG.standard_normal((s['nN'], m))
    dt = 0.037
    K1, F1 = orc.local_system(dim, p, 1, x, D=D, ndD=1, tau=t)
    tau = np.einsum('j,ab->jba', t, np.eye(m)).reshape(-1)
    K, F = bo.burgers_local(dim, p, x, D, tau, np.zeros(lu), np.zeros(ll))
    K2, F2 = bo.implicit_euler(dim, m, s['nN'], K, F, dt, u_old.reshape(-1))
    for a in range(m):
        Ke, Fe = orc.apply_euler(dim, p, dt, u_old[:, a], K1, F1)
        sc = np.abs(Ke).max()
        assert np.allclose(K2[a::m, a::m], Ke, rtol=0, atol=1e-12 * sc)
        assert np.allclose(F2[a::m], Fe, rtol=0, atol=1e-12 * max(np.abs(Fe).max(), 1.0))
    assert np.allclose(K2, np.kron(orc.apply_euler(dim, p, dt, u_old[:, 0], K1, F1)[0], np.eye(m)), rtol=0,
                       atol=1e-12 * np.abs(K2).max())
