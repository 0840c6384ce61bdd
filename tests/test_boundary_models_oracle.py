"""IntegratedDirichletModel (src/model/IntegratedDirichletModel.cpp:35-56,
applied by HDGSolver::applyBoundaryConditions, HDGSolver.cpp:489-502) in the
CPU oracle, pinned by what the mathematics fixes.  No GPU.

* The face mass is the Mass operator on the face element with the face's own
  nodes: on straight P1 faces it is the textbook |f|/6 [[2,1],[1,2]] and
  |f|/12 (1 + I) mass; on any face its entries sum to the face measure, which
  for a curved (p = 2) edge is the arc length computed by an independent
  quadrature.
* The boundary rows become M lambda = M g, so the discrete solution is the
  one of DirichletModel (lambda = g on the boundary) to round-off, and only
  the boundary rows of the trace system change.
"""
import numpy as np
import pytest
from scipy import integrate

import hyperfox_amd as hf
import oracle as orc


def test_p1_face_masses_are_the_textbook_ones():
    x = np.array([[0.3, 0.1], [1.1, 0.7]])
    L = np.linalg.norm(x[1] - x[0])
    assert np.allclose(orc.face_mass(2, 1, x), L / 6 * np.array([[2, 1], [1, 2]]), rtol=1e-14)
    t = np.array([[0.0, 0.0, 0.1], [1.0, 0.2, 0.0], [0.3, 0.9, 0.5]])
    A = 0.5 * np.linalg.norm(np.cross(t[1] - t[0], t[2] - t[0]))
    assert np.allclose(orc.face_mass(3, 1, t), A / 12 * (np.ones((3, 3)) + np.eye(3)), rtol=1e-14)


@pytest.mark.parametrize('dim,p', [(2, 2), (2, 4), (3, 2), (3, 3)])
def test_face_mass_is_symmetric_positive_and_sums_to_the_measure(dim, p):
    m = hf.Mesh.structured(dim, p, 2, 0.1, 3)
    a = m.arrays()
    F = a['faces'][0]
    xf = a['coords'][F]
    M = orc.face_mass(dim, p, xf)
    assert np.allclose(M, M.T, rtol=0, atol=1e-15)
    assert np.all(np.linalg.eigvalsh(M) > 0)
    v = xf[:dim] if dim == 3 else xf[:2]
    meas = np.linalg.norm(v[1] - v[0]) if dim == 2 else 0.5 * np.linalg.norm(np.cross(v[1] - v[0], v[2] - v[0]))
    assert np.isclose(M.sum(), meas, rtol=1e-13)


def test_curved_edge_mass_sums_to_the_arc_length():
    # p = 2 edge, face-node order (end, end, middle) as the face element
    # carries it; the middle node is lifted off the chord
    xf = np.array([[0.0, 0.0], [1.0, 0.0], [0.5, 0.2]])
    # parametrise the quadratic through the three nodes at s = 0, 1, 1/2
    def x(s):
        l0, l1, l2 = 2 * (s - 0.5) * (s - 1), 2 * s * (s - 0.5), 4 * s * (1 - s)
        return l0 * xf[0] + l1 * xf[1] + l2 * xf[2]

    def speed(s, h=1e-6):
        return np.linalg.norm((x(s + h) - x(s - h)) / (2 * h))
    arc = integrate.quad(speed, 0, 1, epsabs=1e-13)[0]
    M = orc.face_mass(2, 2, xf)
    # the 2p = 4 face rule integrates sqrt(det J J^T) only approximately
    assert abs(M.sum() - arc) < 1e-3 * arc
    assert M.sum() > 1.0 + 1e-3        # longer than the chord: the measure is not the chord's


def system(dim, p, seed=2):
    m = hf.Mesh.structured(dim, p, 3 if dim == 2 else 2, 0.12 if dim == 2 else 0.05, seed)
    a = m.arrays()
    nNf = m.sizes['nNf']
    rng = np.random.default_rng(seed)
    x = a['coords']
    b = np.stack([np.sin(2 * x[:, 1] + 0.3), np.cos(x[:, 0])] + ([np.sin(x[:, 0] * x[:, 1])] if dim == 3 else []), 1)
    tau = rng.uniform(0.5, 2.0, (m.n_faces, nNf))
    g = rng.normal(size=(m.n_faces, nNf))
    xip = np.stack([orc.ip_coords(dim, p, x[c]) for c in a['cells']])
    f = np.cos(xip[..., 0] + xip[..., 1])
    mk = lambda: orc.System(dim, p, 2, a, tau, tau_sides=1, D=0.2 + 0.05 * np.sin(3 * x[:, 0]), ndD=1, b=b,
                            f_ip=f, dirichlet=g)
    return m, a, g, mk


@pytest.mark.parametrize('dim,p', [(2, 1), (2, 3), (3, 2)])
def test_integrated_dirichlet_gives_the_dirichlet_solution(dim, p):
    m, a, g, mk = system(dim, p)
    s0, s1 = mk(), mk()
    s1.set_boundary_model(1)
    rp, cols, v0, r0 = s0.csr()
    rp1, cols1, v1, r1 = s1.csr()
    assert np.array_equal(rp, rp1) and np.array_equal(cols, cols1)
    nNf = m.sizes['nNf']
    bnd = np.zeros(m.n_faces, bool)
    bnd[np.asarray(a['face2cell'])[:, 1] < 0] = True
    rows_b = np.repeat(bnd, nNf)
    row_of = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    inter = ~rows_b[row_of]
    assert np.array_equal(v0[inter], v1[inter]) and np.array_equal(r0[~rows_b], r1[~rows_b])
    # boundary rows: the face mass of the face's own nodes, rhs M g
    for F in np.nonzero(bnd)[0][:5]:
        M = orc.face_mass(dim, p, a['coords'][a['faces'][F]])
        for aa in range(nNf):
            r = F * nNf + aa
            c = cols[rp[r]:rp[r + 1]]
            blk = v1[rp[r]:rp[r + 1]][(c >= F * nNf) & (c < (F + 1) * nNf)]
            assert np.allclose(blk, M[aa], rtol=1e-14, atol=1e-17)
        assert np.allclose(r1[F * nNf:(F + 1) * nNf], M @ g[F], rtol=1e-13, atol=1e-16)
    l0 = s0.solve(method=1)[0]
    l1 = s1.solve(method=1)[0]
    assert np.allclose(l1, l0, rtol=0, atol=1e-10 * np.abs(l0).max())
    assert np.allclose(l1.reshape(m.n_faces, nNf)[bnd], g[bnd], rtol=0, atol=1e-10 * np.abs(g).max())


def test_boundary_model_per_face_and_checks():
    m, a, g, mk = system(2, 2)
    s = mk()
    bnd = np.nonzero(np.asarray(a['face2cell'])[:, 1] < 0)[0]
    s.set_boundary_model(1, bnd[::2])
    l = s.solve(method=1)[0]
    l0 = mk().solve(method=1)[0]
    assert np.allclose(l, l0, rtol=0, atol=1e-10 * np.abs(l0).max())
    interior = np.nonzero(np.asarray(a['face2cell'])[:, 1] >= 0)[0]
    with pytest.raises(orc.OracleError, match='boundary face'):
        s.set_boundary_model(1, interior[:1])
    with pytest.raises(orc.OracleError, match='unknown kind'):
        s.set_boundary_model(3)


# ---------------------------------------------------------------- HDG type
# The HDG-type branch of applyBoundaryConditions (HDGSolver.cpp:503-518): a
# boundary model's local matrix [B_u | B_q | B_lambda] and rhs r act on the
# face's rows through the cell's U, Q, U0, Q0.


def hdg_B(dim, p, nfaces, seed, scale=0.3):
    rng = np.random.default_rng(seed)
    nNf = orc.sizes(dim, p)['nNf']
    B = scale * rng.normal(size=(nfaces, nNf, nNf * (2 + dim)))
    B[:, :, nNf * (1 + dim):] += 2.0 * np.eye(nNf)     # keep the face rows well posed
    return B, rng.normal(size=(nfaces, nNf))


def test_hdg_type_model_can_state_dirichlet():
    m, a, g, mk = system(2, 2)
    bf = np.nonzero(np.asarray(a['face2cell'])[:, 1] < 0)[0]
    nNf = m.sizes['nNf']
    B = np.zeros((bf.size, nNf, nNf * 4))
    B[:, :, 3 * nNf:] = np.eye(nNf)
    s1 = mk()
    s1.set_boundary_hdg(bf, B, g[bf])
    v0, r0 = mk().csr()[2:]
    v1, r1 = s1.csr()[2:]
    assert np.array_equal(v1, v0) and np.array_equal(r1, r0)


@pytest.mark.parametrize('dim,p', [(2, 1), (2, 3), (3, 2)])
def test_hdg_type_rows_impose_their_constraint(dim, p):
    """After the solve, B_u u_F + B_q q_F + B_lambda lambda_F = r on every
    such face, with u, q the reconstructed fields."""
    m, a, g, mk = system(dim, p)
    bf = np.nonzero(np.asarray(a['face2cell'])[:, 1] < 0)[0][::2]
    B, r = hdg_B(dim, p, bf.size, 7)
    s = mk()
    s.set_boundary_hdg(bf, B, r)
    lam = s.solve(method=1)[0]
    u, q = s.reconstruct(lam)
    nNf = m.sizes['nNf']
    fn = orc.refel(dim, p)['face_nodes']
    for i, F in enumerate(bf):
        e = a['face2cell'][F, 0]
        lf = list(a['cell2face'][e]).index(F)
        nodes = fn[lf]
        uF = u[e, nodes]
        qF = q[e].reshape(-1, dim)[nodes].reshape(-1)
        lF = lam[F * nNf:(F + 1) * nNf]
        res = B[i, :, :nNf] @ uF + B[i, :, nNf:nNf * (1 + dim)] @ qF + B[i, :, nNf * (1 + dim):] @ lF - r[i]
        assert np.abs(res).max() < 1e-9 * max(1.0, np.abs(r[i]).max()), (F, np.abs(res).max())


def test_hdg_type_add_is_set_plus_the_natural_rows():
    m, a, g, mk = system(2, 3)
    bf = np.nonzero(np.asarray(a['face2cell'])[:, 1] < 0)[0][:5]
    B, r = hdg_B(2, 3, bf.size, 9)
    out = {}
    for name, BB, rr, st in (('set', B, r, True), ('add', B, r, False), ('nat', 0 * B, 0 * r, False)):
        s = mk()
        s.set_boundary_hdg(bf, BB, rr, matrix_set=st, rhs_set=st)
        out[name] = s.csr()
    rp, cols = out['set'][:2]
    nNf = m.sizes['nNf']
    rows = np.concatenate([np.arange(F * nNf, (F + 1) * nNf) for F in bf])
    for rw in rows:
        sl = slice(rp[rw], rp[rw + 1])
        assert np.allclose(out['add'][2][sl], out['set'][2][sl] + out['nat'][2][sl], rtol=1e-13, atol=1e-14)
        assert np.isclose(out['add'][3][rw], out['set'][3][rw] + out['nat'][3][rw], rtol=1e-13, atol=1e-14)
    # the natural rows are the cell's own flux rows: no identity block left
    F = bf[0]
    sl = slice(rp[F * nNf], rp[F * nNf + 1])
    assert not np.allclose(out['nat'][2][sl][cols[sl] == F * nNf], 1.0)
