"""The oracle's HDGBohmModel (oracle/bohm_oracle.py, restating
src/model/HDGBohmModel.cpp:97-268 with the transfer functions of
tests/regression/HDG/TestHDGBohm.cpp:37-130), pinned by:

- the Neumann transfer (T = 0): both rows are tau_a. (u - lambda) - (D n).q
  face integrals, i.e. face mass blocks scaled by the constant fields (the
  face mass is pinned on the textbook |f|/6 [[2,1],[1,2]] elsewhere);
- the Dirichlet transfer (T = 1): the Gamma rows become
  u_Gamma b.n - u_n c_s |b.n| mass blocks with no q or lambda part;
- the smooth (tanh) transfer near its switch, curved face, variable tensor
  fields: B is the Jacobian, by central differences, of the face residual
  written from the fields at the cubature points (not from the matrix), and
  r = B U0 - N(U0), which makes B U = r the Newton step of N(U) = 0;
- the step (heaviside) transfer: the same with T frozen (zero derivative)."""
import numpy as np
import pytest

import bohm_oracle as bo
import oracle as orc


def straight_face(p, a=(0.2, 0.1), b=(0.7, 0.45)):
    t = np.linspace(0.0, 1.0, p + 1)
    s = (bo.face_tables(p)[0].shape[1])
    assert s == p + 1
    ref = orc.refel(2, p)
    fn = ref['face_nodes'][0]
    lam = (ref['nodes'][fn][:, 0] + 1.0) / 2.0           # face-node positions along the face
    del t
    return np.asarray(a) + lam[:, None] * (np.asarray(b) - np.asarray(a))[None, :]


def fields(p, rng, const=False):
    nN = p + 1
    xf = straight_face(p)
    if not const:
        xf = xf + 0.02 * np.stack([np.sin(3 * xf[:, 1]), np.cos(2 * xf[:, 0])], 1) * np.sin(np.pi * np.linspace(0, 1, nN))[:, None]
    t = xf[-1] - xf[0]
    n0 = np.array([t[1], -t[0]]) / np.linalg.norm(t)
    if const:
        normals = np.tile(n0, (nN, 1))
        tau = np.tile([1.3, 0.2, -0.1, 0.9], (nN, 1))
        D, G = np.full(nN, 0.7), np.full(nN, 0.4)
        b = np.tile([0.6, -0.3], (nN, 1))
        cs = np.full(nN, 1.1)
    else:
        normals = n0 + 0.05 * rng.standard_normal((nN, 2))
        tau = np.tile([1.3, 0.2, -0.1, 0.9], (nN, 1)) + 0.1 * rng.standard_normal((nN, 4))
        D = np.tile([0.7, 0.05, -0.02, 0.5], (nN, 1)) + 0.05 * rng.standard_normal((nN, 4))
        G = np.tile([0.4, -0.03, 0.06, 0.6], (nN, 1)) + 0.05 * rng.standard_normal((nN, 4))
        b = np.tile([0.6, -0.3], (nN, 1)) + 0.1 * rng.standard_normal((nN, 2))
        cs = 1.1 + 0.1 * rng.standard_normal(nN)
    state = dict(sol=1.0 + 0.2 * rng.standard_normal((nN, 2)), trace=1.0 + 0.2 * rng.standard_normal((nN, 2)),
                 flux=0.3 * rng.standard_normal((nN, 4)))
    return dict(xf=xf, normals=normals, b=b, tau=tau, D=D, G=G, cs=cs, **state)


def call(f, p, fn, kind, **over):
    g = dict(f)
    g.update(over)
    return fn(p, g['xf'], g['normals'], g['b'], g['trace'], g['flux'], g['tau'], g['D'], g['G'], g['sol'], g['cs'], kind)


@pytest.mark.parametrize('p', [1, 2, 3, 4])
@pytest.mark.parametrize('kind', [0, 1])
def test_neumann_and_dirichlet_transfers_are_face_mass_blocks(p, kind):
    rng = np.random.default_rng(p)
    f = fields(p, rng, const=True)
    B, r, orig = call(f, p, bo.bohm_local, kind)
    nN = p + 1
    M = orc.face_mass(2, p, f['xf'])
    n = f['normals'][0]
    bn = f['b'][0] @ n
    tau = np.array([[1.3, -0.1], [0.2, 0.9]])            # column-major [1.3, 0.2, -0.1, 0.9]
    ex = np.zeros_like(B)
    for a in range(2):
        rows = np.arange(nN) * 2 + a
        coef = f['D'][0] if a == 0 else f['G'][0]
        if a == 0 or kind == 0:
            for c in range(2):
                ex[np.ix_(rows, np.arange(nN) * 2 + c)] += tau[a, c] * M
                ex[np.ix_(rows, 6 * nN + np.arange(nN) * 2 + c)] -= tau[a, c] * M
            for k in range(2):
                ex[np.ix_(rows, 2 * nN + (np.arange(nN) * 2 + k) * 2 + a)] -= coef * n[k] * M
        else:
            ex[np.ix_(rows, np.arange(nN) * 2)] -= f['cs'][0] * abs(bn) * M
            ex[np.ix_(rows, np.arange(nN) * 2 + 1)] += bn * M
    assert np.allclose(B, ex, rtol=0, atol=1e-14 * np.abs(ex).max())
    assert np.allclose(r, 0.0, atol=1e-15) and np.array_equal(B, orig)


def near_switch_state(p, rng, f, gap=3e-4):
    """Constant (n, Gamma) on the face with x = Gamma b.n and y = n c_s b.n
    gap apart and both positive: the tanh transfer's derivatives are O(1e3)."""
    phi = bo.face_tables(p)[0]
    nq = phi.shape[0]
    # at the cubature points b.n and c_s vary slightly; choose Gamma so that
    # x - y is about `gap` at the first point
    sh = phi[0]
    bn = (sh @ f['b']) @ (sh @ f['normals'])
    cs = sh @ f['cs']
    nval = 0.8
    gam = (nval * cs * bn + gap) / bn
    sol = np.tile([nval, gam], (p + 1, 1)) + 1e-5 * rng.standard_normal((p + 1, 2))
    del nq
    return sol


@pytest.mark.parametrize('p', [1, 2, 3])
@pytest.mark.parametrize('kind', [2, 3])
def test_rows_are_the_jacobian_of_the_face_residual(p, kind):
    rng = np.random.default_rng(10 + p + kind)
    f = fields(p, rng)
    f['b'] = np.abs(f['b'])                              # b.n > 0 on this face
    if f['b'][0] @ f['normals'][0] < 0:
        f['normals'] = -f['normals']
    if kind == 3:
        f['sol'] = near_switch_state(p, rng, f)
    B, r, orig = call(f, p, bo.bohm_local, kind)
    nN = p + 1
    U0 = np.concatenate([f['sol'].ravel(), f['flux'].ravel(), f['trace'].ravel()])

    def N(U):
        return call(f, p, bo.bohm_residual, kind, sol=U[:2 * nN].reshape(nN, 2), flux=U[2 * nN:6 * nN].reshape(nN, 4),
                    trace=U[6 * nN:].reshape(nN, 2))
    h = 1e-7
    Jfd = np.zeros_like(B)
    for j in range(U0.size):
        e = np.zeros_like(U0)
        e[j] = h
        Jfd[:, j] = (N(U0 + e) - N(U0 - e)) / (2 * h)
    scale = np.abs(B).max()
    assert np.abs(B - Jfd).max() < 1e-5 * scale, np.abs(B - Jfd).max() / scale
    if kind == 3:
        # the transfer's derivative terms are really exercised: B differs from
        # the frozen-transfer rows by much more than the bar
        assert np.abs(B - orig).max() > 1e-2 * scale
    assert np.allclose(B @ U0 - r, N(U0), rtol=0, atol=1e-12 * np.abs(B).max() * np.abs(U0).max())
