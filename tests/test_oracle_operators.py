"""Local HDG operators of the oracle against closed forms: the known-answer
tests of tests/unittests/operator/Test{Mass,Convection,Source,HDGBase,
HDGDiffusion,HDGConvection}.cpp, restated as bilinear-form identities
u^T K v = (exact integral) for polynomials u, v of degree <= p, on reference,
rotated and jittered simplices, plus a brute-force check of the static
condensation against the monolithic local solve."""
import numpy as np
import pytest

import oracle as orc
from hdg_helpers import (Blocks, dmono, element_nodes, face_vertex_coords, integrate, integrate_face, mono,
                         monomials, outward_normal)

ELEMS = {
    2: {'ref': [[-1, -1], [1, -1], [-1, 1]],
        'rot': [[0.3, 0.1], [1.1, 0.6], [-0.2, 0.9]],
        'jit': [[0.02, -0.01], [0.27, 0.03], [0.04, 0.22]]},
    3: {'ref': [[-1, -1, -1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]],
        'rot': [[0.1, 0.2, 0.0], [1.0, 0.4, 0.1], [0.2, 1.1, 0.3], [0.3, 0.2, 0.9]],
        'jit': [[0.01, 0.0, 0.02], [0.26, 0.02, 0.0], [0.24, 0.23, 0.01], [0.26, 0.24, 0.27]]},
}
PMAX = {2: 10, 3: 3}
CASES = [(d, p, e) for d in (2, 3) for p in range(1, PMAX[d] + 1) for e in ELEMS[d]]


def nonsymmetric_D(dim):
    """A constant diffusion tensor with D != D^T (positive definite symmetric
    part), so that reading the column-major node values transposed
    (HDGConvectionDiffusionReactionSource.cpp:122-124) changes every D term."""
    Dm = np.diag(np.arange(1, dim + 1, dtype=float))
    Dm += 0.3 * np.triu(np.ones((dim, dim)), 1) - 0.2 * np.tril(np.ones((dim, dim)), -1)
    return Dm


def _tau(dim, p, val=1.0):
    s = orc.sizes(dim, p)
    return np.full(s['nF'] * s['nNf'], val)


@pytest.mark.parametrize('dim,p,el', CASES)
def test_mass_and_gradient_blocks(dim, p, el):
    V = ELEMS[dim][el]
    X = element_nodes(dim, p, V)
    K, _ = orc.local_system(dim, p, 1, X, tau=_tau(dim, p))
    B = Blocks(dim, p)
    o = 2 * p
    M = K[np.ix_(B.q(0), B.q(0))]
    vol = integrate(V, lambda x: np.ones(len(x)), 1)
    assert abs(np.ones(B.nN) @ M @ np.ones(B.nN) - vol) < 1e-13 * max(vol, 1)
    for k in range(dim):
        assert np.allclose(K[np.ix_(B.q(k), B.q(k))], M, atol=1e-15)
    for eu in monomials(dim, p)[::2]:
        for ev in monomials(dim, p)[1::3]:
            u = mono(eu)(X)
            v = mono(ev)(X)
            ex = integrate(V, lambda x: mono(eu)(x) * mono(ev)(x), o)
            assert abs(u @ M @ v - ex) < 1e-12 * max(1, abs(ex))
            for k in range(dim):
                Squ = K[np.ix_(B.q(k), B.u())]
                ex = integrate(V, lambda x: dmono(eu, k)(x) * mono(ev)(x), o)
                assert abs(u @ Squ @ v - ex) < 1e-11 * max(1, abs(ex))


@pytest.mark.parametrize('dim,p,el', CASES)
def test_face_blocks(dim, p, el):
    """tau, normal and D.n face terms (HDGBase.cpp:106-137, HDGDiffusion.cpp:109-128)."""
    V = ELEMS[dim][el]
    X = element_nodes(dim, p, V)
    Dm = nonsymmetric_D(dim)
    Dnodes = np.tile(Dm.T.reshape(-1), (X.shape[0], 1))  # column-major per node
    # constant tau on each face keeps every integrand within the face rule's
    # exactness (degree 2p); variable tau is covered by the GPU parity tests
    nNf = orc.sizes(dim, p)['nNf']
    tau = np.repeat(np.linspace(0.5, 2.0, dim + 1), nNf)
    K, _ = orc.local_system(dim, p, 1, X, D=Dnodes, ndD=dim * dim, tau=tau)
    B = Blocks(dim, p)
    fn_all = orc.refel(dim, p)['face_nodes']
    o = 3 * p
    eu, ew = monomials(dim, p)[-1], monomials(dim, p)[1]
    u = mono(eu)(X)
    for f in range(B.nF):
        fv, fx = face_vertex_coords(dim, p, V, f)
        n = outward_normal(V, fv)
        w = mono(ew)(fx)                           # trace values at face nodes
        tc = tau[f * B.nNf]
        tf = lambda x: np.full(len(x), tc)
        Sll = K[np.ix_(B.lam(f), B.lam(f))]
        ex = integrate_face(fv, lambda x: tf(x) * mono(ew)(x) * mono(ew)(x), o + p)
        assert abs(w @ Sll @ w + ex) < 1e-11 * max(1, abs(ex))
        Slu = K[np.ix_(B.lam(f), B.u())]
        ex = integrate_face(fv, lambda x: tf(x) * mono(ew)(x) * mono(eu)(x), o + p)
        assert abs(w @ Slu @ u - ex) < 1e-11 * max(1, abs(ex))
        for k in range(dim):
            Sql = K[np.ix_(B.q(k), B.lam(f))]
            ex = integrate_face(fv, lambda x: n[k] * mono(eu)(x) * mono(ew)(x), o)
            assert abs(u @ Sql @ w + ex) < 1e-11 * max(1, abs(ex))
            Slq = K[np.ix_(B.lam(f), B.q(k))]
            ex = integrate_face(fv, lambda x: (Dm @ n)[k] * mono(ew)(x) * mono(eu)(x), o)
            assert abs(w @ Slq @ u + ex) < 1e-11 * max(1, abs(ex))
        assert set(fn_all[f]) <= set(range(B.nN))


@pytest.mark.parametrize('dim,p,el', CASES)
def test_diffusion_volume_and_convection(dim, p, el):
    """Suq = int (D grad u)_k v - faces (HDGDiffusion.cpp:129-144) and
    Suu = -int (b.grad u) v (HDGConvection.cpp:79-81) with constant D, b."""
    V = ELEMS[dim][el]
    X = element_nodes(dim, p, V)
    nN = X.shape[0]
    bvec = np.array([0.7, -0.4, 0.25][:dim])
    Dm = np.eye(dim) * 0.5
    B = Blocks(dim, p)
    o = 2 * p
    zero_tau = _tau(dim, p, 0.0)
    K, _ = orc.local_system(dim, p, 2, X, D=np.full(nN, 0.5), ndD=1, b=np.tile(bvec, (nN, 1)), tau=zero_tau)
    eu, ev = monomials(dim, p)[-1], monomials(dim, p)[2]
    u = mono(eu)(X)
    v = mono(ev)(X)
    Suu = K[np.ix_(B.u(), B.u())]
    ex = -integrate(V, lambda x: sum(bvec[k] * dmono(eu, k)(x) for k in range(dim)) * mono(ev)(x), o)
    assert abs(u @ Suu @ v - ex) < 1e-11 * max(1, abs(ex))
    for k in range(dim):
        Suq = K[np.ix_(B.u(), B.q(k))]
        ex = integrate(V, lambda x: 0.5 * dmono(eu, k)(x) * mono(ev)(x), o)
        for f in range(B.nF):
            fv, _ = face_vertex_coords(dim, p, V, f)
            n = outward_normal(V, fv)
            ex -= integrate_face(fv, lambda x: 0.5 * n[k] * mono(eu)(x) * mono(ev)(x), o)
        assert abs(u @ Suq @ v - ex) < 1e-11 * max(1, abs(ex))
    for f in range(B.nF):
        fv, fx = face_vertex_coords(dim, p, V, f)
        n = outward_normal(V, fv)
        w = mono(ev)(fx)
        Sul = K[np.ix_(B.u(), B.lam(f))]
        ex = integrate_face(fv, lambda x: (bvec @ n) * mono(eu)(x) * mono(ev)(x), o)
        assert abs(u @ Sul @ w - ex) < 1e-11 * max(1, abs(ex))


@pytest.mark.parametrize('dim,p,el', [c for c in CASES if c[2] != 'ref'])
def test_source_vector(dim, p, el):
    """F_u = int f phi (Source.cpp:24-39): F^T v = int f v."""
    V = ELEMS[dim][el]
    X = element_nodes(dim, p, V)
    xip = orc.ip_coords(dim, p, X)
    ef, ev = monomials(dim, p)[1], monomials(dim, p)[-1]
    _, F = orc.local_system(dim, p, 1, X, tau=_tau(dim, p), f_ip=mono(ef)(xip))
    B = Blocks(dim, p)
    ex = integrate(V, lambda x: mono(ef)(x) * mono(ev)(x), 2 * p)
    assert abs(F[B.u()] @ mono(ev)(X) - ex) < 1e-12 * max(1, abs(ex))
    assert np.all(F[B.nN:] == 0.0)


@pytest.mark.parametrize('dim,p,el', CASES)
def test_condensation_against_monolithic_solve(dim, p, el):
    """U lambda + U0 and Q lambda + Q0 equal the direct solve of the local
    [u, q] system for a given lambda; S lambda - S0 equals the lambda-rows
    residual (HDGSolver.cpp:331-348 by brute force)."""
    V = ELEMS[dim][el]
    X = element_nodes(dim, p, V)
    nN = X.shape[0]
    rng = np.random.default_rng(dim * 100 + p)
    s = orc.sizes(dim, p)
    tau = rng.uniform(0.5, 2.0, s['nF'] * s['nNf'])
    b = rng.normal(size=(nN, dim)) * 0.3
    Dn = rng.uniform(0.5, 1.5, nN)
    f = rng.normal(size=s['nIP'])
    K, F = orc.local_system(dim, p, 2, X, D=Dn, ndD=1, b=b, tau=tau, f_ip=f)
    c = orc.condense(dim, p, K, F)
    nUQ = nN * (dim + 1)
    lam = rng.normal(size=s['lenL'])
    A = K[:nUQ, :nUQ]
    rhs = F[:nUQ] - K[:nUQ, nUQ:] @ lam
    uq = np.linalg.solve(A, rhs)
    assert np.allclose(c['U'] @ lam + c['U0'], uq[:nN], rtol=1e-11, atol=1e-11)
    assert np.allclose(c['Q'] @ lam + c['Q0'], uq[nN:], rtol=1e-11, atol=1e-11)
    lrow = K[nUQ:, :nUQ] @ uq + K[nUQ:, nUQ:] @ lam - F[nUQ:]
    assert np.allclose(c['S'] @ lam - c['S0'], lrow, rtol=1e-10, atol=1e-11)


@pytest.mark.parametrize('dim,p,el', CASES)
def test_nonsymmetric_diffusion_volume(dim, p, el):
    """Suq_k = int (D grad u)_k v - sum_f int (D n)_k u v with a constant,
    non-symmetric D given column-major per node (HDGDiffusion.cpp:129-144 with
    the field layout of HDGConvectionDiffusionReactionSource.cpp:122-124):
    the transposed tensor gives different values, which this bar separates."""
    V = ELEMS[dim][el]
    X = element_nodes(dim, p, V)
    Dm = nonsymmetric_D(dim)
    Dnodes = np.tile(Dm.T.reshape(-1), (X.shape[0], 1))
    K, _ = orc.local_system(dim, p, 1, X, D=Dnodes, ndD=dim * dim, tau=_tau(dim, p, 0.0))
    B = Blocks(dim, p)
    o = 2 * p
    mons = monomials(dim, p)
    gap = scale = 0.0
    for eu in (mons[-1], mons[1]):
        for ev in mons[1:]:
            u, v = mono(eu)(X), mono(ev)(X)
            for k in range(dim):
                Suq = K[np.ix_(B.u(), B.q(k))]

                def exact(D):
                    ex = integrate(V, lambda x: sum(D[k, c] * dmono(eu, c)(x) for c in range(dim)) * mono(ev)(x), o)
                    for f in range(B.nF):
                        fv, _ = face_vertex_coords(dim, p, V, f)
                        n = outward_normal(V, fv)
                        ex -= integrate_face(fv, lambda x: (D @ n)[k] * mono(eu)(x) * mono(ev)(x), o)
                    return ex
                ex, ex_t = exact(Dm), exact(Dm.T)
                assert abs(u @ Suq @ v - ex) < 1e-11 * max(1, abs(ex))
                gap = max(gap, abs(ex - ex_t))
                scale = max(scale, abs(ex))
    assert gap > 1e-3 * scale   # the transposed tensor fails the bar above somewhere


@pytest.mark.parametrize('dim,p,el', CASES)
def test_reaction_block(dim, p, el):
    """Reaction (Reaction.cpp:5-36) is a Mass with the measure r dV, added to
    the u-u block of HDGConvectionDiffusionReactionSource only
    (HDGConvectionDiffusionReactionSource.cpp:90-95).  With r(x) = r0 + g.x
    evaluated at the element IPs: the local system with r minus the one
    without is zero outside Suu and in F, u^T dSuu v = int r u v for
    monomials with deg(u v) <= 2p - 1 (the integrand then stays within the
    rule's degree 2p), and a constant r gives r0 times the mass block.  The
    other models have no Reaction operator and refuse r."""
    V = ELEMS[dim][el]
    X = element_nodes(dim, p, V)
    nN = X.shape[0]
    rng = np.random.default_rng(7 * dim + p)
    s = orc.sizes(dim, p)
    tau = rng.uniform(0.5, 2.0, s['nF'] * s['nNf'])
    b = rng.normal(size=(nN, dim)) * 0.3
    Dn = rng.uniform(0.5, 1.5, nN)
    f = rng.normal(size=s['nIP'])
    xip = orc.ip_coords(dim, p, X)
    g = np.array([0.6, -0.9, 0.4][:dim])
    r0 = 1.3
    args = dict(D=Dn, ndD=1, b=b, tau=tau, f_ip=f)
    K0, F0 = orc.local_system(dim, p, 2, X, **args)
    K1, F1 = orc.local_system(dim, p, 2, X, r_ip=r0 + xip @ g, **args)
    Kc, _ = orc.local_system(dim, p, 2, X, r_ip=np.full(s['nIP'], r0), **args)
    B = Blocks(dim, p)
    dK = K1 - K0
    outside = dK.copy()
    outside[np.ix_(B.u(), B.u())] = 0.0
    assert np.all(outside == 0.0) and np.array_equal(F1, F0)
    dS = dK[np.ix_(B.u(), B.u())]
    M = K0[np.ix_(B.q(0), B.q(0))]
    uu = np.ix_(B.u(), B.u())
    # the difference cancels the (larger) convection and tau entries of Suu:
    # compare at their round-off
    assert np.allclose((Kc - K0)[uu], r0 * M, rtol=0, atol=1e-14 * np.abs(Kc[uu]).max())
    mons = monomials(dim, p)
    for eu in mons:
        for ev in mons:
            if eu.sum() + ev.sum() > 2 * p - 1:
                continue
            ex = integrate(V, lambda x: (r0 + x @ g) * mono(eu)(x) * mono(ev)(x), 2 * p)
            assert abs(mono(eu)(X) @ dS @ mono(ev)(X) - ex) < 1e-12 * max(1, abs(ex))
    for kind in (0, 1):
        with pytest.raises(orc.OracleError):
            orc.local_system(dim, p, kind, X, tau=tau, r_ip=r0 + xip @ g)
