"""Pinning the FP64 numpy oracle of the learned preconditioner (CPU).

There is no reference counterpart, so the oracle is pinned by mathematics:
* its hand-written backward pass against central finite differences of its
  own loss, on a real HDG trace system (catches a dropped term, a wrong
  sign, a transposed weight or a mis-routed edge gradient);
* zero decoders give exactly block Jacobi (loss = ||z - D^-1 A z||^2 computed
  densely);
* the normalised diagonal blocks have mean |diagonal| 1 and Dirichlet rows
  stay identity rows of M;
* the block matrix-vector product equals the dense product.
"""
import numpy as np
import pytest

import hyperfox_amd as hf
import hyperfox_amd.problems as pr
import lpc_oracle as lo
import oracle as orc


def small_system(p=2, n=3, seed=5):
    w = pr.diffusion_grf_2d(n, p, seed)
    a = w.mesh.arrays()
    nip = orc.sizes(2, p)['nIP']
    f = np.ones((w.mesh.n_cells, nip))
    s = orc.System(2, p, 1, a, w.tau, D=w.D, ndD=1, f_ip=f, dirichlet=w.dirichlet)
    rowptr, cols, vals, rhs = s.csr()
    lam, _, conv, _ = s.solve(method=1)
    assert conv
    b = w.mesh.sizes['nNf']
    dirf = np.zeros(w.mesh.n_faces, dtype=bool)
    dirf[a['boundary']] = True
    g = lo.graph_from_scalar_csr(rowptr, cols, b, dirf)
    A = lo.blocks_from_scalar_csr(g, rowptr, cols, vals, b)
    dense = np.zeros((len(rowptr) - 1,) * 2)
    for r in range(len(rowptr) - 1):
        dense[r, cols[rowptr[r]:rowptr[r + 1]]] = vals[rowptr[r]:rowptr[r + 1]]
    return g, A, lam, b, dense, rhs


def random_params(b, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    L = lo.layout(b)
    p = np.zeros(L['total'])
    for k, (o, n) in ((k, v) for k, v in L.items() if k != 'total'):
        fan = {'Wn': b * b + 1, 'We': b * b, 'Wdn': lo.H, 'Wde': lo.H}.get(k, 3 * lo.H if k.startswith('Wm') else 2 * lo.H)
        if k.startswith('W'):
            p[o:o + n] = rng.uniform(-1, 1, n) * np.sqrt(6.0 / fan) * scale
        else:
            p[o:o + n] = rng.uniform(-0.1, 0.1, n)
    for k in ('Wdn', 'Wde'):
        o, n = L[k]
        p[o:o + n] *= 0.05
    return p


def test_block_mv_and_normalisation():
    g, A, lam, b, dense, rhs = small_system()
    x = np.random.default_rng(0).standard_normal(g.nb * b)
    assert np.allclose(lo.bsr_mv(g, A, x), dense @ x, rtol=0, atol=1e-12 * np.abs(dense).max())
    assert np.allclose(dense @ lam, rhs, atol=1e-9 * np.abs(rhs).max())
    sinv, ahat, dinv = lo.normalise(g, A)
    md = np.abs(np.diagonal(ahat[g.diag], axis1=1, axis2=2)).mean(axis=1)
    assert np.allclose(md, 1.0)
    assert np.allclose(ahat[g.diag][g.dir], np.eye(b))  # Dirichlet rows are identity rows
    for k in g.eblk[g.dir[g.esrc]]:
        assert np.all(ahat[k] == 0.0)


def test_zero_decoders_are_block_jacobi():
    g, A, lam, b, dense, rhs = small_system()
    p = random_params(b, 1)
    L = lo.layout(b)
    for k in ('Wdn', 'bdn', 'Wde', 'bde'):
        o, n = L[k]
        p[o:o + n] = 0.0
    loss, grad, acts = lo.loss_grad(p, b, g, A, lam)
    sinv, ahat, dinv = lo.normalise(g, A)
    S = np.repeat(sinv, b)
    Ah = dense * S[:, None] * S[None, :]
    Dinv = np.zeros_like(Ah)
    for F in range(g.nb):
        Dinv[F * b:(F + 1) * b, F * b:(F + 1) * b] = np.linalg.inv(Ah[F * b:(F + 1) * b, F * b:(F + 1) * b])
    z = lam / S
    z /= np.linalg.norm(z)
    r = z - Dinv @ (Ah @ z)
    assert loss == pytest.approx(float(r @ r), rel=1e-12)


@pytest.mark.parametrize('n_random', [0, 2])
def test_gradient_matches_finite_differences(n_random):
    g, A, lam, b, dense, rhs = small_system()
    p = random_params(b, 3)
    loss, grad, _ = lo.loss_grad(p, b, g, A, lam, n_random=n_random, seed=11)
    assert loss > 0
    L = lo.layout(b)
    rng = np.random.default_rng(4)
    eps = 1e-6
    worst = 0.0
    for k, (o, n) in ((k, v) for k, v in L.items() if k != 'total'):
        for idx in o + rng.choice(n, size=min(n, 4), replace=False):
            pp, pm = p.copy(), p.copy()
            pp[idx] += eps
            pm[idx] -= eps
            fd = (lo.loss_grad(pp, b, g, A, lam, n_random, 11)[0] - lo.loss_grad(pm, b, g, A, lam, n_random, 11)[0]) / (2 * eps)
            err = abs(fd - grad[idx]) / (abs(fd) + abs(grad[idx]) + 1e-8 * abs(loss))
            worst = max(worst, err)
            assert err < 1e-4, (k, idx, fd, grad[idx])
    print('worst relative FD mismatch %.2e' % worst)


def test_rademacher_stream_is_balanced():
    z = lo.rademacher(4000, lo.probe_seed(7, 0))
    assert set(np.unique(z)) == {-1.0, 1.0}
    assert abs(z.mean()) < 0.05
