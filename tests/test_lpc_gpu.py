"""Learned trace preconditioner on the GPU against its FP64 numpy oracle.

* forward pass (FP32 SIMT and BF16 MFMA) equals the oracle's network outputs;
* loss and the full parameter gradient of the GPU backward pass equal the
  oracle's (itself pinned by finite differences in test_lpc_oracle.py);
* training lowers the loss and is bit-reproducible;
* FGMRES with the learned preconditioner reaches the same true FP64 residual
  and the same solution as block Jacobi;
* the committed trained network takes fewer iterations than block Jacobi on
  held-out systems.
"""
import os

import numpy as np
import pytest

import hyperfox_amd as hf
import hyperfox_amd.problems as pr
from hyperfox_amd.lpc import LearnedPreconditioner, pc_setup_ms
import lpc_oracle as lo
from test_lpc_oracle import random_params

from conftest import ROOT

pytestmark = pytest.mark.gpu

WEIGHTS = os.path.join(ROOT, 'hyperfox_amd', 'data', 'lpc_grf2d_p3.npy')


def gpu_problem(n=6, p=3, seed=5, mesh=None):
    w = pr.diffusion_grf_2d(n, p, seed, mesh=mesh)
    prob = hf.HDGProblem(2, p)
    pr.setup_gpu(w, prob)
    prob.assemble()
    its, reason, rel = prob.solve(pc=hf.PC_BJACOBI, rtol=1e-12, maxits=20000)
    assert reason == 2
    return w, prob


def oracle_inputs(w, prob):
    rowptr, cols, vals, rhs = prob.trace_csr()
    a = w.mesh.arrays()
    dirf = np.zeros(w.mesh.n_faces, dtype=bool)
    dirf[a['boundary']] = True
    b = w.mesh.sizes['nNf']
    g = lo.graph_from_scalar_csr(rowptr, cols, b, dirf)
    A = lo.blocks_from_scalar_csr(g, rowptr, cols, vals, b)
    lam = prob.solution()[2]
    return g, A, lam, b


@pytest.fixture(scope='module')
def case():
    w, prob = gpu_problem()
    net = LearnedPreconditioner(4)
    net.add_sample(prob)
    g, A, lam, b = oracle_inputs(w, prob)
    sz = net.graph_sizes()
    assert (sz['nodes'], sz['blocks'], sz['edges']) == (g.nb, g.col.size, g.E)
    return dict(w=w, prob=prob, net=net, g=g, A=A, lam=lam, b=b)


@pytest.mark.parametrize('bf16', [False, True])
def test_forward_matches_oracle(case, bf16):
    b = case['b']
    p = random_params(b, 21).astype(np.float32)
    case['net'].set_params(p)
    cn, ce = case['net'].outputs(0, bf16=bf16)
    P = lo.unpack(p.astype(np.float64), b)
    sinv, ahat, dinv = lo.normalise(case['g'], case['A'])
    acts = lo.forward(P, case['g'], ahat)
    scale = max(np.abs(acts['cn']).max(), np.abs(acts['ce']).max())
    tol = 3e-2 if bf16 else 1e-4
    e_n = np.abs(cn - acts['cn']).max() / scale
    e_e = np.abs(ce - acts['ce']).max() / scale
    print('bf16=%s: max rel err nodes %.2e edges %.2e' % (bf16, e_n, e_e))
    assert e_n < tol and e_e < tol
    assert np.all(cn[case['g'].dir] == 0) and np.all(ce[case['g'].dir[case['g'].esrc]] == 0)


@pytest.mark.parametrize('n_random', [0, 2])
def test_loss_and_gradient_match_oracle(case, n_random):
    b = case['b']
    p = random_params(b, 22).astype(np.float32)
    case['net'].set_params(p)
    loss, grad = case['net'].loss_grad(0, n_random=n_random, seed=9)
    loss_o, grad_o, _ = lo.loss_grad(p.astype(np.float64), b, case['g'], case['A'], case['lam'], n_random, 9)
    assert loss == pytest.approx(loss_o, rel=1e-4)
    L = lo.layout(b)
    for k, (o, n) in ((k, v) for k, v in L.items() if k != 'total'):
        go, gg = grad_o[o:o + n], grad[o:o + n].astype(np.float64)
        den = np.linalg.norm(go)
        if den == 0:
            assert np.linalg.norm(gg) < 1e-6
            continue
        err = np.linalg.norm(gg - go) / den
        assert err < 2e-3, (k, err)


def test_training_lowers_loss_and_is_reproducible():
    nets, hist = [], []
    probs = [gpu_problem(6, 3, s) for s in (31, 32, 33)]
    for rep in range(2):
        net = LearnedPreconditioner(4)
        for w, prob in probs:
            net.add_sample(prob)
        net.init_params(seed=3)
        order = np.tile(np.arange(3), 40)
        hist.append(net.train(order, lr=1e-3, n_random=1, seed=100))
        nets.append(net)
    h = hist[0]
    print('loss first %.4e last %.4e' % (h[:6].mean(), h[-6:].mean()))
    assert h[-6:].mean() < 0.7 * h[:6].mean()
    assert np.array_equal(hist[0], hist[1])
    assert np.array_equal(nets[0].params(), nets[1].params())


@pytest.mark.parametrize('bf16', [False, True])
def test_learned_solve_matches_block_jacobi(bf16):
    w, prob = gpu_problem(8, 3, 41)
    lam_ref = prob.solution()[2].copy()
    net = LearnedPreconditioner(4)
    net.add_sample(prob)
    net.init_params(seed=5)
    net.train(np.zeros(30, dtype=np.int32), lr=1e-3, n_random=1, seed=7)
    net.attach(prob, bf16=bf16)
    its, reason, rel = prob.solve(pc=hf.PC_LEARNED, rtol=1e-12, maxits=20000)
    assert reason == 2 and rel <= 1e-12
    lam = prob.solution()[2]
    assert np.abs(lam - lam_ref).max() <= 1e-9 * np.abs(lam_ref).max()
    assert pc_setup_ms(prob) > 0


@pytest.mark.skipif(not os.path.exists(WEIGHTS), reason='no trained network committed')
def test_trained_network_beats_block_jacobi():
    """Held-out GRF coefficients (seeds outside the training range) on a
    training mesh size: fewer BiCGStab iterations to a true residual of
    1e-10 than block Jacobi."""
    mesh = hf.Mesh.structured(2, 3, 64, 0.1, 2024)
    net = LearnedPreconditioner(4)
    net.load(WEIGHTS)
    gains = []
    for seed in (900001, 900002):
        w = pr.diffusion_grf_2d(64, 3, seed, mesh=mesh)
        prob = hf.HDGProblem(2, 3)
        pr.setup_gpu(w, prob)
        prob.assemble()
        prob.set_krylov(hf.KRYLOV_BICGSTAB)
        it_b, r_b, rel_b = prob.solve(pc=hf.PC_BJACOBI, rtol=1e-10, maxits=20000)
        net.attach(prob, bf16=True)
        it_l, r_l, rel_l = prob.solve(pc=hf.PC_LEARNED, rtol=1e-10, maxits=20000)
        print('seed %d: block Jacobi %d its, learned %d its' % (seed, it_b, it_l))
        assert r_b == 2 and r_l == 2 and rel_l <= 1e-10
        gains.append(it_b / it_l)
    assert min(gains) > 1.1
