"""The nGamma oracle (HDGnGammaModel, HDGnGammaModel.cpp:105-170, 2-D,
m = 2 unknowns n and Gamma per node) pinned on the CPU against the scalar
oracle, itself pinned on the reference's fixtures:

* b = 0 leaves HDGBase + one HDGDiffusion per component: the (n, n) block is
  the scalar diffusion system with D, the (Gamma, Gamma) block the one with G,
  and the components do not couple (tau = t I);
* each component block (a, c) of the convection part is the scalar
  HDGConvection of the velocity the reference builds from the BufferSolution:
  0, b, (c_s^2 - (Gamma/n)^2) b and 2 (Gamma/n) b;
* one Source per component;
* the nonlinear loop solves a manufactured problem of the model's equations,
  div(Gamma b) - D lap n = S_n and div((Gamma^2/n + c_s^2 n) b) - G lap Gamma
  = S_Gamma, in a few Newton iterations with the error falling at order
  p + 1 in h.
"""
import os
import sys

import numpy as np
import pytest

import hyperfox_amd as hf

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, '..', 'oracle'))
import oracle as orc  # noqa: E402
import burgers_oracle as bo  # noqa: E402
from test_burgers_oracle import cell  # noqa: E402

RNG = np.random.default_rng(11)


def sizes_2(p):
    s = orc.sizes(2, p)
    return s, s['nN'] * 2, s['nF'] * s['nNf'] * 2


@pytest.mark.parametrize('p', [1, 2, 3])
def test_no_field_line_flow_is_the_scalar_diffusion_per_component(p):
    s, lu, ll = sizes_2(p)
    x = cell(2, p, RNG)
    D = 0.5 + RNG.random(s['nN'])
    G = 0.5 + RNG.random(s['nN'])
    t = 0.5 + RNG.random(s['nF'] * s['nNf'])
    tau = np.einsum('j,ab->jba', t, np.eye(2)).reshape(-1)
    sol = np.stack([1.0 + RNG.random(s['nN']), RNG.standard_normal(s['nN'])], 1)
    K, F = bo.ngamma_local(p, x, D, G, np.zeros((s['nN'], 2)), sol, tau, 1.3)
    KD, _ = orc.local_system(2, p, 1, x, D=D, ndD=1, tau=t)
    KG, _ = orc.local_system(2, p, 1, x, D=G, ndD=1, tau=t)
    sc = np.abs(KD).max()
    assert np.allclose(K[0::2, 0::2], KD, rtol=0, atol=1e-12 * sc)
    assert np.allclose(K[1::2, 1::2], KG, rtol=0, atol=1e-12 * sc)
    assert not K[0::2, 1::2].any() and not K[1::2, 0::2].any() and not F.any()


@pytest.mark.parametrize('p', [1, 2, 3])
def test_convection_blocks_are_the_scalar_convection_of_the_model_velocities(p):
    s, lu, ll = sizes_2(p)
    x = cell(2, p, RNG)
    t = 0.5 + RNG.random(s['nF'] * s['nNf'])
    tau = np.einsum('j,ab->jba', t, np.eye(2)).reshape(-1)
    b = RNG.standard_normal((s['nN'], 2))
    sol = np.stack([1.0 + RNG.random(s['nN']), RNG.standard_normal(s['nN'])], 1)
    cs = 0.7
    K, _ = bo.ngamma_local(p, x, None, None, b, sol, tau, cs)
    K0, _ = bo.ngamma_local(p, x, None, None, np.zeros_like(b), sol, tau, cs)
    r = sol[:, 1] / sol[:, 0]
    vel = {(0, 0): np.zeros_like(b), (0, 1): b, (1, 0): (cs * cs - r * r)[:, None] * b, (1, 1): (2 * r)[:, None] * b}
    Kref0, _ = orc.local_system(2, p, 2, x, tau=t)
    for (a, c), v in vel.items():
        Kv, _ = orc.local_system(2, p, 2, x, b=v, tau=t)
        conv = Kv - Kref0
        got = (K - K0)[a::2, c::2]
        assert np.allclose(got, conv, rtol=0, atol=1e-12 * max(np.abs(conv).max(), 1.0)), (a, c)
    assert np.abs((K - K0)[1::2, 0::2]).max() > 0


@pytest.mark.parametrize('p', [1, 2])
def test_source_per_component(p):
    s, lu, ll = sizes_2(p)
    x = cell(2, p, RNG)
    c = RNG.standard_normal(2)
    tau = np.zeros(s['nF'] * s['nNf'] * 4)
    sol = np.ones((s['nN'], 2))
    _, F = bo.ngamma_local(p, x, None, None, np.zeros((s['nN'], 2)), sol, tau, 1.0, np.tile(c, (s['nIP'], 1)))
    R = orc.refel(2, p)
    J = np.einsum('qik,ic->qkc', R['dphi'], x)[0]
    vol = R['ipw'].sum() * abs(np.linalg.det(J))
    assert np.allclose(F[:lu].reshape(-1, 2).sum(0), c * vol, rtol=1e-12, atol=0)
    assert not F[lu:].any()


# ------------------------------------------------------- manufactured problem
A_N, B_G, CS, DIF, GDIF = 0.3, 0.2, 1.0, 0.1, 0.15
BDIR = np.array([np.cos(0.4), np.sin(0.4)])


def exact(X):
    x, y = X[:, 0], X[:, 1]
    n = 1.0 + A_N * np.sin(np.pi * x) * np.sin(np.pi * y)
    g = B_G * np.cos(np.pi * x) * np.sin(np.pi * y)
    return np.stack([n, g], 1)


def source(X):
    """S_n = b.grad Gamma - D lap n, S_Gamma = b.grad(Gamma^2/n + c^2 n) - G lap Gamma (b constant)."""
    x, y = X[:, 0], X[:, 1]
    s1, c1, s2, c2 = np.sin(np.pi * x), np.cos(np.pi * x), np.sin(np.pi * y), np.cos(np.pi * y)
    n = 1.0 + A_N * s1 * s2
    g = B_G * c1 * s2
    gn = A_N * np.pi * np.stack([c1 * s2, s1 * c2], 1)
    gg = B_G * np.pi * np.stack([-s1 * s2, c1 * c2], 1)
    lap_n = -2 * np.pi ** 2 * A_N * s1 * s2
    lap_g = -2 * np.pi ** 2 * g
    gF = (2 * g / n)[:, None] * gg - (g * g / (n * n))[:, None] * gn + CS * CS * gn
    return np.stack([gg @ BDIR - DIF * lap_n, gF @ BDIR - GDIF * lap_g], 1)


def manufactured(n, p):
    mesh = hf.Mesh.structured(2, p, n)
    a = mesh.arrays()
    s = mesh.sizes
    R = orc.refel(2, p)
    xq = np.einsum('qi,eic->eqc', R['phi'], a['coords'][a['cells']])
    f = source(xq.reshape(-1, 2)).reshape(mesh.n_cells, s['nIP'], 2)
    tau = np.zeros((mesh.n_faces, s['nNf'], 2, 2, 2))
    tau[:] = np.eye(2) * (2.0 + DIF * n)
    g = exact(a['coords'][a['faces'].reshape(-1)]).reshape(mesh.n_faces, s['nNf'], 2)
    nn = mesh.n_nodes
    ng = dict(G=np.full(nn, GDIF), b=np.tile(BDIR, (nn, 1)), cs=CS)
    sysm = bo.BurgersSystem(2, p, a, tau, Dn=np.full(nn, DIF), f_ip=f, dirichlet=g, ngamma=ng)
    return mesh, a, sysm, f, tau, g, ng


@pytest.mark.parametrize('p', [1, 2])
def test_manufactured_problem_converges(p):
    errs = []
    for n in (4, 8):
        mesh, a, sysm, *_ = manufactured(n, p)
        init = np.tile([1.0, 0.0], (mesh.n_cells, sysm.s['nN']))
        sol, buf, flux, lam, its, res = sysm.newton(init, init, np.zeros(sysm.ndofs), max_iters=20, tol=1e-10)
        assert res < 1e-10 and its <= 8
        ex = exact(a['coords'][a['cells'].reshape(-1)]).reshape(mesh.n_cells, -1)
        errs.append(np.linalg.norm(sol - ex) / np.linalg.norm(ex))
    # measured 1.2e-2 -> 3.4e-3 (p = 1), 1.8e-3 -> 2.4e-4 (p = 2), 5-8 Newton iterations
    assert errs[1] < errs[0] / 3 and errs[0] < [2e-2, 2e-3][p - 1]
