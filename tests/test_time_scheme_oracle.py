"""Implicit Euler on the u rows (Euler::apply, isExplicit = false,
src/operator/Euler.cpp:18-37, applied by HDGModel::compute,
src/model/HDGModel.cpp:39-47) in the CPU oracle, pinned by what the
mathematics fixes.  No GPU.

* Heat equation u_t = Laplace u on the unit square, u = 0 on the boundary,
  u(0) = sin(pi x) sin(pi y): backward Euler multiplies this eigenfunction by
  1 / (1 + 2 pi^2 dt) per step, so after n steps u = (1 + 2 pi^2 dt)^-n u(0)
  up to the HDG space error, which falls at the HDG rate O(h^(p+1)).
* A huge step forgets u_old: the solution is the steady one.
* The scheme acts on the u rows only: q and trace rows are unchanged.
"""
import numpy as np
import pytest

import hyperfox_amd as hf
import oracle as orc


def heat_system(n, p, f_steady=False):
    m = hf.Mesh.structured(2, p, n, 0.1, 11)
    a = m.arrays()
    nNf = m.sizes['nNf']
    f_ip = None
    if f_steady:
        xip = np.stack([orc.ip_coords(2, p, a['coords'][c]) for c in a['cells']])
        f_ip = 2 * np.pi ** 2 * np.sin(np.pi * xip[..., 0]) * np.sin(np.pi * xip[..., 1])
    s = orc.System(2, p, 1, a, np.ones((m.n_faces, nNf)), tau_sides=1, f_ip=f_ip,
                   dirichlet=np.zeros((m.n_faces, nNf)))
    X = a['coords'][a['cells']]
    return s, np.sin(np.pi * X[..., 0]) * np.sin(np.pi * X[..., 1])


def euler_steps(s, u, dt, steps):
    for _ in range(steps):
        s.set_euler(dt, u)
        lam = s.solve(method=1)[0]
        u = s.reconstruct(lam)[0]
    return u


def test_heat_eigenfunction_decays_at_the_backward_euler_rate():
    dt, steps = 0.01, 5
    amp = (1 + 2 * np.pi ** 2 * dt) ** -steps
    errs = []
    for n in (4, 8):
        s, u0 = heat_system(n, 3)
        u = euler_steps(s, u0, dt, steps)
        errs.append(np.abs(u - amp * u0).max() / (amp * np.abs(u0).max()))
    print('relative errors at h = 1/4, 1/8: %.2e %.2e' % tuple(errs))
    assert errs[1] < 1e-3
    assert errs[0] / errs[1] > 8  # HDG rate h^(p+1) = h^4 (a dropped term would not converge)


def test_huge_step_gives_the_steady_solution():
    s, u0 = heat_system(6, 2, f_steady=True)
    lam_steady = s.solve(method=1)[0]
    u_steady = s.reconstruct(lam_steady)[0]
    rng = np.random.default_rng(3)
    u = euler_steps(s, rng.normal(size=u0.shape), 1e10, 1)
    assert np.abs(u - u_steady).max() <= 1e-7 * np.abs(u_steady).max()
    assert np.abs(u_steady - u0).max() < 5e-2  # the steady solution approximates sin sin


@pytest.mark.parametrize('dim,p', [(2, 2), (3, 2)])
def test_scheme_touches_the_u_rows_only(dim, p):
    rng = np.random.default_rng(4)
    s = orc.sizes(dim, p)
    n, nN = s['n_local'], s['nN']
    m = hf.Mesh.structured(dim, p, 1, 0.0, 1)
    a = m.arrays()
    x = a['coords'][a['cells'][0]]
    tau = rng.uniform(0.5, 2, s['nF'] * s['nNf'])
    K, F = orc.local_system(dim, p, 1, x, tau=tau, f_ip=rng.normal(size=s['nIP']))
    uo = rng.normal(size=nN)
    K2, F2 = orc.apply_euler(dim, p, 0.25, uo, K, F)
    assert np.array_equal(K2[nN:], K[nN:]) and np.array_equal(F2[nN:], F[nN:])
    # u rows: 0.25 x (model rows) + [M | 0 | 0], M = the first block of Sqq = I_d (x) M
    M = K[nN::dim, nN::dim][:nN, :nN]
    assert np.allclose(K2[:nN, nN:], 0.25 * K[:nN, nN:], rtol=1e-15, atol=0)
    assert np.allclose(K2[:nN, :nN] - 0.25 * K[:nN, :nN], M, rtol=1e-13, atol=1e-16)
    assert np.allclose(F2[:nN], 0.25 * F[:nN] + M @ uo, rtol=1e-13, atol=1e-15)
    with pytest.raises(orc.OracleError, match='time step'):
        orc.apply_euler(dim, p, 0.0, uo, K, F)
