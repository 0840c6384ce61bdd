"""FP64 numpy restatement of HDGBohmModel (src/model/HDGBohmModel.cpp:97-268),
the Bohm-Chodura boundary model of the 2-D (n, Gamma) problem, and of the
transfer functions of its regression test
(tests/regression/HDG/TestHDGBohm.cpp:37-130).

TEST INFRASTRUCTURE ONLY: the checker of the GPU path's Bohm rows; the
product never imports it.

One boundary face at a time, on the face reference element (nNf nodes, the
face cubature of the 2-D element), with the face-local fields that
HDGSolver::applyBoundaryConditions hands the model (HDGSolver.cpp:404-476):
ExteriorNormals, Trace and Tau (side 0) in face-node order; Solution and
Flux of the face's cell at the face nodes; b, D, G and the sound speed at the
face nodes.  Returns the model's localMatrix B = [B_u | B_q | B_lambda]
(2 nNf x 8 nNf; columns u (j 2 + a), q ((j 2 + k) 2 + a), lambda (j 2 + a))
and localRHS r = (B - originalSystem) U0, U0 = [Solution | Flux | Trace].
"""
import numpy as np

import oracle as orc

TOL = 1e-5       # the |y| tolerance of the test's transfer functions
DELTA = 0.001    # the tanh width of numHeavyside


def num_heaviside(x):
    return 0.5 * (1.0 + np.tanh(x / DELTA))


def num_deriv_heaviside(x):
    return (0.5 / DELTA) * (1.0 - np.tanh(x / DELTA) ** 2)


def transfer(kind, x, y):
    """Transfer function value and [d/dx, d/dy] (TestHDGBohm.cpp:37-130):
    0 neumann (0), 1 dirichlet (1), 2 heaviside (a step, zero derivative),
    3 tanh (the smoothed step)."""
    if kind == 0:
        return 0.0, (0.0, 0.0)
    if kind == 1:
        return 1.0, (0.0, 0.0)
    if kind == 2:
        res = 0.0
        if abs(y) > TOL:
            if (x > 0) != (y > 0):
                res = 1.0
            elif abs(x) - abs(y) < 0.0:
                res = 1.0
        return res, (0.0, 0.0)
    if kind == 3:
        H = num_heaviside
        dH = num_deriv_heaviside
        Hx, Hy, Hmx, Hmy = H(x), H(y), H(-x), H(-y)
        dHx, dHy, dHmx, dHmy = dH(x), dH(y), dH(-x), dH(-y)
        Htol, dHtol = H(abs(TOL) - abs(y)), dH(abs(TOL) - abs(y))
        base = Hx * Hmy + Hmx * Hy + Hx * Hy * H(y - x) + Hmx * Hmy * H(x - y)
        val = base * (1.0 - Htol)
        d0 = (dHx * Hmy - dHmx * Hy + dHx * Hy * H(y - x) - Hx * Hy * dH(y - x)
              - dHmx * Hmy * H(x - y) + Hmx * Hmy * dH(x - y)) * (1.0 - Htol)
        d1 = (-Hx * dHmy + Hmx * dHy + Hx * dHy * H(y - x) + Hx * Hy * dH(y - x)
              - Hmx * dHmy * H(x - y) - Hmx * Hmy * dH(x - y)) * (1.0 - Htol) + base * dHtol
        return val, (d0, d1)
    raise ValueError('unknown transfer function')


def face_tables(p):
    """Face reference element of the 2-D order-p simplex: Lagrange basis on
    its nNf nodes (reference coordinate in [-1, 1], from the 2-D element's
    first face) and its derivative at the face cubature points, with the
    weights (the oracle's 1-D rule of the face element)."""
    t = orc.refel(2, p)
    fn = t['face_nodes'][0]
    x2 = t['nodes'][fn]                                  # on the face (-1,-1) -> (1,-1)
    s = x2[:, 0]                                         # reference coordinate along the face
    nNf = len(fn)
    V = np.vander(s, nNf, increasing=True)
    C = np.linalg.inv(V)                                 # phi_j(t) = sum_k C[k, j] t^k
    tq = t['fipx'][:, 0]
    P = np.vander(tq, nNf, increasing=True)
    dP = np.zeros_like(P)
    dP[:, 1:] = P[:, :-1] * np.arange(1, nNf)
    return P @ C, dP @ C, t['fipw']


def bohm_local(p, xf, normals, b, trace, flux, tau, D, G, sol, cs, kind):
    """HDGBohmModel::computeLocalMatrix and computeLocalRHS on one face.
    xf, normals, b, trace, sol: (nNf, 2); flux (nNf, 4), node values
    [k 2 + a] (the local flux order); tau (nNf, 4), column-major 2 x 2 per
    node; D, G: (nNf,) scalars (identity times the value) or (nNf, 4); cs:
    (nNf,) sound speed.  Returns (B, r, original system)."""
    phi, dphi, w = face_tables(p)
    nN = phi.shape[1]
    nq = len(w)
    J = dphi @ xf                                        # (nq, 2) face Jacobians
    dV = w * np.sqrt(np.einsum('qc,qc->q', J, J))        # w sqrt(det(J J^T)) (Operator.cpp:61-81)

    def mat(v):                                          # node values -> (nN, 4) column-major 2 x 2
        v = np.asarray(v, dtype=float)
        if v.ndim == 1:
            return v[:, None] * np.array([1.0, 0.0, 0.0, 1.0])[None, :]
        return v

    def cm(v):                                           # 4 values, column-major -> 2 x 2
        return np.array([[v[0], v[2]], [v[1], v[3]]])
    Dm, Gm = mat(D), mat(G)
    B = np.zeros((2 * nN, 8 * nN))
    orig = np.zeros_like(B)
    for q in range(nq):
        sh = phi[q]
        n = sh @ normals
        csq = sh @ cs
        bn = (sh @ b) @ n
        tr = sh @ trace
        F = cm(sh @ flux)                                # F[a][k] = q_{a,k}
        ta = cm(sh @ tau)
        Dq, Gq = cm(sh @ Dm), cm(sh @ Gm)
        u = sh @ sol
        T, dT = transfer(kind, u[1] * bn, u[0] * csq * bn)
        Dn, Gn = Dq @ n, Gq @ n
        for nd in range(nN):
            for md in range(nN):
                db = dV[q] * sh[nd] * sh[md]
                # n condition
                orig[md * 2, nd * 2] += ta[0, 0] * db
                orig[md * 2, nd * 2 + 1] += ta[0, 1] * db
                orig[md * 2, nN * 6 + nd * 2] -= ta[0, 0] * db
                orig[md * 2, nN * 6 + nd * 2 + 1] -= ta[0, 1] * db
                for d in range(2):
                    orig[md * 2, 2 * nN + (nd * 2 + d) * 2] -= db * Dn[d]
                # Gamma condition: Neumann part, Dirichlet part, and the
                # derivative of the transfer function
                orig[md * 2 + 1, nd * 2] += ta[1, 0] * (1 - T) * db
                orig[md * 2 + 1, nd * 2 + 1] += ta[1, 1] * (1 - T) * db
                orig[md * 2 + 1, nN * 6 + nd * 2] -= ta[1, 0] * (1 - T) * db
                orig[md * 2 + 1, nN * 6 + nd * 2 + 1] -= ta[1, 1] * (1 - T) * db
                for d in range(2):
                    orig[md * 2 + 1, 2 * nN + (nd * 2 + d) * 2 + 1] -= (1 - T) * db * Gn[d]
                orig[md * 2 + 1, nd * 2] -= T * db * csq * abs(bn)
                orig[md * 2 + 1, nd * 2 + 1] += T * db * bn
                dirv = db * (u[1] * bn - u[0] * csq * abs(bn))
                neuv = db * (-(Gn @ F[1]) + ta[1, 1] * (u[1] - tr[1]) + ta[1, 0] * (u[0] - tr[0]))
                B[md * 2 + 1, nd * 2] += (csq * bn * dT[1]) * (dirv - neuv)
                B[md * 2 + 1, nd * 2 + 1] += (bn * dT[0]) * (dirv - neuv)
    B += orig
    U0 = np.concatenate([np.ravel(sol), np.ravel(flux), np.ravel(trace)])
    return B, (B - orig) @ U0, orig


def bohm_residual(p, xf, normals, b, trace, flux, tau, D, G, sol, cs, kind):
    """The nonlinear face residual that the Bohm rows linearise, written from
    the fields at the face cubature points (not from the matrix): row (m, 0)
    int phi_m [tau_0. (u - lambda) - (D n).q_n], row (m, 1)
    int phi_m [(1 - T) (tau_1. (u - lambda) - (G n).q_Gamma)
    + T (u_Gamma b.n - u_n c_s |b.n|)] with T = T(u_Gamma b.n, u_n c_s b.n)."""
    phi, dphi, w = face_tables(p)
    nN = phi.shape[1]
    J = dphi @ xf
    dV = w * np.sqrt(np.einsum('qc,qc->q', J, J))
    Dm = D[:, None] * np.array([1.0, 0.0, 0.0, 1.0]) if np.ndim(D) == 1 else D
    Gm = G[:, None] * np.array([1.0, 0.0, 0.0, 1.0]) if np.ndim(G) == 1 else G
    out = np.zeros(2 * nN)
    for q in range(len(w)):
        sh = phi[q]
        n = sh @ normals
        csq = sh @ cs
        bn = (sh @ b) @ n
        u, lam = sh @ sol, sh @ trace
        fl = sh @ flux                                   # [k 2 + a]
        qn = np.array([fl[0], fl[2]])                    # q of n, components k
        qg = np.array([fl[1], fl[3]])                    # q of Gamma
        tv = sh @ tau
        tau0 = np.array([tv[0], tv[2]])                  # row 0 of the column-major 2 x 2
        tau1 = np.array([tv[1], tv[3]])
        dv = sh @ Dm
        gv = sh @ Gm
        Dn = np.array([dv[0] * n[0] + dv[2] * n[1], dv[1] * n[0] + dv[3] * n[1]])
        Gn = np.array([gv[0] * n[0] + gv[2] * n[1], gv[1] * n[0] + gv[3] * n[1]])
        T, _ = transfer(kind, u[1] * bn, u[0] * csq * bn)
        r0 = tau0 @ (u - lam) - Dn @ qn
        r1 = (1 - T) * (tau1 @ (u - lam) - Gn @ qg) + T * (u[1] * bn - u[0] * csq * abs(bn))
        out[0::2] += dV[q] * sh * r0
        out[1::2] += dV[q] * sh * r1
    return out
