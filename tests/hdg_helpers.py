"""Shared helpers for the HDG tests: physical elements, exact integration,
block views of the local system, and error norms."""
import itertools

import numpy as np

import oracle as orc


def element_nodes(dim, p, verts):
    """Nodes of the straight order-p simplex with the given vertices (affine
    map of the reference nodes, tools/convertGmsh2H5HO.cpp:367-397)."""
    ref = orc.refel(dim, p)['nodes']
    lam = (ref + 1.0) / 2.0
    V = np.asarray(verts, dtype=float)
    return V[0] + lam @ (V[1:] - V[0])


def rule(dim, order):
    """Cubature rule pinned by the exactness tests (used as the exact integral)."""
    return orc.cubature(dim, order)


def integrate(verts, fn, order):
    """Integral over the physical simplex of fn(x) (x: (n, dim) points)."""
    V = np.asarray(verts, dtype=float)
    dim = V.shape[1]
    pts, w = rule(dim, order)
    lam = (pts + 1.0) / 2.0
    x = V[0] + lam @ (V[1:] - V[0])
    jac = abs(np.linalg.det(V[1:] - V[0])) / 2.0 ** dim
    return float(np.sum(w * fn(x)) * jac)


def integrate_face(fverts, fn, order):
    """Integral over a straight face (dim-1 simplex embedded in dim)."""
    F = np.asarray(fverts, dtype=float)
    fd = F.shape[0] - 1
    pts, w = rule(fd, order) if fd > 1 else orc.cubature(1, order)
    lam = (pts + 1.0) / 2.0
    x = F[0] + lam @ (F[1:] - F[0])
    T = (F[1:] - F[0]).T
    meas = np.sqrt(abs(np.linalg.det(T.T @ T))) / 2.0 ** fd
    return float(np.sum(w * fn(x)) * meas)


def monomials(dim, p):
    return [np.array(e) for e in itertools.product(range(p + 1), repeat=dim) if sum(e) <= p]


def mono(e):
    return lambda x: np.prod(x ** e, axis=-1)


def dmono(e, k):
    def f(x):
        if e[k] == 0:
            return np.zeros(x.shape[:-1])
        ek = e.copy()
        ek[k] -= 1
        return e[k] * np.prod(x ** ek, axis=-1)
    return f


class Blocks:
    """Index helper for the local unknowns [u | q (i*dim+k) | lambda (f*nNf+a)]."""

    def __init__(self, dim, p):
        s = orc.sizes(dim, p)
        self.dim, self.nN, self.nNf, self.nF = dim, s['nN'], s['nNf'], s['nF']
        self.oq = self.nN
        self.ol = self.nN * (dim + 1)

    def u(self):
        return np.arange(self.nN)

    def q(self, k):
        return self.oq + np.arange(self.nN) * self.dim + k

    def lam(self, f):
        return self.ol + f * self.nNf + np.arange(self.nNf)


def face_vertex_coords(dim, p, verts, f):
    """Physical vertices of local face f (its first dim face nodes are vertices)."""
    fn = orc.refel(dim, p)['face_nodes'][f]
    X = element_nodes(dim, p, verts)
    return X[fn[:dim]], X[fn]


def outward_normal(verts, fverts):
    V = np.asarray(verts, dtype=float)
    F = np.asarray(fverts, dtype=float)
    dim = V.shape[1]
    if dim == 2:
        t = F[1] - F[0]
        n = np.array([t[1], -t[0]])
    else:
        n = np.cross(F[1] - F[0], F[2] - F[0])
    n /= np.linalg.norm(n)
    if np.dot(F[0] - V.mean(axis=0), n) < 0:
        n = -n
    return n


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def l2_projection_error(dim, p, mesh_arrays, u_cell, u_exact_nodes):
    """sqrt(sum_e r_e^T M_e r_e) with r the nodal residual
    (TestUtils.h.in:152-191, l2ProjectionCellField)."""
    t = orc.refel(dim, p)
    phi, w, dphi = t['phi'], t['ipw'], t['dphi']
    X = mesh_arrays['coords'][mesh_arrays['cells']]            # (E, nN, dim)
    J = np.einsum('qna,enb->eqab', dphi, X)
    det = np.abs(np.linalg.det(J))
    r = u_exact_nodes - u_cell
    rq = np.einsum('qn,en->eq', phi, r)
    return float(np.sqrt(np.sum(w[None, :] * det * rq * rq)))


def refined_sparse_solve(rowptr, cols, vals, b, steps=2):
    """Solve the CSR system A x = b by sparse LU (SuperLU) plus `steps` of
    iterative refinement with the residual b - A x accumulated in long
    double: the LU alone stops at a relative residual of ~1e-11 on the
    full-size 2-D p = 4 trace system, refinement brings it to ~1e-15.
    Returns (x, final relative residual)."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    n = len(rowptr) - 1
    lu = spla.splu(sp.csr_matrix((vals, cols, rowptr), shape=(n, n)).tocsc())
    row_start = np.asarray(rowptr[:-1], dtype=np.int64)
    bl = np.asarray(b, dtype=np.longdouble)

    def resid(x):
        prod = np.asarray(vals, dtype=np.longdouble) * x[cols]
        return bl - np.add.reduceat(prod, row_start)
    x = lu.solve(b).astype(np.longdouble)
    for _ in range(steps):
        x = x + lu.solve(resid(x).astype(np.float64)).astype(np.longdouble)
    r = resid(x)
    return x.astype(np.float64), float(np.sqrt(np.sum(r * r)) / np.linalg.norm(b))
