"""CPU oracle of the multi-DOF nonlinear path: HDGBurgersModel (m = dim) and
HDGnGammaModel (2-D, m = 2) on HDGSolver (IMPLICIT) under NonLinearWrapper.

TEST INFRASTRUCTURE ONLY: loaded by tests/, never by the product
(hyperfox_amd).  The local systems come from liboracle_hdg.so
(orc_burgers_local, HDGBurgersModel.cpp:86-132; orc_ngamma_local,
HDGnGammaModel.cpp:105-170; orc_condense_sizes,
HDGSolver.cpp:331-348); the trace system is assembled here with numpy and
solved with a sparse direct solver:

* local trace index (f nNf + j) m + a -> global (F nNf + p) m + a, p the
  position of the cell's face node j in the face's node list
  (HDGSolver.cpp:531-675);
* boundary faces: IntegratedDirichletModel rows M (x) I_m and M g
  (IntegratedDirichletModel.cpp:35-56, the Mass operator with m unknowns per
  node) or DirichletModel rows I and g (DirichletModel.cpp); faces of an
  HDG-type model (HDGSolver.cpp:503-518) get B_u U_F + B_q Q_F over the
  cell's trace columns plus B_lambda on their own block, and
  r - B_u U0_F - B_q Q0_F, Set (rows zeroed first) or Add;
* implicit Euler (RungeKutta BEuler, RungeKutta.cpp:113-159, its first and
  only stage): the u rows of K and F_u times dt, then (M (x) I_m) added to the
  u-u block and (M (x) I_m) u_old to F_u, M the scalar mass (the first
  component block of Sqq = I (x) M);
* u = U lambda_e + U0, q = Q lambda_e + Q0 (HDGSolver.cpp:741-775);
* NonLinearWrapper::solve (NonLinearWrapper.cpp:41-79) with the vanilla
  residual ||u - u_buf|| / ||u_buf|| (||u - u_buf|| when u_buf = 0) and the
  damped update u = u_buf = (1 - d) u + d u_buf; the next linearisation
  reads the last solved trace.
"""
import ctypes

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from oracle import _chk, _p, face_mass, lib, refel, sizes

lib.orc_burgers_local.argtypes = [ctypes.c_int, ctypes.c_int] + [ctypes.c_void_p] * 8
lib.orc_burgers_local.restype = ctypes.c_int
lib.orc_ngamma_local.argtypes = [ctypes.c_int] + [ctypes.c_void_p] * 6 + [ctypes.c_double] + [ctypes.c_void_p] * 3
lib.orc_ngamma_local.restype = ctypes.c_int
lib.orc_condense_sizes.argtypes = [ctypes.c_int] * 3 + [ctypes.c_void_p] * 8
lib.orc_condense_sizes.restype = ctypes.c_int


def _c(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def burgers_local(dim, p, x, Dn, tau, sol, trace, f_ip=None):
    """(K, F) of one cell; see orc_burgers_local for the layouts."""
    s = sizes(dim, p)
    m = dim
    n = s['nN'] * m * (dim + 1) + s['nF'] * s['nNf'] * m
    K = np.zeros((n, n))
    F = np.zeros(n)
    args = [_c(v) for v in (x, Dn, tau, sol, trace, f_ip)]
    _chk(lib.orc_burgers_local(dim, p, *[_p(v) for v in args], _p(K), _p(F)))
    return K, F


def ngamma_local(p, x, D, G, b, sol, tau, cs, f_ip=None):
    """(K, F) of one 2-D cell of HDGnGammaModel; see orc_ngamma_local."""
    s = sizes(2, p)
    n = 2 * (s['nN'] * 3 + s['nF'] * s['nNf'])
    K = np.zeros((n, n))
    F = np.zeros(n)
    args = [_c(v) for v in (x, D, G, b, sol, tau)]
    fi = _c(f_ip)
    _chk(lib.orc_ngamma_local(p, *[_p(v) for v in args], float(cs), _p(fi), _p(K), _p(F)))
    return K, F


def condense_sizes(lu, lq, ll, K, F):
    """(S, S0, U, Q, U0, Q0) of HDGSolver's static condensation."""
    K = _c(K)
    F = _c(F)
    S, S0 = np.zeros((ll, ll)), np.zeros(ll)
    U, Q, U0, Q0 = np.zeros((lu, ll)), np.zeros((lq, ll)), np.zeros(lu), np.zeros(lq)
    _chk(lib.orc_condense_sizes(lu, lq, ll, _p(K), _p(F), _p(S), _p(S0), _p(U), _p(Q), _p(U0), _p(Q0)))
    return S, S0, U, Q, U0, Q0


def implicit_euler(dim, m, nN, K, F, dt, u_old):
    """RungeKutta::apply for BEuler on the u rows of an m-unknown local system."""
    lu = nN * m
    oq = lu
    Msc = K[oq:oq + nN * dim * m:dim * m, oq:oq + nN * dim * m:dim * m]  # Sqq = I (x) M: first component block
    Mm = np.kron(Msc, np.eye(m))
    K2, F2 = K.copy(), F.copy()
    K2[:lu] *= dt
    F2[:lu] *= dt
    K2[:lu, :lu] += Mm
    F2[:lu] += Mm @ u_old
    return K2, F2


def face_positions(dim, p, cells, faces, cell2face):
    """pos[e, f, j]: position of cell e's local face f node j in the node list
    of its face (the orientation HDGSolver reads from the mesh)."""
    fmap = refel(dim, p)['face_nodes']
    nC, nF = cell2face.shape
    pos = np.zeros((nC, nF, fmap.shape[1]), dtype=np.int64)
    for e in range(nC):
        for f in range(nF):
            row = list(faces[cell2face[e, f]])
            for j, ln in enumerate(fmap[f]):
                pos[e, f, j] = row.index(cells[e, ln])
    return pos


class BurgersSystem:
    """HDGSolver over HDGBurgersModel on a whole mesh (numpy arrays of the
    mesh: coords, cells, faces, cell2face, face2cell).

    tau: (nFaces, nNf, sides, m, m) with tau[F, p, s][a, b] = tau_ab of the
    face's node p on the side of cell face2cell[F, s]; Dn: nNodes scalars or
    None; f_ip: (nCells, nIP, m) or None; dirichlet: (nFaces, nNf, m);
    boundary: 'integrated' or 'dirichlet'.  ngamma: a dict with the
    HDGnGammaModel fields G (nNodes scalars or None), b (nNodes, 2) and cs
    (sound speed); Dn is then its D field."""

    def __init__(self, dim, p, mesh, tau, Dn=None, f_ip=None, dirichlet=None, boundary='integrated', ngamma=None,
                 hdg=None):
        self.dim, self.p, self.m = dim, p, dim
        self.ngamma = ngamma
        self.hdg = hdg  # dict(faces, B (n, bs, bs (2 + dim)), r (n, bs), mset, rset)
        self.bohm = None  # dict(faces, normals (n, nNf, 2), kind): HDGBohmModel rows rebuilt before each assembly
        self.dt = None  # implicit Euler step, with self.u_old (nCells x nN m)
        self.u_old = None
        self.s = sizes(dim, p)
        self.mesh = {k: np.asarray(mesh[k]) for k in ('coords', 'cells', 'faces', 'cell2face', 'face2cell')}
        M = self.mesh
        self.nC, self.nFaces = M['cells'].shape[0], M['faces'].shape[0]
        self.pos = face_positions(dim, p, M['cells'], M['faces'], M['cell2face'])
        self.tau = np.asarray(tau, dtype=np.float64)
        self.Dn = None if Dn is None else np.asarray(Dn, dtype=np.float64)
        self.f_ip = None if f_ip is None else np.asarray(f_ip, dtype=np.float64)
        nNf, m = self.s['nNf'], self.m
        self.g = np.zeros((self.nFaces, nNf, m)) if dirichlet is None else np.asarray(dirichlet, dtype=np.float64)
        self.bfaces = np.nonzero(M['face2cell'][:, 1] < 0)[0]
        self.boundary = boundary
        self.nL = self.s['nF'] * nNf * m
        self.ndofs = self.nFaces * nNf * m
        # global trace index of every local trace entry of every cell
        f = np.arange(self.s['nF'])[None, :, None, None]
        a = np.arange(m)[None, None, None, :]
        F = M['cell2face'][:, :, None, None]
        self.gidx = ((F * nNf + self.pos[:, :, :, None]) * m + a).reshape(self.nC, -1)
        del f

    def local_fields(self, e, sol, trace):
        """Cell e's tau (side of e, local node order), BufferSolution and trace."""
        M, s, m = self.mesh, self.s, self.m
        nNf = s['nNf']
        taus = np.zeros((s['nF'], nNf, m * m))
        for f in range(s['nF']):
            F = M['cell2face'][e, f]
            side = 0 if (self.tau.shape[2] == 1 or M['face2cell'][F, 0] == e) else 1
            for j in range(nNf):
                taus[f, j] = self.tau[F, self.pos[e, f, j], side].T.reshape(-1)  # column-major m x m
        lam = np.asarray(trace).reshape(-1)[self.gidx[e]]
        return taus.reshape(-1), np.asarray(sol).reshape(self.nC, -1)[e], lam

    def assemble(self, sol, trace):
        """Condensed blocks of every cell and the global trace system."""
        M, s, m, dim = self.mesh, self.s, self.m, self.dim
        nN, nNf = s['nN'], s['nNf']
        lu, lq, ll = nN * m, nN * dim * m, self.nL
        self.U = np.zeros((self.nC, lu, ll))
        self.Q = np.zeros((self.nC, lq, ll))
        self.U0 = np.zeros((self.nC, lu))
        self.Q0 = np.zeros((self.nC, lq))
        self.S = np.zeros((self.nC, ll, ll))
        self.S0 = np.zeros((self.nC, ll))
        for e in range(self.nC):
            x = M['coords'][M['cells'][e]]
            taus, so, lam = self.local_fields(e, sol, trace)
            Dn = None if self.Dn is None else self.Dn[M['cells'][e]]
            fi = None if self.f_ip is None else self.f_ip[e]
            if self.ngamma is None:
                K, F = burgers_local(dim, self.p, x, Dn, taus, so, lam, fi)
            else:
                ng = self.ngamma
                Gn = None if ng.get('G') is None else np.asarray(ng['G'])[M['cells'][e]]
                K, F = ngamma_local(self.p, x, Dn, Gn, np.asarray(ng['b'])[M['cells'][e]], so, taus, ng['cs'], fi)
            if self.dt:
                K, F = implicit_euler(dim, m, nN, K, F, self.dt, np.asarray(self.u_old).reshape(self.nC, -1)[e])
            (self.S[e], self.S0[e], self.U[e], self.Q[e], self.U0[e], self.Q0[e]) = condense_sizes(lu, lq, ll, K, F)
        rows = np.repeat(self.gidx, ll, axis=1).reshape(-1)
        cols = np.tile(self.gidx, (1, ll)).reshape(-1)
        A = sp.coo_matrix((self.S.reshape(-1), (rows, cols)), shape=(self.ndofs, self.ndofs)).tocsr()
        b = np.zeros(self.ndofs)
        np.add.at(b, self.gidx.reshape(-1), self.S0.reshape(-1))
        # boundary rows
        A = A.tolil()
        bs = nNf * m
        hdg_faces = set() if self.hdg is None else set(int(F) for F in self.hdg['faces'])
        if self.hdg is not None:
            fmap = refel(dim, self.p)['face_nodes']
            for ib, F in enumerate(self.hdg['faces']):
                e = M['face2cell'][F, 0]
                lf = list(M['cell2face'][e]).index(F)
                node_at = {int(self.pos[e, lf, j]): int(fmap[lf][j]) for j in range(nNf)}
                rows_u = [node_at[pj] * m + b for pj in range(nNf) for b in range(m)]
                rows_q = [(node_at[pj] * dim + k) * m + b for pj in range(nNf) for k in range(dim) for b in range(m)]
                Bi = np.asarray(self.hdg['B'][ib])
                Bu, Bq, Bl = Bi[:, :bs], Bi[:, bs:bs * (1 + dim)], Bi[:, bs * (1 + dim):]
                rowvals = Bu @ self.U[e][rows_u] + Bq @ self.Q[e][rows_q]  # bs x local trace columns
                rhsv = np.asarray(self.hdg['r'][ib]) - Bu @ self.U0[e][rows_u] - Bq @ self.Q0[e][rows_q]
                r0 = F * bs
                if self.hdg.get('mset', True):
                    A[r0:r0 + bs, :] = 0.0
                blk = np.zeros((bs, self.ndofs))
                np.add.at(blk, (slice(None), self.gidx[e]), rowvals)
                blk[:, r0:r0 + bs] += Bl
                A[r0:r0 + bs, :] = A[r0:r0 + bs, :].toarray() + blk
                b[r0:r0 + bs] = rhsv if self.hdg.get('rset', True) else b[r0:r0 + bs] + rhsv
        for F in self.bfaces:
            if int(F) in hdg_faces:
                continue
            r0 = F * bs
            A[r0:r0 + bs, :] = 0.0
            gF = self.g[F].reshape(-1)
            if self.boundary == 'integrated':
                Mf = face_mass(dim, self.p, M['coords'][M['faces'][F]])
                blk = np.kron(Mf, np.eye(m))
                A[r0:r0 + bs, r0:r0 + bs] = blk
                b[r0:r0 + bs] = blk @ gF
            else:
                A[r0:r0 + bs, r0:r0 + bs] = np.eye(bs)
                b[r0:r0 + bs] = gF
        self.A = A.tocsr()
        self.b = b
        return self.A, self.b

    def solve(self):
        return spla.spsolve(self.A.tocsc(), self.b)

    def reconstruct(self, lam):
        L = np.asarray(lam).reshape(-1)[self.gidx]
        u = np.einsum('eij,ej->ei', self.U, L) + self.U0
        q = np.einsum('eij,ej->ei', self.Q, L) + self.Q0
        return u, q

    def bohm_rows(self, sol, flux, trace):
        """HDGBohmModel::compute on every face of self.bohm with the face-local
        fields of HDGSolver::applyBoundaryConditions (HDGSolver.cpp:404-476):
        Solution and Flux of the face's cell at the face nodes, Trace, Tau of
        side 0 and the ExteriorNormals in face-node order, b, D, G at the face
        nodes; returns the HDG-type rows as self.hdg takes them (Set / Set)."""
        import bohm_oracle
        M, nN, nNf = self.mesh, self.s['nN'], self.s['nNf']
        fmap = refel(self.dim, self.p)['face_nodes']
        ng = self.ngamma
        sol = np.asarray(sol).reshape(self.nC, nN, 2)
        flux = np.asarray(flux).reshape(self.nC, nN, 4)
        trace = np.asarray(trace).reshape(self.nFaces, nNf, 2)
        Bs, rs = [], []
        for ib, F in enumerate(self.bohm['faces']):
            e = M['face2cell'][F, 0]
            lf = list(M['cell2face'][e]).index(F)
            node_at = {int(self.pos[e, lf, j]): int(fmap[lf][j]) for j in range(nNf)}
            cn = [node_at[j] for j in range(nNf)]
            nodes = M['faces'][F]
            B, r, _ = bohm_oracle.bohm_local(
                self.p, M['coords'][nodes], np.asarray(self.bohm['normals'][ib]), np.asarray(ng['b'])[nodes],
                trace[F], flux[e][cn], np.stack([self.tau[F, j, 0].T.reshape(-1) for j in range(nNf)]),
                self.Dn[nodes], np.asarray(ng['G'])[nodes], sol[e][cn], np.full(nNf, ng['cs']), self.bohm['kind'])
            Bs.append(B)
            rs.append(r)
        return dict(faces=list(self.bohm['faces']), B=np.array(Bs), r=np.array(rs), mset=True, rset=True)

    def newton(self, sol, buf, trace, max_iters=20, tol=1e-6, damp=0.0, flux=None):
        """NonLinearWrapper::solve: returns (sol, buf, flux, trace, iterations,
        residual).  With a Bohm model, its rows come from the Solution, Flux
        (flux: the starting Flux field) and Trace of each iteration."""
        sol = np.array(sol, dtype=np.float64).reshape(self.nC, -1)
        buf = np.array(buf, dtype=np.float64).reshape(self.nC, -1)
        trace = np.array(trace, dtype=np.float64).reshape(-1)
        if flux is None:
            flux = np.zeros((self.nC, self.s['nN'] * self.dim * self.m))
        res = np.inf
        it = 0
        for it in range(1, max_iters + 1):
            if self.bohm is not None:
                self.hdg = self.bohm_rows(sol, flux, trace)
            self.assemble(buf, trace)
            trace = self.solve()
            sol, flux = self.reconstruct(trace)
            dn, rn = np.sum((sol - buf) ** 2), np.sum(buf ** 2)
            res = np.sqrt(dn / rn) if rn != 0 else np.sqrt(dn)
            if res < tol:
                break
            sol = (1.0 - damp) * sol + damp * buf
            buf = sol.copy()
        return sol, buf, flux, trace, it, res


class MultiSystem(BurgersSystem):
    """A scalar HDG model at m = dim DOFs per node (HDGModel::allocate with
    nDOFsPerNode = m, HDGOperator::multiplyDOFs, HDGOperator.cpp:17-28):
    HDGLaplaceModel when Dn is None (D = 1), the diffusion-source model with
    the scalar nodal Dn otherwise, one source per component and the m x m tau
    on the faces.  It is the BurgersSystem with no solution and no trace,
    where HDGUNabU vanishes: the no-flow local system is the scalar oracle's
    times I_m and a tau matrix enters as kron(scalar tau part, tau), both
    pinned in test_burgers_oracle.py.  Boundary rows: DirichletModel
    ('dirichlet', as TestHDGLaplaceMulti.cpp) or 'integrated'."""

    def __init__(self, dim, p, mesh, tau, Dn=None, f_ip=None, dirichlet=None, boundary='dirichlet'):
        nn = np.asarray(mesh['coords']).shape[0]
        super().__init__(dim, p, mesh, tau, Dn=np.ones(nn) if Dn is None else Dn, f_ip=f_ip, dirichlet=dirichlet,
                         boundary=boundary)

    def assemble_linear(self):
        nN, m = self.s['nN'], self.m
        return self.assemble(np.zeros((self.nC, nN * m)), np.zeros(self.ndofs))
