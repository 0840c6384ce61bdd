"""The multi-DOF path on the GPU: HDGBurgersModel
(src/model/HDGBurgersModel.cpp, m = dim unknowns per node), HDGnGammaModel
(src/model/HDGnGammaModel.cpp, 2-D, m = 2) and the scalar models at m = dim
DOFs per node (MultiProblem) on HDGSolver's IMPLICIT assembly,
and NonLinearWrapper's loop (src/solver/NonLinearWrapper.cpp), through the
hfx_burgers_* / hfx_ngamma_* entry points of libhfx_hip.so.

Field shapes follow the reference's Field values: tau (nFaces, nNf, sides,
m, m) with tau[F, p, s][a, b] = tau_ab of the face's node p seen from the
cell face2cell[F, s]; D (nNodes,) or None; dirichlet (nFaces, nNf, m);
source values at the cell IPs (nCells, nIP, m); Solution and
BufferSolution (nCells, nN, m); Flux (nCells, nN, dim, m); Trace
(nFaces, nNf, m)."""
import ctypes

import numpy as np

from ._lib import check, lib
from .hdg import _f64, _ptr, _sized

BC_DIRICHLET, BC_INTEGRATED_DIRICHLET = 0, 1


class BurgersProblem:
    """HDGSolver + HDGBurgersModel resident on one GPU."""

    def __init__(self, dim, order, device=0):
        h = ctypes.c_void_p()
        check(lib.hfx_burgers_create(int(device), int(dim), int(order), ctypes.byref(h)))
        self._h = h
        self.dim, self.order, self.m = dim, order, dim
        self.mesh = None

    def __del__(self):
        if getattr(self, '_h', None):
            lib.hfx_burgers_destroy(self._h)
            self._h = None

    def _check_tau(self, t):
        if self.mesh is None:
            raise ValueError('set the mesh first')
        if t.shape[:2] != (self.mesh.n_faces, self.nNf):
            raise ValueError('tau must be (nFaces, nNf, sides, m, m) = (%d, %d, ...), got %s'
                             % (self.mesh.n_faces, self.nNf, t.shape))

    def set_mesh(self, mesh):
        check(lib.hfx_burgers_set_mesh(self._h, mesh._h))
        self.mesh = mesh
        s = mesh.sizes
        self.nN, self.nNf = s['nN'], s['nNf']
        self.nIP = s['nIP']
        out = np.zeros(5, dtype=np.int64)
        check(lib.hfx_burgers_sizes(self._h, _ptr(out)))
        self.nb, self.bs, self.nnzb, self.lu, self.ll = (int(v) for v in out)

    def set_fields(self, tau, D=None, dirichlet=None, source=None):
        """tau (nFaces, nNf, sides, m, m); the other fields as the module says."""
        t = np.asarray(tau, dtype=np.float64)
        if t.ndim != 5 or t.shape[3:] != (self.m, self.m) or t.shape[2] not in (1, 2):
            raise ValueError('tau must be (nFaces, nNf, sides, m, m) with sides 1 or 2')
        self._check_tau(t)
        tcol = _f64(np.swapaxes(t, 3, 4))  # column-major m x m per node
        m = self.mesh
        Dd = _sized(D, m.n_nodes, 'DiffusionTensor')
        g = _sized(dirichlet, m.n_faces * self.nNf * self.m, 'Dirichlet')
        f = _sized(source, m.n_cells * self.nIP * self.m, 'source')
        self._keep = (tcol, Dd, g, f)
        check(lib.hfx_burgers_set_fields(self._h, _ptr(tcol), t.shape[2], _ptr(Dd), _ptr(g), _ptr(f)))

    def set_boundary_model(self, kind):
        check(lib.hfx_burgers_set_boundary_model(self._h, int(kind)))

    def set_boundary_hdg(self, faces, B, r, matrix_set=True, rhs_set=True):
        """HDG-type boundary model on `faces`: B (n, bs, bs (2 + dim)), r (n, bs)."""
        f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1)
        Bd = _sized(B, f.size * self.bs * self.bs * (2 + self.dim), 'set_boundary_hdg B')
        rd = _sized(r, f.size * self.bs, 'set_boundary_hdg r')
        self._keep_hdg = (f, Bd, rd)
        self._n_hdg = f.size
        check(lib.hfx_burgers_set_boundary_hdg(self._h, _ptr(f), int(f.size), _ptr(Bd), _ptr(rd), int(bool(matrix_set)),
                                               int(bool(rhs_set))))

    def boundary_rows(self):
        """(B, r) of the HDG-type boundary faces at the last assembly
        (caller-set, or built by the Bohm model): (n, bs, bs (2 + dim)), (n, bs)."""
        n = int(getattr(self, '_n_hdg', 0))
        B = np.zeros((n, self.bs, self.bs * (2 + self.dim)))
        r = np.zeros((n, self.bs))
        check(lib.hfx_burgers_boundary_rows(self._h, _ptr(B), _ptr(r)))
        return B, r

    def set_time_scheme(self, scheme, dt=0.0):
        """0: steady; 1 (TS_BEULER): implicit Euler with step dt."""
        check(lib.hfx_burgers_set_time_scheme(self._h, int(scheme), float(dt)))

    def start_step(self):
        """OldSolution = BufferSolution = Solution."""
        check(lib.hfx_burgers_start_step(self._h))

    def set_dirichlet(self, dirichlet):
        g = _sized(dirichlet, self.mesh.n_faces * self.nNf * self.m, 'Dirichlet')
        self._keep_g = g
        check(lib.hfx_burgers_set_dirichlet(self._h, _ptr(g)))

    def set_solver_graphs(self, on):
        check(lib.hfx_burgers_set_solver_graphs(self._h, int(bool(on))))

    def set_krylov(self, krylov):
        check(lib.hfx_burgers_set_krylov(self._h, int(krylov)))

    def set_state(self, sol=None, buffer=None, trace=None):
        nC, nF = self.mesh.n_cells, self.mesh.n_faces
        a = [_sized(sol, nC * self.nN * self.m, 'Solution'), _sized(buffer, nC * self.nN * self.m, 'BufferSolution'),
             _sized(trace, nF * self.nNf * self.m, 'Trace')]
        check(lib.hfx_burgers_set_state(self._h, *[_ptr(v) for v in a]))

    def state(self):
        """(Solution, BufferSolution, Flux, Trace) copied to the host."""
        nC, nF = self.mesh.n_cells, self.mesh.n_faces
        sol = np.zeros((nC, self.nN, self.m))
        buf = np.zeros_like(sol)
        flux = np.zeros((nC, self.nN, self.dim, self.m))
        tr = np.zeros((nF, self.nNf, self.m))
        check(lib.hfx_burgers_get_state(self._h, _ptr(sol), _ptr(buf), _ptr(flux), _ptr(tr)))
        return sol, buf, flux, tr

    def assemble(self):
        check(lib.hfx_burgers_assemble(self._h))

    def condensed(self):
        """(S, S0, U, Q, U0, Q0) of every cell, the trace in cell-local order."""
        nC, lu, ll = self.mesh.n_cells, self.lu, self.ll
        S, S0 = np.zeros((nC, ll, ll)), np.zeros((nC, ll))
        U, Q = np.zeros((nC, lu, ll)), np.zeros((nC, lu * self.dim, ll))
        U0, Q0 = np.zeros((nC, lu)), np.zeros((nC, lu * self.dim))
        check(lib.hfx_burgers_condensed(self._h, *[_ptr(v) for v in (S, S0, U, Q, U0, Q0)]))
        return S, S0, U, Q, U0, Q0

    def trace_system(self):
        """(rowptr, col, vals (nnzb, bs, bs), rhs) of the assembled trace system."""
        rowptr = np.zeros(self.nb + 1, dtype=np.int32)
        col = np.zeros(self.nnzb, dtype=np.int32)
        vals = np.zeros((self.nnzb, self.bs, self.bs))
        rhs = np.zeros(self.nb * self.bs)
        check(lib.hfx_burgers_trace_system(self._h, _ptr(rowptr), _ptr(col), _ptr(vals), _ptr(rhs)))
        return rowptr, col, vals, rhs

    def solve(self, pc=2, rtol=1e-12, maxits=10000, restart=30):
        """Trace solve + reconstruction; returns (iterations, reason, true relative residual)."""
        it, reason, rr = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        check(lib.hfx_burgers_solve(self._h, int(pc), float(rtol), int(maxits), int(restart), ctypes.byref(it),
                                    ctypes.byref(reason), ctypes.byref(rr)))
        return it.value, reason.value, rr.value

    def nonlinear_solve(self, max_iters=20, res_tol=1e-6, damp=0.0, pc=2, rtol=1e-12, maxits=10000, restart=30):
        """NonLinearWrapper::solve; returns (iterations, residual, total linear iterations)."""
        it, res, lin = ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
        check(lib.hfx_burgers_nonlinear_solve(self._h, int(max_iters), float(res_tol), float(damp), int(pc),
                                              float(rtol), int(maxits), int(restart), ctypes.byref(it),
                                              ctypes.byref(res), ctypes.byref(lin)))
        return it.value, res.value, lin.value


class MultiProblem(BurgersProblem):
    """A scalar HDG model at m = dim DOFs per node on one GPU
    (HDGOperator::multiplyDOFs, src/operator/HDGOperator.cpp:17-28, as
    tests/regression/HDG/TestHDGLaplaceMulti.cpp runs HDGLaplaceModel):
    set_fields(tau, D=None, dirichlet, source) with D None for
    HDGLaplaceModel (D = I) or a scalar nodal D; tau (nFaces, nNf, sides, m,
    m) couples the components on the faces.  Linear: assemble(), solve(),
    state() give the solution."""

    def __init__(self, dim, order, device=0):
        h = ctypes.c_void_p()
        check(lib.hfx_multi_create(int(device), int(dim), int(order), ctypes.byref(h)))
        self._h = h
        self.dim, self.order, self.m = dim, order, dim
        self.mesh = None


class NGammaProblem(BurgersProblem):
    """HDGSolver + HDGnGammaModel (2-D; unknowns n and Gamma) on one GPU; the
    state fields hold (n, Gamma) per node."""

    def __init__(self, order, device=0):
        h = ctypes.c_void_p()
        check(lib.hfx_ngamma_create(int(device), int(order), ctypes.byref(h)))
        self._h = h
        self.dim, self.order, self.m = 2, order, 2
        self.mesh = None

    def set_fields(self, tau, D=None, G=None, b=None, sound_speed=1.0, dirichlet=None, source=None):
        """tau (nFaces, nNf, sides, 2, 2); D, G (nNodes,) or None; b (nNodes, 2)."""
        t = np.asarray(tau, dtype=np.float64)
        if t.ndim != 5 or t.shape[3:] != (2, 2) or t.shape[2] not in (1, 2):
            raise ValueError('tau must be (nFaces, nNf, sides, 2, 2) with sides 1 or 2')
        self._check_tau(t)
        tcol = _f64(np.swapaxes(t, 3, 4))
        m = self.mesh
        arrs = [_sized(D, m.n_nodes, 'D'), _sized(G, m.n_nodes, 'G'), _sized(b, m.n_nodes * 2, 'Velocity'),
                _sized(dirichlet, m.n_faces * self.nNf * 2, 'Dirichlet'), _sized(source, m.n_cells * self.nIP * 2, 'source')]
        self._keep = (tcol, *arrs)
        Dd, Gd, bd, g, f = arrs
        check(lib.hfx_ngamma_set_fields(self._h, _ptr(tcol), t.shape[2], _ptr(Dd), _ptr(Gd), _ptr(bd),
                                        float(sound_speed), _ptr(g), _ptr(f)))

    def set_boundary_bohm(self, faces, normals, transfer=2):
        """HDGBohmModel on the boundary `faces` with the exterior normals
        (n, nNf, 2) in face-node order; transfer 0 neumann, 1 dirichlet,
        2 heaviside, 3 tanh (the reference regression's functions)."""
        f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1)
        nr = _sized(normals, f.size * self.nNf * 2, 'set_boundary_bohm normals')
        self._keep_bohm = (f, nr)
        check(lib.hfx_ngamma_set_boundary_bohm(self._h, _ptr(f), int(f.size), _ptr(nr), int(transfer)))
        self._n_hdg = f.size
