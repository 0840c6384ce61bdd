"""ctypes binding of the FP64 CPU oracle (liboracle_hdg.so).

TEST INFRASTRUCTURE ONLY: loaded by tests/, __graft_entry__.smoke() and the
benchmark's CPU-baseline leg, never by the product (hyperfox_amd).  See
hdg_oracle.h for what it restates from the reference and where.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'liboracle_hdg.so')


class OracleError(RuntimeError):
    pass


class Model(ctypes.Structure):
    _fields_ = [('kind', ctypes.c_int), ('ndD', ctypes.c_int)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError('oracle not built: run make -C oracle')
    L = ctypes.CDLL(LIB_PATH)
    P, I, D, LL = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_longlong
    sig = {
        'orc_last_error': ([], ctypes.c_char_p),
        'orc_sizes': ([I, I, P], I),
        'orc_refel': ([I, I] + [P] * 9, I),
        'orc_cubature': ([I, I, P, P], I),
        'orc_local_system': ([I, I, Model] + [P] * 8, I),
        'orc_ip_coords': ([I, I, P, P], I),
        'orc_condense': ([I, I] + [P] * 8, I),
        'orc_apply_euler': ([I, I, D, P, P, P], I),
        'orc_explicit_blocks': ([I, I] + [P] * 6, I),
        'orc_system_set_euler': ([P, D, P, I], I),
        'orc_system_set_hdg_type': ([P, I, P, P, I], I),
        'orc_apply_euler_explicit': ([I, I, D, P, P, P], I),
        'orc_apply_rk': ([I, I, D, I, P, I, P, P, P, P], I),
        'orc_system_set_time_scheme': ([P, I, D, I, P, I, P, P, P, I], I),
        'orc_system_local_vectors': ([P] * 5, I),
        'orc_system_set_boundary_model': ([P, I, P, I, I], I),
        'orc_face_mass': ([I, I, P, P], I),
        'orc_system_set_boundary_hdg': ([P, P, I, P, P, I, I, I], I),
        'orc_system_build': ([I, I, Model, I, P, I, P, I, P, P, P, P, P, P, I, P, P, P, I], P),
        'orc_system_free': ([P], None),
        'orc_system_sizes': ([P, P], I),
        'orc_system_csr': ([P] * 5, I),
        'orc_system_condensed': ([P] * 7, I),
        'orc_system_solve': ([P, I, D, I, I, P, P, P], I),
        'orc_system_reconstruct': ([P, P, P, P], I),
        'orc_system_time_assembly': ([P, I, I], D),
        'orc_dense_gmres': ([I, P, P, D, I, I, P, P], I),
    }
    for n, (a, r) in sig.items():
        f = getattr(L, n)
        f.argtypes = a
        f.restype = r
    return L


lib = _load()


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _chk(st):
    if st != 0:
        raise OracleError(lib.orc_last_error().decode())


def sizes(dim, p):
    out = np.zeros(7, dtype=np.int32)
    _chk(lib.orc_sizes(dim, p, _p(out)))
    return dict(zip(('nN', 'nF', 'nNf', 'nIP', 'nIPf', 'n_local', 'lenL'), (int(v) for v in out)))


def refel(dim, p):
    s = sizes(dim, p)
    nN, nF, nNf, nIP, nIPf = s['nN'], s['nF'], s['nNf'], s['nIP'], s['nIPf']
    t = dict(nodes=np.zeros((nN, dim)), face_nodes=np.zeros((nF, nNf), dtype=np.int32),
             ipx=np.zeros((nIP, dim)), ipw=np.zeros(nIP), phi=np.zeros((nIP, nN)),
             dphi=np.zeros((nIP, nN, dim)), fipx=np.zeros((nIPf, dim - 1)), fipw=np.zeros(nIPf),
             fphi=np.zeros((nIPf, nNf)))
    _chk(lib.orc_refel(dim, p, *[_p(t[k]) for k in ('nodes', 'face_nodes', 'ipx', 'ipw', 'phi', 'dphi',
                                                      'fipx', 'fipw', 'fphi')]))
    return t


def cubature(dim, order):
    n = lib.orc_cubature(dim, order, None, None)
    if n < 0:
        raise OracleError(lib.orc_last_error().decode())
    pts, w = np.zeros((n, dim)), np.zeros(n)
    lib.orc_cubature(dim, order, _p(pts), _p(w))
    return pts, w


def local_system(dim, p, kind, x, D=None, ndD=0, b=None, tau=None, f_ip=None, r_ip=None):
    s = sizes(dim, p)
    n = s['n_local']
    K, F = np.zeros((n, n)), np.zeros(n)
    arrs = [np.ascontiguousarray(v, dtype=np.float64) if v is not None else None for v in (x, D, b, tau, f_ip, r_ip)]
    _chk(lib.orc_local_system(dim, p, Model(kind, ndD if D is not None else 0), *[_p(a) for a in arrs], _p(K), _p(F)))
    return K, F


def apply_euler(dim, p, dt, u_old, K, F):
    """Implicit Euler on the u rows of a local system (Euler::apply,
    Euler.cpp:18-37); returns new (K, F)."""
    K2, F2 = np.array(K, dtype=np.float64, order='C'), np.array(F, dtype=np.float64)
    _chk(lib.orc_apply_euler(dim, p, dt, _p(np.ascontiguousarray(u_old, dtype=np.float64)), _p(K2), _p(F2)))
    return K2, F2


def apply_euler_explicit(dim, p, dt, u_old, K, F):
    """Explicit Euler on the u rows of a local system (Euler::apply with
    isExplicit = true, Euler.cpp:30-35); returns new (K, F)."""
    K2, F2 = np.array(K, dtype=np.float64, order='C'), np.array(F, dtype=np.float64)
    _chk(lib.orc_apply_euler_explicit(dim, p, dt, _p(np.ascontiguousarray(u_old, dtype=np.float64)), _p(K2), _p(F2)))
    return K2, F2


def face_mass(dim, p, xf):
    """Face mass of IntegratedDirichletModel (Mass on the face element with
    the face's own nodes xf, nNf x dim, IntegratedDirichletModel.cpp:35-44)."""
    n = sizes(dim, p)['nNf']
    M = np.zeros((n, n))
    _chk(lib.orc_face_mass(dim, p, _p(np.ascontiguousarray(xf, dtype=np.float64)), _p(M)))
    return M


def apply_rk(dim, p, dt, a_row, stage, ut, kst, K, F):
    """One Runge-Kutta stage on the u rows of a local system
    (RungeKutta::apply, RungeKutta.cpp:113-159, auxiliary fields Flux and
    Trace): a_row = Butcher row `stage` without its first column, ut = the
    cell-local old vector [u | q | lambda], kst = (nst, n) stage vectors;
    returns new (K, F)."""
    a = np.ascontiguousarray(a_row, dtype=np.float64)
    K2, F2 = np.array(K, dtype=np.float64, order='C'), np.array(F, dtype=np.float64)
    ut = np.ascontiguousarray(ut, dtype=np.float64)
    kst = np.ascontiguousarray(kst, dtype=np.float64)
    _chk(lib.orc_apply_rk(dim, p, dt, a.size, _p(a), stage, _p(ut), _p(kst), _p(K2), _p(F2)))
    return K2, F2


# Butcher tables of RungeKutta::setUpDB (RungeKutta.cpp:197-285), in RKType
# order (RKType.h): row i < nst is [c_i | a_i,0 .. a_i,nst-1], the last row
# [0 | b_0 .. b_nst-1].
_G = 1.0 - np.sqrt(2.0) / 2.0
RK_TABLES = {
    'FEuler': [[0, 0], [0, 1]],
    'EMidpoint': [[0, 0, 0], [0.5, 0.5, 0], [0, 0, 1]],
    'Heun': [[0, 0, 0], [1, 1, 0], [0, 0.5, 0.5]],
    'Kutta3': [[0, 0, 0, 0], [0.5, 0.5, 0, 0], [1, -1, 2, 0], [0, 1 / 6, 2 / 3, 1 / 6]],
    'Heun3': [[0, 0, 0, 0], [1 / 3, 1 / 3, 0, 0], [2 / 3, 0, 2 / 3, 0], [0, 1 / 4, 0, 3 / 4]],
    'SSPRK3': [[0, 0, 0, 0], [1, 1, 0, 0], [1 / 2, 1 / 4, 1 / 4, 0], [0, 1 / 6, 1 / 6, 2 / 3]],
    'RK4': [[0, 0, 0, 0, 0], [1 / 2, 1 / 2, 0, 0, 0], [1 / 2, 0, 1 / 2, 0, 0], [1, 0, 0, 1, 0],
            [0, 1 / 6, 1 / 3, 1 / 3, 1 / 6]],
    'BEuler': [[1, 1], [0, 1]],
    'IMidpoint': [[1 / 2, 1 / 2], [0, 1]],
    'CrankNicolson': [[0, 0, 0], [1, 1 / 2, 1 / 2], [0, 1 / 2, 1 / 2]],
    'KS2': [[1 / 2, 1 / 2, 0], [3 / 2, -1 / 2, 2], [0, -1 / 2, 3 / 2]],
    'QZ2': [[1 / 4, 1 / 4, 0], [3 / 4, 1 / 2, 1 / 4], [0, 1 / 2, 1 / 2]],
    'ALX2': [[_G, _G, 0], [1, 1 - _G, _G], [0, 1 - _G, _G]],
    'RK43': [[1 / 2, 1 / 2, 0, 0, 0], [2 / 3, 1 / 6, 1 / 2, 0, 0], [1 / 2, -1 / 2, 1 / 2, 1 / 2, 0],
             [1, 3 / 2, -3 / 2, 1 / 2, 1 / 2], [0, 3 / 2, -3 / 2, 1 / 2, 1 / 2]],
}
RK_NAMES = list(RK_TABLES)


class RungeKutta:
    """Stage bookkeeping of RungeKutta (RungeKutta.cpp:161-253) over whole-mesh
    fields: fields = {'Solution': u, 'Flux': q, 'Trace': lambda} (numpy
    arrays updated in place), with Old* copies and per-stage RKStage arrays."""

    def __init__(self, table, dt):
        if isinstance(table, str):  # setButcherTable(RKType): a table of the database
            t = np.asarray(RK_TABLES[table], dtype=np.float64)
        else:  # setButcherTable(EMatrix) and its checks (RungeKutta.cpp:17-31)
            t = np.asarray(table, dtype=np.float64)
            if t.ndim != 2 or t.shape[0] != t.shape[1]:
                raise OracleError('RungeKutta: the Butcher table should be square')
            if np.any(t[np.triu_indices(t.shape[0], 1)] != 0):
                raise OracleError('RungeKutta: the upper triangular part of the Butcher table should be null')
        self.table, self.dt = t, float(dt)
        self.nst = t.shape[0] - 1
        self.stage = 0

    def a_row(self, s):
        return self.table[s, 1:]

    def start_step(self, fields):
        self.old = {k: np.array(v, dtype=np.float64) for k, v in fields.items()}
        self.k = {k: np.zeros((self.nst,) + np.shape(v)) for k, v in fields.items()}
        self.stage = 0

    def compute_stage(self, fields):
        if self.stage >= self.nst:
            raise OracleError('RungeKutta: cannot compute more stages than the method allows')
        a = self.a_row(self.stage)
        for name, cur in fields.items():
            self.k[name][self.stage] = (cur - self.old[name]) / self.dt
            buf = np.zeros_like(cur)
            for j in range(self.stage + 1):
                buf += a[j] * self.k[name][j]
            cur[...] = self.old[name] + self.dt * buf
        self.stage += 1

    def compute_solution(self, fields):
        if self.stage != self.nst:
            raise OracleError('RungeKutta: all stages must be computed before computing the solution')
        b = self.table[self.nst, 1:]
        for name, cur in fields.items():
            cur[...] = self.old[name]
            for j in range(self.nst):
                cur += (b[j] * self.dt) * self.k[name][j]
        self.stage = 0


def explicit_blocks(dim, p, K, F, sol, flux):
    """Explicit HDG modes on one local system: (S, S0) with S = Sll and
    S0 = F_lambda - Slu sol - Slq flux (HDGSolver.cpp:346-353)."""
    s = sizes(dim, p)
    nL = s['lenL']
    S, S0 = np.zeros((nL, nL)), np.zeros(nL)
    arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (K, F, sol, flux)]
    _chk(lib.orc_explicit_blocks(dim, p, *[_p(x) for x in arrs], _p(S), _p(S0)))
    return S, S0


def ip_coords(dim, p, x):
    s = sizes(dim, p)
    out = np.zeros((s['nIP'], dim))
    _chk(lib.orc_ip_coords(dim, p, _p(np.ascontiguousarray(x, dtype=np.float64)), _p(out)))
    return out


def condense(dim, p, K, F):
    s = sizes(dim, p)
    nN, nL = s['nN'], s['lenL']
    out = dict(S=np.zeros((nL, nL)), S0=np.zeros(nL), U=np.zeros((nN, nL)), Q=np.zeros((nN * dim, nL)),
               U0=np.zeros(nN), Q0=np.zeros(nN * dim))
    _chk(lib.orc_condense(dim, p, _p(np.ascontiguousarray(K)), _p(np.ascontiguousarray(F)),
                          *[_p(out[k]) for k in ('S', 'S0', 'U', 'Q', 'U0', 'Q0')]))
    return out


class System:
    """Whole-mesh oracle pipeline (HDGSolver.cpp:176-779 restated)."""

    def __init__(self, dim, p, kind, mesh_arrays, tau, tau_sides=1, D=None, ndD=0, b=None, f_ip=None, r_ip=None,
                 dirichlet=None, nthreads=0):
        m = mesh_arrays
        self.dim, self.p = dim, p
        self.s = sizes(dim, p)
        self.n_cells = m['cells'].shape[0]
        self._keep = [np.ascontiguousarray(v, dtype=np.float64) if v is not None else None
                      for v in (m['coords'], D, b, tau, f_ip, r_ip, dirichlet)]
        self._ikeep = [np.ascontiguousarray(m[k], dtype=np.int32) for k in ('cells', 'faces', 'cell2face', 'face2cell')]
        co, Dd, bb, tt, ff, rr, gg = self._keep
        ce, fa, c2f, f2c = self._ikeep
        h = lib.orc_system_build(dim, p, Model(kind, ndD if D is not None else 0), co.shape[0], _p(co), ce.shape[0],
                                 _p(ce), fa.shape[0], _p(fa), _p(c2f), _p(f2c), _p(Dd), _p(bb), _p(tt), tau_sides,
                                 _p(ff), _p(rr), _p(gg), nthreads)
        if not h:
            raise OracleError(lib.orc_last_error().decode())
        self._h = h
        out = np.zeros(2, dtype=np.int64)
        lib.orc_system_sizes(self._h, _p(out))
        self.ndofs, self.nnz = int(out[0]), int(out[1])

    def __del__(self):
        if getattr(self, '_h', None):
            lib.orc_system_free(self._h)
            self._h = None

    def set_euler(self, dt, u_old=None, nthreads=0):
        """Implicit Euler with step dt (0: none) and solution field u_old
        (nCells x nN) on every element, then reassemble."""
        self._uold = None if u_old is None else np.ascontiguousarray(u_old, dtype=np.float64)
        _chk(lib.orc_system_set_euler(self._h, dt, _p(self._uold), nthreads))

    def set_hdg_type(self, kind, sol=None, flux=None, nthreads=0):
        """HDGSolverType 0 IMPLICIT, 1 WEXPLICIT, 2 SEXPLICIT (explicit modes
        use the solution field sol (nCells x nN) and flux (nCells x nN*dim)),
        then reassemble."""
        self._sol = None if sol is None else np.ascontiguousarray(sol, dtype=np.float64)
        self._flux = None if flux is None else np.ascontiguousarray(flux, dtype=np.float64)
        _chk(lib.orc_system_set_hdg_type(self._h, kind, _p(self._sol), _p(self._flux), nthreads))

    def set_time_scheme(self, kind, dt=0.0, a_row=None, stage=0, u_old=None, old_local=None, stage_local=None,
                        nthreads=0):
        """Time scheme on every element, then reassemble: kind 0 none,
        1 implicit Euler, 2 explicit Euler (u_old), 3 Runge-Kutta stage
        `stage` with Butcher row a_row, old_local (nCells x n) and
        stage_local (nst x nCells x n) cell-local vectors."""
        self._ts = [None if v is None else np.ascontiguousarray(v, dtype=np.float64)
                    for v in (a_row, u_old, old_local, stage_local)]
        a, uo, ol, sl = self._ts
        nst = 0 if a is None else a.size
        _chk(lib.orc_system_set_time_scheme(self._h, kind, dt, nst, _p(a), stage, _p(uo), _p(ol), _p(sl), nthreads))

    def set_boundary_model(self, kind, faces=None, nthreads=0):
        """Boundary model of the listed boundary faces (None: all), then
        reassemble: 0 DirichletModel, 1 IntegratedDirichletModel."""
        self._bf = None if faces is None else np.ascontiguousarray(faces, dtype=np.int32)
        n = 0 if self._bf is None else self._bf.size
        _chk(lib.orc_system_set_boundary_model(self._h, kind, _p(self._bf), n, nthreads))

    def set_boundary_hdg(self, faces, B, r, matrix_set=True, rhs_set=True, nthreads=0):
        """HDG-type boundary model on the listed boundary faces: B (n, nNf,
        nNf (2 + dim)) = [B_u | B_q | B_lambda] in face-node order, r (n, nNf),
        Set or Add assembly; then reassemble."""
        self._hf = np.ascontiguousarray(faces, dtype=np.int32)
        self._hb = np.ascontiguousarray(B, dtype=np.float64)
        self._hr = np.ascontiguousarray(r, dtype=np.float64)
        _chk(lib.orc_system_set_boundary_hdg(self._h, _p(self._hf), self._hf.size, _p(self._hb), _p(self._hr),
                                             int(matrix_set), int(rhs_set), nthreads))

    def local_vectors(self, u, q, lam):
        """Cell-local [u | q | lambda] vectors (nCells x n) of whole-mesh fields."""
        out = np.zeros((self.n_cells, self.s['n_local']))
        arrs = [np.ascontiguousarray(v, dtype=np.float64) for v in (u, q, lam)]
        _chk(lib.orc_system_local_vectors(self._h, *[_p(x) for x in arrs], _p(out)))
        return out

    def csr(self):
        rowptr = np.zeros(self.ndofs + 1, dtype=np.int64)
        cols = np.zeros(self.nnz, dtype=np.int32)
        vals = np.zeros(self.nnz)
        rhs = np.zeros(self.ndofs)
        lib.orc_system_csr(self._h, _p(rowptr), _p(cols), _p(vals), _p(rhs))
        return rowptr, cols, vals, rhs

    def condensed(self):
        s, ne, d = self.s, self.n_cells, self.dim
        nN, nL = s['nN'], s['lenL']
        out = dict(S=np.zeros((ne, nL, nL)), S0=np.zeros((ne, nL)), U=np.zeros((ne, nN, nL)),
                   Q=np.zeros((ne, nN * d, nL)), U0=np.zeros((ne, nN)), Q0=np.zeros((ne, nN * d)))
        lib.orc_system_condensed(self._h, *[_p(out[k]) for k in ('S', 'S0', 'U', 'Q', 'U0', 'Q0')])
        return out

    def solve(self, method=0, rtol=1e-12, maxits=20000, restart=30):
        lam = np.zeros(self.ndofs)
        info = np.zeros(2, dtype=np.int32)
        rel = np.zeros(1)
        _chk(lib.orc_system_solve(self._h, method, rtol, maxits, restart, _p(lam), _p(info), _p(rel)))
        return lam, int(info[0]), bool(info[1]), float(rel[0])

    def reconstruct(self, lam):
        s, ne, d = self.s, self.n_cells, self.dim
        u = np.zeros((ne, s['nN']))
        q = np.zeros((ne, s['nN'] * d))
        _chk(lib.orc_system_reconstruct(self._h, _p(np.ascontiguousarray(lam)), _p(u), _p(q)))
        return u, q

    def time_assembly(self, n_sample, nthreads=0):
        t = lib.orc_system_time_assembly(self._h, int(n_sample), int(nthreads))
        if t < 0:
            raise OracleError(lib.orc_last_error().decode())
        return t


def dense_gmres(A, b, rtol=1e-12, maxits=1000, restart=30):
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(b.shape[0])
    it = np.zeros(1, dtype=np.int32)
    _chk(lib.orc_dense_gmres(b.shape[0], _p(A), _p(b), rtol, maxits, restart, _p(x), _p(it)))
    return x, int(it[0])
