"""Python handles over the C ABI of libhfx_hip.so (include/hfx_hip.h).

These mirror the reference's set-up sequence (Mesh -> fields -> Solver
initialize/allocate/assemble/solve, src/solver/HDGSolver.cpp:5-779) for the
tests and the benchmark.  All compute runs in the HIP library.
"""
import ctypes

import numpy as np

from ._lib import HfxError, check, lib

# Butcher tables of the reference's RungeKutta database, in RKType order
# (src/operator/RKType.h); RK_CUSTOM selects a caller's table
RK_NAMES = ['FEuler', 'EMidpoint', 'Heun', 'Kutta3', 'Heun3', 'SSPRK3', 'RK4', 'BEuler', 'IMidpoint',
            'CrankNicolson', 'KS2', 'QZ2', 'ALX2', 'RK43']
RK_CUSTOM = -1


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.int32)


def _sized(a, n, name):
    """a as a contiguous FP64 array of exactly n values (None stays None);
    the C side reads n values, so any other size is refused here."""
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=np.float64)
    if a.size != n:
        raise ValueError('%s: expected %d values, got %d (shape %s)' % (name, n, a.size, a.shape))
    return a


def refel_sizes(dim, order):
    out = np.zeros(7, dtype=np.int32)
    check(lib.hfx_refel_sizes(dim, order, _ptr(out)))
    keys = ('nN', 'nF', 'nNf', 'nIP', 'nIPf', 'n_local', 'lenL')
    return dict(zip(keys, (int(v) for v in out)))


def refel_tables(dim, order):
    s = refel_sizes(dim, order)
    nN, nF, nNf, nIP, nIPf = s['nN'], s['nF'], s['nNf'], s['nIP'], s['nIPf']
    t = dict(nodes=np.zeros((nN, dim)), face_nodes=np.zeros((nF, nNf), dtype=np.int32),
             ipx=np.zeros((nIP, dim)), ipw=np.zeros(nIP), phi=np.zeros((nIP, nN)),
             dphi=np.zeros((nIP, nN, dim)), fipx=np.zeros((nIPf, dim - 1)), fipw=np.zeros(nIPf),
             fphi=np.zeros((nIPf, nNf)))
    check(lib.hfx_refel_tables(dim, order, *[_ptr(t[k]) for k in
                                             ('nodes', 'face_nodes', 'ipx', 'ipw', 'phi', 'dphi',
                                              'fipx', 'fipw', 'fphi')]))
    return t


def cubature(dim, order):
    n = lib.hfx_cubature(dim, order, None, None)
    if n < 0:
        check(-1)
    pts = np.zeros((n, dim))
    w = np.zeros(n)
    lib.hfx_cubature(dim, order, _ptr(pts), _ptr(w))
    return pts, w


class Mesh:
    """Simplex mesh with the reference's topology conventions."""

    def __init__(self, handle, dim, order):
        self._h = handle
        self.dim = dim
        self.order = order
        out = np.zeros(5, dtype=np.int64)
        check(lib.hfx_mesh_sizes(self._h, _ptr(out)))
        self.n_nodes, self.n_cells, self.n_faces, self.n_boundary, self.n_trace = (int(v) for v in out)
        self.sizes = refel_sizes(dim, order)
        self._arrays = None

    @classmethod
    def structured(cls, dim, order, n, jitter=0.0, seed=0, device=None):
        """device: build the topology on that GPU (same arrays as the host build)."""
        h = ctypes.c_void_p()
        if device is None:
            check(lib.hfx_mesh_structured(dim, order, n, float(jitter), ctypes.c_ulonglong(seed), ctypes.byref(h)))
        else:
            check(lib.hfx_mesh_structured_gpu(int(device), dim, order, n, float(jitter), ctypes.c_ulonglong(seed),
                                              ctypes.byref(h)))
        return cls(h, dim, order)

    @classmethod
    def from_cells(cls, dim, order, nodes, cells, device=None):
        nodes = _f64(nodes)
        cells = _i32(cells)
        h = ctypes.c_void_p()
        if device is None:
            check(lib.hfx_mesh_from_cells(dim, order, nodes.shape[0], _ptr(nodes), cells.shape[0], _ptr(cells),
                                          ctypes.byref(h)))
        else:
            check(lib.hfx_mesh_from_cells_gpu(int(device), dim, order, nodes.shape[0], _ptr(nodes), cells.shape[0],
                                              _ptr(cells), ctypes.byref(h)))
        return cls(h, dim, order)

    def arrays(self):
        if self._arrays is None:
            s = self.sizes
            a = dict(coords=np.zeros((self.n_nodes, self.dim)),
                     cells=np.zeros((self.n_cells, s['nN']), dtype=np.int32),
                     faces=np.zeros((self.n_faces, s['nNf']), dtype=np.int32),
                     cell2face=np.zeros((self.n_cells, s['nF']), dtype=np.int32),
                     face2cell=np.zeros((self.n_faces, 2), dtype=np.int32),
                     boundary=np.zeros(self.n_boundary, dtype=np.int32),
                     face_perm=np.zeros((self.n_cells, s['nF'], s['nNf']), dtype=np.int8))
            check(lib.hfx_mesh_get(self._h, *[_ptr(a[k]) for k in
                                             ('coords', 'cells', 'faces', 'cell2face', 'face2cell',
                                              'boundary', 'face_perm')]))
            self._arrays = a
        return self._arrays

    def __del__(self):
        if getattr(self, '_h', None):
            lib.hfx_mesh_free(self._h)
            self._h = None


class HDGProblem:
    """One HDG problem resident on a GPU."""

    def __init__(self, dim, order, device=0):
        self.dim = dim
        self.order = order
        self.sizes = refel_sizes(dim, order)
        h = ctypes.c_void_p()
        check(lib.hfx_hdg_create(device, dim, order, ctypes.byref(h)))
        self._h = h
        self.mesh = None

    def __del__(self):
        if getattr(self, '_h', None):
            lib.hfx_hdg_destroy(self._h)
            self._h = None

    def set_mesh(self, mesh):
        check(lib.hfx_hdg_set_mesh(self._h, mesh._h))
        self.mesh = mesh

    def _need_mesh(self, what):
        if self.mesh is None:
            raise ValueError('%s: set the mesh first' % what)
        return self.mesh

    def set_model(self, kind):
        check(lib.hfx_hdg_set_model(self._h, kind))

    def set_fields(self, tau, tau_sides=1, D=None, ndD=0, b=None, dirichlet=None):
        m, nNf = self._need_mesh('set_fields'), self.sizes['nNf']
        if tau is None:
            raise ValueError('set_fields: the field map must have a Tau field')
        self._keep = [_sized(tau, m.n_faces * nNf * tau_sides, 'Tau'),
                      _sized(D, m.n_nodes * ndD, 'DiffusionTensor') if D is not None and ndD > 0 else _f64(D),
                      _sized(b, m.n_nodes * self.dim, 'Velocity'),
                      _sized(dirichlet, m.n_faces * nNf, 'Dirichlet')]
        t, d, bb, g = self._keep
        check(lib.hfx_hdg_set_fields(self._h, _ptr(t), tau_sides, _ptr(d), ndD if d is not None else 0,
                                     _ptr(bb), _ptr(g)))

    def ip_coords(self):
        out = np.zeros((self.mesh.n_cells, self.sizes['nIP'], self.dim))
        check(lib.hfx_hdg_ip_coords(self._h, _ptr(out)))
        return out

    def set_source(self, f_ip=None, r_ip=None):
        n = self._need_mesh('set_source').n_cells * self.sizes['nIP']
        f = _sized(f_ip, n, 'source')
        r = _sized(r_ip, n, 'reaction')
        check(lib.hfx_hdg_set_source(self._h, _ptr(f), _ptr(r)))

    def set_launch(self, threads=0, grid_cap=0):
        check(lib.hfx_hdg_set_launch(self._h, threads, grid_cap))

    def set_kernel(self, which):
        """0 automatic, 1 general (per-IP Jacobians), 2 affine-cell kernel."""
        check(lib.hfx_hdg_set_kernel(self._h, which))

    def kernel_info(self):
        out = np.zeros(4, dtype=np.int32)
        check(lib.hfx_hdg_kernel_info(self._h, _ptr(out)))
        return dict(kernel='affine' if out[0] == 2 else 'general', threads=int(out[1]), mesh_affine=bool(out[2]),
                    lds_bytes=int(out[3]))

    def assemble(self):
        check(lib.hfx_hdg_assemble(self._h))

    def local_matrices(self, first, count):
        n = self.sizes['n_local']
        K = np.zeros((count, n, n))
        F = np.zeros((count, n))
        check(lib.hfx_hdg_local_matrices(self._h, first, count, _ptr(K), _ptr(F)))
        return K, F

    def condensed(self):
        s, ne, d = self.sizes, self.mesh.n_cells, self.dim
        nN, nL = s['nN'], s['lenL']
        out = dict(S=np.zeros((ne, nL, nL)), S0=np.zeros((ne, nL)), U=np.zeros((ne, nN, nL)),
                   Q=np.zeros((ne, nN * d, nL)), U0=np.zeros((ne, nN)), Q0=np.zeros((ne, nN * d)))
        check(lib.hfx_hdg_get_condensed(self._h, *[_ptr(out[k]) for k in ('S', 'S0', 'U', 'Q', 'U0', 'Q0')]))
        return out

    def trace_sizes(self):
        out = np.zeros(3, dtype=np.int64)
        check(lib.hfx_hdg_trace_sizes(self._h, _ptr(out)))
        return tuple(int(v) for v in out)

    def trace_csr(self):
        n, nnz, _ = self.trace_sizes()
        rowptr = np.zeros(n + 1, dtype=np.int64)
        cols = np.zeros(nnz, dtype=np.int32)
        vals = np.zeros(nnz)
        rhs = np.zeros(n)
        check(lib.hfx_hdg_get_trace_csr(self._h, _ptr(rowptr), _ptr(cols), _ptr(vals), _ptr(rhs)))
        return rowptr, cols, vals, rhs

    def set_time_scheme(self, scheme, dt=0.0):
        """0 none, 1 implicit Euler, 2 explicit Euler (the reference's Euler,
        isExplicit false / true; u_old is the current solution field), 3
        Runge-Kutta (set_butcher, rk_*), on the u rows with step dt."""
        check(lib.hfx_hdg_set_time_scheme(self._h, scheme, dt))

    def set_butcher(self, table):
        """Butcher table: a name of RK_NAMES (the reference's database) or a
        square array [[c | A], [0 | b]] with zeros above the diagonal."""
        if isinstance(table, str):
            check(lib.hfx_hdg_set_butcher(self._h, RK_NAMES.index(table), None, 0))
        else:
            t = _f64(table)
            if t.ndim != 2 or t.shape[0] != t.shape[1]:
                raise HfxError('RungeKutta: the Butcher table should be square')
            check(lib.hfx_hdg_set_butcher(self._h, RK_CUSTOM, _ptr(t), t.shape[0]))

    def rk_start_step(self):
        """Old fields = current u, q, lambda; stage fields zero."""
        check(lib.hfx_hdg_rk_start_step(self._h))

    def rk_compute_stage(self):
        check(lib.hfx_hdg_rk_compute_stage(self._h))

    def rk_compute_solution(self):
        check(lib.hfx_hdg_rk_compute_solution(self._h))

    def rk_stages(self):
        """(number of stages, stage counter)."""
        out = np.zeros(2, dtype=np.int32)
        check(lib.hfx_hdg_rk_stages(self._h, _ptr(out)))
        return int(out[0]), int(out[1])

    def rk_step(self, pc=2, rtol=1e-10, maxits=10000, restart=30):
        """One Runge-Kutta time step as the reference's drivers take it: per
        stage assemble, solve and computeStage, then computeSolution.
        Returns the solve info of every stage."""
        self.rk_start_step()
        infos = []
        for _ in range(self.rk_stages()[0]):
            self.assemble()
            infos.append(self.solve(pc, rtol, maxits, restart))
            self.rk_compute_stage()
        self.rk_compute_solution()
        return infos

    def set_boundary_model(self, kind, faces=None):
        """Boundary model of the listed boundary faces (None: all): 0
        DirichletModel (default), 1 IntegratedDirichletModel."""
        if faces is None:
            check(lib.hfx_hdg_set_boundary_model(self._h, kind, None, 0))
        else:
            f = np.ascontiguousarray(faces, dtype=np.int32)
            check(lib.hfx_hdg_set_boundary_model(self._h, kind, _ptr(f), f.size))

    def set_boundary_hdg(self, faces, B, r, matrix_set=True, rhs_set=True):
        """HDG-type boundary model on the listed boundary faces: B (n, nNf,
        nNf (2 + dim)) = [B_u | B_q | B_lambda] in face-node order, r (n,
        nNf), Set or Add assembly; an empty list clears it."""
        f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1)
        nNf = self.sizes['nNf']
        B = _sized(B, f.size * nNf * nNf * (2 + self.dim), 'set_boundary_hdg B')
        r = _sized(r, f.size * nNf, 'set_boundary_hdg r')
        check(lib.hfx_hdg_set_boundary_hdg(self._h, _ptr(f), f.size, _ptr(B), _ptr(r), int(bool(matrix_set)),
                                           int(bool(rhs_set))))

    def set_trace(self, lam):
        """Set the trace field (nFaces x nNf)."""
        lam = _sized(lam, self._need_mesh('set_trace').n_trace, 'Trace')
        check(lib.hfx_hdg_set_trace(self._h, _ptr(lam)))

    def set_solution(self, u):
        """Set the cell-wise nodal solution field (nCells x nN)."""
        u = _sized(u, self._need_mesh('set_solution').n_cells * self.sizes['nN'], 'Solution')
        check(lib.hfx_hdg_set_solution(self._h, _ptr(u)))

    def set_solver_type(self, kind):
        """HDGSolverOpts::type: 0 IMPLICIT, 1 WEXPLICIT, 2 SEXPLICIT."""
        check(lib.hfx_hdg_set_solver_type(self._h, kind))

    def set_flux(self, q):
        """Set the cell-wise nodal flux field (nCells x nN*dim)."""
        q = _sized(q, self._need_mesh('set_flux').n_cells * self.sizes['nN'] * self.dim, 'Flux')
        check(lib.hfx_hdg_set_flux(self._h, _ptr(q)))

    def set_fused_scatter(self, on):
        """Fused trace scatter on straight-sided cells (default on): the
        element kernel writes the trace matrix in place instead of storing S."""
        check(lib.hfx_hdg_set_fused_scatter(self._h, 1 if on else 0))

    def set_solver_graphs(self, on):
        """BiCGStab steady iterations as one captured HIP graph (default off:
        no faster on MI355X)."""
        check(lib.hfx_hdg_set_solver_graphs(self._h, 1 if on else 0))

    def set_krylov(self, method):
        """0 restarted FGMRES (default), 1 BiCGStab."""
        check(lib.hfx_hdg_set_krylov(self._h, method))

    def solve(self, pc=2, rtol=1e-10, maxits=10000, restart=30):
        info = np.zeros(2, dtype=np.int32)
        rel = np.zeros(1)
        check(lib.hfx_hdg_solve(self._h, pc, rtol, maxits, restart, _ptr(info), _ptr(rel)))
        return int(info[0]), int(info[1]), float(rel[0])

    def solve_noalloc(self, pc=2, rtol=1e-10, maxits=10000):
        """The BiCGStab trace solve of solve() with no device allocation
        (hfx_hdg_solve_noalloc: a fixed per-GPU workspace; pc 0, 1 or 2)."""
        info = np.zeros(2, dtype=np.int32)
        rel = np.zeros(1)
        check(lib.hfx_hdg_solve_noalloc(self._h, pc, rtol, maxits, _ptr(info), _ptr(rel)))
        return int(info[0]), int(info[1]), float(rel[0])

    def pc_apply(self, pc, x):
        """y = M^-1 x for preconditioner `pc` set up on the assembled system."""
        x = _sized(x, self._need_mesh('pc_apply').n_trace, 'pc_apply x')
        y = np.zeros_like(x)
        check(lib.hfx_hdg_pc_apply(self._h, pc, _ptr(x), _ptr(y)))
        return y

    def ilu_factors(self, ordering=0):
        """Block ILU(0) factors (ordering 0 multicolour, 1 natural): lu
        (nnzb, b, b) in trace-CSR block order, dinv (nFaces, b, b), and
        (colours, forward levels, backward levels)."""
        n, nnz, _ = self.trace_sizes()
        b = self.sizes['nNf']
        lu = np.zeros((nnz // (b * b), b, b))
        dinv = np.zeros((n // b, b, b))
        info = np.zeros(3, dtype=np.int32)
        check(lib.hfx_hdg_ilu_factors(self._h, ordering, _ptr(lu), _ptr(dinv), _ptr(info)))
        return lu, dinv, tuple(int(v) for v in info)

    def solution(self):
        s, ne, d = self.sizes, self.mesh.n_cells, self.dim
        u = np.zeros((ne, s['nN']))
        q = np.zeros((ne, s['nN'] * d))
        lam = np.zeros(self.mesh.n_trace)
        check(lib.hfx_hdg_get_solution(self._h, _ptr(u), _ptr(q), _ptr(lam)))
        return u, q, lam

    def time(self, what, warmup=1, reps=3):
        ms = np.zeros(reps)
        check(lib.hfx_hdg_time(self._h, what, warmup, reps, _ptr(ms)))
        return ms


def rk_table(name):
    """The library's Butcher table of the database entry `name` (n x n)."""
    t = np.zeros(25)
    n = np.zeros(1, dtype=np.int32)
    check(lib.hfx_rk_table(RK_NAMES.index(name), _ptr(t), _ptr(n)))
    return t[:n[0] * n[0]].reshape(n[0], n[0])


def ilu_analyse(rowptr, col, ordering):
    """Host analysis of a block pattern for ILU(0) (no GPU): dict of rank,
    levf, levb, colour arrays and the counts (colours, forward levels,
    backward levels, update pairs)."""
    rowptr, col = _i32(rowptr), _i32(col)
    nb = rowptr.size - 1
    out = {k: np.zeros(nb, dtype=np.int32) for k in ('rank', 'levf', 'levb', 'colour')}
    info = np.zeros(4, dtype=np.int64)
    check(lib.hfx_ilu_analyse(nb, _ptr(rowptr), _ptr(col), ordering, _ptr(out['rank']), _ptr(out['levf']),
                              _ptr(out['levb']), _ptr(out['colour']), _ptr(info)))
    out.update(ncolours=int(info[0]), nlev_f=int(info[1]), nlev_b=int(info[2]), npairs=int(info[3]))
    return out


class Batch:
    """Many 2-D problems assembled by one persistent launch (hfx_batch_*)."""

    def __init__(self, device=0):
        h = ctypes.c_void_p()
        check(lib.hfx_batch_create(device, ctypes.byref(h)))
        self._h = h
        self.problems = []

    def __del__(self):
        if getattr(self, '_h', None):
            lib.hfx_batch_free(self._h)
            self._h = None

    def add(self, prob):
        check(lib.hfx_batch_add(self._h, prob._h))
        self.problems.append(prob)  # the batch holds raw handles: keep them alive

    def size(self):
        n, c = ctypes.c_int(), ctypes.c_longlong()
        check(lib.hfx_batch_size(self._h, ctypes.byref(n), ctypes.byref(c)))
        return n.value, c.value

    def set_config(self, waves, tile, order=0):
        """Launch shape of one order (1..5) or of all (0)."""
        check(lib.hfx_batch_set_config(self._h, order, waves, tile))

    def config(self):
        """{order: dict(waves, tile, items, workgroups)} for the orders present."""
        out = np.zeros(21, dtype=np.int32)
        check(lib.hfx_batch_get_config(self._h, _ptr(out)))
        self.concurrent = bool(out[20] & 1)
        self.graph_concurrent = bool(out[20] & 2)
        return {p: dict(waves=int(out[4 * (p - 1)]), tile=int(out[4 * (p - 1) + 1]), items=int(out[4 * (p - 1) + 2]),
                        workgroups=int(out[4 * (p - 1) + 3])) for p in range(1, 6) if out[4 * (p - 1) + 2] > 0}

    def tune(self, reps=3):
        v = ctypes.c_double()
        check(lib.hfx_batch_tune(self._h, reps, ctypes.byref(v)))
        return v.value

    def assemble(self, use_graph=False):
        check(lib.hfx_batch_assemble(self._h, 1 if use_graph else 0))

    def time(self, what, warmup=1, reps=3):
        ms = np.zeros(reps)
        check(lib.hfx_batch_time(self._h, what, warmup, reps, _ptr(ms)))
        return ms
