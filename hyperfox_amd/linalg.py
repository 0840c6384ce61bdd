"""Python mirror of the reference's linear-algebra plug point.

``HipLinAlgebra`` has the methods of ``LinAlgebraInterface``
(src/resolution/LinAlgebraInterface.h:23-162) with the behaviour of its PETSc
implementation (src/resolution/PetscInterface.cpp:5-265); ``HipOpts`` has the
fields of ``PetscOpts`` (src/resolution/PetscOpts.h:9-31).  Every call goes
through the hfx_la_* C ABI; the solve runs on the GPU.  Misuse (a call out of
order, mixing ADD and INSERT before a flush) raises ``HfxError`` where the
reference throws ``ErrorHandle``.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import check, lib

GMRES = 0
BCGS = 1
PC_NONE, PC_JACOBI, PC_ILU0, PC_ILU0_NATURAL = 0, 1, 4, 5


@dataclass
class HipOpts:
    """PetscOpts: solver type, preconditioner, rtol, maxits (+ GMRES restart)."""
    solverType: int = GMRES
    preconditionnerType: int = PC_JACOBI
    rtol: float = 1e-6
    maxits: int = 1000
    restart: int = 30
    verbose: bool = False


def _i32(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32).ravel())


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class HipLinAlgebra:
    """One sparse FP64 system A x = b; solve on GPU ``device``."""

    def __init__(self, options=None, device=0):
        self.options = options if options is not None else HipOpts()
        h = ctypes.c_void_p()
        check(lib.hfx_la_create(device, ctypes.byref(h)))
        self._h = h
        self.last_info = None

    def __del__(self):
        if getattr(self, '_h', None):
            lib.hfx_la_free(self._h)
            self._h = None

    # state machine -----------------------------------------------------
    def initialize(self):
        check(lib.hfx_la_initialize(self._h))

    def configure(self):
        o = self.options
        check(lib.hfx_la_configure(self._h, o.solverType, o.preconditionnerType, float(o.rtol), int(o.maxits),
                                   int(o.restart)))

    def allocate(self, ndofs, diagSparsePattern=None, offSparsePattern=None):
        d = None if diagSparsePattern is None else _i32(diagSparsePattern)
        o = None if offSparsePattern is None else _i32(offSparsePattern)
        check(lib.hfx_la_allocate(self._h, int(ndofs), None if d is None else _p(d), None if o is None else _p(o)))

    # values -----------------------------------------------------------------
    def _vals(self, fn, i_s, j_s, vals):
        i_s, j_s = _i32(i_s), _i32(j_s)
        v = _f64(vals)
        if v.size < i_s.size * j_s.size:
            raise ValueError('value block smaller than len(is) * len(js)')
        check(fn(self._h, i_s.size, _p(i_s), j_s.size, _p(j_s), _p(v)))

    def _rhs(self, fn, i_s, vals):
        i_s = _i32(i_s)
        v = _f64(vals)
        if v.size < i_s.size:
            raise ValueError('fewer values than indexes')
        check(fn(self._h, i_s.size, _p(i_s), _p(v)))

    def addValMatrix(self, i, j, val):
        self._vals(lib.hfx_la_add_vals, [i], [j], [val])

    def addValsMatrix(self, i_s, j_s, vals):
        self._vals(lib.hfx_la_add_vals, i_s, j_s, vals)

    def setValMatrix(self, i, j, val):
        self._vals(lib.hfx_la_set_vals, [i], [j], [val])

    def setValsMatrix(self, i_s, j_s, vals):
        self._vals(lib.hfx_la_set_vals, i_s, j_s, vals)

    def addValRHS(self, i, val):
        self._rhs(lib.hfx_la_add_rhs, [i], [val])

    def addValsRHS(self, i_s, vals):
        self._rhs(lib.hfx_la_add_rhs, i_s, vals)

    def setValRHS(self, i, val):
        self._rhs(lib.hfx_la_set_rhs, [i], [val])

    def setValsRHS(self, i_s, vals):
        self._rhs(lib.hfx_la_set_rhs, i_s, vals)

    def zeroOutRows(self, i_s):
        i_s = _i32(i_s)
        check(lib.hfx_la_zero_rows(self._h, i_s.size, _p(i_s)))

    # assembly / solve ---------------------------------------------------------
    def assemble(self):
        check(lib.hfx_la_assemble(self._h))

    def assembleFlush(self):
        check(lib.hfx_la_assemble_flush(self._h))

    def solve(self, solution=None):
        """Solve; fills (and returns) a list like the reference's
        std::vector<double>* argument.  ``last_info`` = (iters, reason, relres)."""
        n = self.getNumDofs()
        x = np.zeros(max(n, 1))
        info = np.zeros(2, dtype=np.int32)
        rel = np.zeros(1)
        check(lib.hfx_la_solve(self._h, _p(x), n, _p(info), _p(rel)))
        self.last_info = (int(info[0]), int(info[1]), float(rel[0]))
        if self.options.verbose:
            print('HipLinAlgebra solve (rtol = %g, maxIter = %d, nDOFs = %d): %d iterations, reason %d, '
                  'relres %.3e' % (self.options.rtol, self.options.maxits, n, *self.last_info))
        out = x[:n].tolist()
        if solution is not None:
            solution[:] = out
        return out

    def getSolutionOwnership(self, rng=None):
        lo, hi = ctypes.c_int(), ctypes.c_int()
        check(lib.hfx_la_ownership(self._h, ctypes.byref(lo), ctypes.byref(hi)))
        out = list(range(lo.value, hi.value))
        if rng is not None:
            rng[:] = out
        return out

    def clearSystem(self):
        check(lib.hfx_la_clear(self._h))

    def destroySystem(self):
        check(lib.hfx_la_destroy_system(self._h))

    def getNumDofs(self):
        n = ctypes.c_int()
        check(lib.hfx_la_num_dofs(self._h, ctypes.byref(n)))
        return n.value

    # reads of the assembled system (the tests' MatGetValues / VecGetValues) ----
    def getValues(self, i_s, j_s):
        i_s, j_s = _i32(i_s), _i32(j_s)
        out = np.zeros((i_s.size, j_s.size))
        check(lib.hfx_la_get_values(self._h, i_s.size, _p(i_s), j_s.size, _p(j_s), _p(out)))
        return out

    def getRHS(self, i_s):
        i_s = _i32(i_s)
        out = np.zeros(i_s.size)
        check(lib.hfx_la_get_rhs(self._h, i_s.size, _p(i_s), _p(out)))
        return out

    def nnz(self):
        v = ctypes.c_longlong()
        check(lib.hfx_la_nnz(self._h, ctypes.byref(v)))
        return v.value

    def csr(self):
        """(rowptr int64, cols int32, vals, rhs) of the assembled system."""
        n = self.getNumDofs()
        nnz = self.nnz()
        rowptr = np.zeros(n + 1, dtype=np.int64)
        cols = np.zeros(max(nnz, 1), dtype=np.int32)
        vals = np.zeros(max(nnz, 1))
        rhs = np.zeros(max(n, 1))
        check(lib.hfx_la_get_csr(self._h, _p(rowptr), _p(cols), _p(vals), _p(rhs)))
        return rowptr, cols[:nnz], vals[:nnz], rhs[:n]
