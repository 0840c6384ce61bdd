"""HDF5 mesh and field files in the reference's layout (HDF5Io,
src/io/HDF5Io.cpp:13-390) through libhfx_io.so (host only, HDF5 C library):
/Mesh/Nodes (nNodes x dim), /Mesh/Cells (nCells x nodes per cell) and
/FieldData/NAME (entities x objects x values, with an int "ftype"
attribute)."""
import ctypes
import os

import numpy as np

FT_NONE, FT_NODE, FT_EDGE, FT_FACE, FT_CELL = -1, 0, 1, 2, 3   # FieldType (src/field/FieldTypes.h)
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'libhfx_io.so')


class HDF5IoError(RuntimeError):
    pass


_lib = None


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HDF5IoError('libhfx_io.so is not built (it needs the HDF5 C library; run the build)')
        L = ctypes.CDLL(LIB_PATH)
        P, I, LL = ctypes.c_void_p, ctypes.c_int, ctypes.c_longlong
        sig = {
            'hfx_io_last_error': ([], ctypes.c_char_p),
            'hfx_h5_mesh_sizes': ([ctypes.c_char_p, P], I),
            'hfx_h5_read_mesh': ([ctypes.c_char_p, P, P], I),
            'hfx_h5_field_info': ([ctypes.c_char_p, ctypes.c_char_p, P, P], I),
            'hfx_h5_read_field': ([ctypes.c_char_p, ctypes.c_char_p, P], I),
            'hfx_h5_write': ([ctypes.c_char_p, I, I, P, I, I, P, I, P, P, P, P], I),
        }
        for n, (a, r) in sig.items():
            f = getattr(L, n)
            f.argtypes = a
            f.restype = r
        _lib = L
    return _lib


def _chk(st):
    if st != 0:
        raise HDF5IoError(_load().hfx_io_last_error().decode())


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def read_mesh(path):
    """(nodes nNodes x dim, cells nCells x nodes per cell) of /Mesh."""
    L = _load()
    s = np.zeros(4, dtype=np.int64)
    _chk(L.hfx_h5_mesh_sizes(os.fsencode(path), _p(s)))
    nodes = np.zeros((int(s[0]), int(s[1])))
    cells = np.zeros((int(s[2]), int(s[3])), dtype=np.int32)
    _chk(L.hfx_h5_read_mesh(os.fsencode(path), _p(nodes), _p(cells)))
    return nodes, cells


def read_field(path, name):
    """(values of shape (entities, objects, values), ftype) of /FieldData/name."""
    L = _load()
    shape = np.zeros(3, dtype=np.int64)
    ft = ctypes.c_int(0)
    _chk(L.hfx_h5_field_info(os.fsencode(path), name.encode(), _p(shape), ctypes.byref(ft)))
    vals = np.zeros(tuple(int(v) for v in shape))
    _chk(L.hfx_h5_read_field(os.fsencode(path), name.encode(), _p(vals)))
    return vals, ft.value


def write(path, nodes=None, cells=None, fields=None):
    """A new file with the mesh (if nodes and cells are given) and the fields,
    a dict name -> (values of shape (entities, objects, values), ftype)."""
    L = _load()
    nodes = None if nodes is None else np.ascontiguousarray(nodes, dtype=np.float64)
    cells = None if cells is None else np.ascontiguousarray(cells, dtype=np.int32)
    fields = fields or {}
    names = list(fields)
    vals = []
    shapes = np.zeros((len(names), 3), dtype=np.int64)
    ftypes = np.zeros(len(names), dtype=np.int32)
    for i, n in enumerate(names):
        v, ft = fields[n]
        v = np.ascontiguousarray(v, dtype=np.float64)
        if v.ndim != 3:
            raise HDF5IoError('HDF5Io: field %s must be (entities, objects, values)' % n)
        vals.append(v)
        shapes[i] = v.shape
        ftypes[i] = ft
    cnames = (ctypes.c_char_p * max(1, len(names)))(*[n.encode() for n in names])
    cvals = (ctypes.c_void_p * max(1, len(vals)))(*[v.ctypes.data for v in vals])
    _chk(L.hfx_h5_write(os.fsencode(path), 0 if nodes is None else nodes.shape[0],
                        0 if nodes is None else nodes.shape[1], _p(nodes), 0 if cells is None else cells.shape[0],
                        0 if cells is None else cells.shape[1], _p(cells), len(names), cnames, _p(shapes),
                        _p(ftypes), cvals))


class HDF5Io:
    """The reference's HDF5Io over a mesh: set_field registers a field (its
    values, for writing, or only its name, for loading); load reads the mesh
    when the file has one (topology on GPU `device` if given) and every
    registered field; write writes the mesh and the registered fields."""

    def __init__(self, mesh=None):
        self.mesh = mesh
        self.fields = {}

    def set_field(self, name, values=None, ftype=FT_CELL):
        self.fields[name] = (None if values is None else np.asarray(values, dtype=np.float64), ftype)

    def load(self, path, device=None):
        from .hdg import Mesh, refel_sizes
        L = _load()
        s = np.zeros(4, dtype=np.int64)
        if L.hfx_h5_mesh_sizes(os.fsencode(path), _p(s)) == 0:
            nodes, cells = read_mesh(path)
            dim = nodes.shape[1]
            order = next((q for q in range(1, 6) if refel_sizes(dim, q)['nN'] == cells.shape[1]), None)
            if order is None:
                raise HDF5IoError('HDF5Io: %d nodes per cell is no simplex of order 1-5' % cells.shape[1])
            self.mesh = Mesh.from_cells(dim, order, nodes, cells, device=device)
        else:
            err = L.hfx_io_last_error().decode()
            if 'not an HDF5 file' in err:
                raise HDF5IoError(err)
        for name in list(self.fields):
            self.fields[name] = read_field(path, name)
        return self

    def write(self, path):
        nodes = cells = None
        if self.mesh is not None:
            a = self.mesh.arrays()
            nodes, cells = a['coords'], a['cells']
        flds = {}
        for n, (v, ft) in self.fields.items():
            if v is None:
                raise HDF5IoError('HDF5Io: field %s has no values to write' % n)
            flds[n] = (v.reshape(v.shape[0], -1, 1) if v.ndim < 3 else v, ft)
        write(path, nodes, cells, flds)


def problem_fields(prob):
    """The solved fields of an HDGProblem in the reference's field shapes:
    Solution (Cell, nCells x nN x 1), Flux (Cell, nCells x nN x dim), Trace
    (Face, nFaces x nNf x 1)."""
    u, q, lam = prob.solution()
    nN, nNf = prob.sizes['nN'], prob.sizes['nNf']
    return {'Solution': (u.reshape(-1, nN, 1), FT_CELL),
            'Flux': (q.reshape(-1, nN, prob.dim), FT_CELL),
            'Trace': (lam.reshape(-1, nNf, 1), FT_FACE)}
