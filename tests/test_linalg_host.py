"""Host-side behaviour of the LinAlgebraInterface mirror (no GPU needed).

Follows the reference's own tests of the plug point:
* call ordering: tests/unittests/resolution/TestLinAlgebraInterfaces.cpp:23-68
  (every check there except the final solve, which needs a GPU);
* setVal / setVals / addVal / addVals read back through MatGetValues /
  VecGetValues: tests/unittests/resolution/TestPetscInterface.cpp:22-117;
and PETSc semantics the HDG assembly relies on (HDGSolver.cpp:531-675):
ADD sums duplicates in arrival order, INSERT keeps the last value, the two
cannot be mixed before a flush, zeroOutRows keeps the pattern.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

import hyperfox_amd as hf
import oracle as orc
from hyperfox_amd import HfxError, HipLinAlgebra, HipOpts

from conftest import ROOT


def ready(n, opts=None):
    la = HipLinAlgebra(opts or HipOpts(rtol=1e-16, verbose=False))
    la.initialize()
    la.configure()
    la.allocate(n)
    return la


def test_call_ordering():
    la = HipLinAlgebra(HipOpts(rtol=1e-16))
    dv, iv = [1.0], [0]

    def all_value_calls_throw():
        with pytest.raises(HfxError):
            la.solve(dv)
        with pytest.raises(HfxError):
            la.assemble()
        with pytest.raises(HfxError):
            la.setValsRHS(iv, dv)
        with pytest.raises(HfxError):
            la.setValRHS(0, 1.0)
        with pytest.raises(HfxError):
            la.setValsMatrix(iv, iv, dv)
        with pytest.raises(HfxError):
            la.setValMatrix(0, 0, 1.0)

    all_value_calls_throw()
    with pytest.raises(HfxError):
        la.allocate(1)
    with pytest.raises(HfxError, match='not initialized'):
        la.configure()
    la.initialize()
    all_value_calls_throw()
    with pytest.raises(HfxError):
        la.allocate(1)
    la.configure()
    all_value_calls_throw()
    la.allocate(1)
    la.setValsRHS(iv, dv)
    la.setValRHS(0, 1.0)
    la.setValsMatrix(iv, iv, dv)
    la.setValMatrix(0, 0, 1.0)
    la.assemble()
    assert la.getValues([0], [0])[0, 0] == 1.0
    la.clearSystem()
    assert la.getValues([0], [0])[0, 0] == 0.0


def test_clear_and_ownership_need_assembly():
    la = ready(3)
    with pytest.raises(HfxError, match='assembled at least once'):
        la.clearSystem()
    with pytest.raises(HfxError, match='ownership'):
        la.getSolutionOwnership()
    la.assemble()
    la.clearSystem()
    assert la.getSolutionOwnership() == []  # no solve yet
    la.destroySystem()
    with pytest.raises(HfxError):
        la.setValMatrix(0, 0, 1.0)
    assert la.getNumDofs() == 0


def test_set_val():
    la = ready(3)
    la.setValMatrix(1, 2, 1.0)
    la.setValRHS(1, 1.0)
    la.assemble()
    assert la.getValues([1], [2])[0, 0] == 1.0
    assert la.getRHS([1])[0] == 1.0


def test_set_vals():
    vals = np.diag([1.0, 2.0, 3.0, 4.0]).ravel()
    idx = [0, 1, 2, 3]
    la = ready(4)
    la.setValsMatrix(idx, idx, vals)
    la.setValsRHS(idx, vals)
    la.assemble()
    assert np.array_equal(la.getValues(idx, idx).ravel(), vals)
    assert np.array_equal(la.getRHS(idx), vals[:4])


def test_add_val():
    la = ready(3)
    la.addValMatrix(1, 2, 1.0)
    la.addValRHS(1, 1.0)
    la.assemble()
    assert la.getValues([1], [2])[0, 0] == 1.0
    assert la.getRHS([1])[0] == 1.0


def test_add_vals():
    vals = np.diag([1.0, 2.0, 3.0, 4.0]).ravel()
    skew = np.array([[0, 2, 0, 0], [-2, 0, 1, 0], [0, -1, 0, -3], [0, 0, 3, 0]], dtype=float).ravel()
    idx = [0, 1, 2, 3]
    la = ready(4)
    la.addValsMatrix(idx, idx, vals)
    la.addValsMatrix(idx, idx, skew)
    la.addValsRHS(idx, vals)
    la.assemble()
    assert np.array_equal(la.getValues(idx, idx).ravel(), vals + skew)
    assert np.array_equal(la.getRHS(idx), vals[:4])


def test_row_major_blocks():
    # MatSetValues' default orientation: vals[i * len(js) + j] goes to (is[i], js[j])
    la = ready(5)
    la.setValsMatrix([4, 1], [0, 3, 2], [1.0, 2.0, 3.0, 4.0, 5.0, 6.0])
    la.assemble()
    A = la.getValues(range(5), range(5))
    assert A[4, 0] == 1.0 and A[4, 3] == 2.0 and A[4, 2] == 3.0
    assert A[1, 0] == 4.0 and A[1, 3] == 5.0 and A[1, 2] == 6.0
    assert la.nnz() == 6


def test_add_insert_mixing_and_flush():
    la = ready(3)
    la.addValMatrix(0, 0, 1.0)
    with pytest.raises(HfxError, match='cannot be mixed'):
        la.setValMatrix(0, 0, 5.0)
    with pytest.raises(HfxError, match='cannot be mixed'):
        la.setValRHS(0, 5.0)
    la.assembleFlush()
    la.setValMatrix(0, 0, 5.0)
    la.setValMatrix(0, 0, 7.0)  # INSERT: last value wins
    la.assembleFlush()
    la.addValMatrix(0, 0, 0.5)
    la.addValMatrix(0, 0, 0.25)
    la.assemble()
    assert la.getValues([0], [0])[0, 0] == 7.75


def test_index_checks():
    la = ready(3)
    with pytest.raises(HfxError, match='row index'):
        la.addValMatrix(3, 0, 1.0)
    with pytest.raises(HfxError, match='column index'):
        la.addValMatrix(0, -1, 1.0)
    with pytest.raises(HfxError, match='row index'):
        la.addValRHS(7, 1.0)
    with pytest.raises(HfxError, match='together'):
        la2 = HipLinAlgebra()
        la2.initialize()
        la2.configure()
        lib = hf.lib
        import ctypes
        d = np.ones(3, dtype=np.int32)
        st = lib.hfx_la_allocate(la2._h, 3, d.ctypes.data_as(ctypes.c_void_p), None)
        hf._lib.check(st)
    with pytest.raises(HfxError, match='restart'):
        la3 = HipLinAlgebra(HipOpts(restart=0))
        la3.initialize()
        la3.configure()


def test_zero_rows_keeps_pattern_and_rhs():
    la = ready(3)
    la.addValsMatrix([0, 1, 2], [0, 1, 2], np.arange(1.0, 10.0))
    la.addValsRHS([0, 1, 2], [1.0, 2.0, 3.0])
    la.zeroOutRows([1])  # flushes pending ADD values first
    la.assemble()
    A = la.getValues(range(3), range(3))
    assert np.array_equal(A[1], np.zeros(3))
    assert np.array_equal(A[0], [1.0, 2.0, 3.0]) and np.array_equal(A[2], [7.0, 8.0, 9.0])
    assert la.nnz() == 9
    assert np.array_equal(la.getRHS([0, 1, 2]), [1.0, 2.0, 3.0])


@pytest.mark.parametrize('seed', [0, 1, 2])
def test_random_add_sequences_match_dense_accumulation(seed):
    rng = np.random.default_rng(seed)
    n = 40
    la = ready(n, HipOpts())
    dense = np.zeros((n, n))
    touched = np.zeros((n, n), dtype=bool)
    rhs = np.zeros(n)
    for _ in range(60):
        r = rng.choice(n, size=rng.integers(1, 6), replace=True)  # duplicates allowed
        c = rng.choice(n, size=rng.integers(1, 6), replace=True)
        v = rng.standard_normal((r.size, c.size))
        la.addValsMatrix(r, c, v)
        for i in range(r.size):
            for j in range(c.size):
                dense[r[i], c[j]] += v[i, j]
                touched[r[i], c[j]] = True
        w = rng.standard_normal(r.size)
        la.addValsRHS(r, w)
        for i in range(r.size):
            rhs[r[i]] += w[i]
        if rng.random() < 0.2:
            la.assembleFlush()
    la.assemble()
    rowptr, cols, vals, b = la.csr()
    assert la.nnz() == touched.sum()
    for i in range(n):
        cs = cols[rowptr[i]:rowptr[i + 1]]
        assert np.all(np.diff(cs) > 0)
        assert np.array_equal(cs, np.nonzero(touched[i])[0])
        assert np.allclose(vals[rowptr[i]:rowptr[i + 1]], dense[i, cs], rtol=0, atol=1e-13)
    assert np.allclose(b, rhs, rtol=0, atol=1e-13)


@pytest.mark.parametrize('dim,p,n', [(2, 2, 3), (3, 1, 2)])
def test_hdg_trace_assembly_through_the_interface_matches_oracle(dim, p, n):
    """The reference's scatter (HDGSolver.cpp:580-625: addValsMatrix of each
    cell's condensed block at its trace rows) through this interface gives the
    oracle's trace CSR bit for bit (same arrival order of the sums)."""
    m = hf.Mesh.structured(dim, p, n, jitter=0.1, seed=7)
    a = m.arrays()
    s = m.sizes
    nNf, nF, lenL = s['nNf'], s['nF'], s['lenL']
    tau = np.ones((m.n_faces, nNf))
    g = np.linspace(0.5, 1.5, m.n_faces * nNf).reshape(m.n_faces, nNf)
    sysm = orc.System(dim, p, 0, a, tau, dirichlet=g)
    cond = sysm.condensed()
    la = ready(m.n_trace, HipOpts())
    for e in range(m.n_cells):
        rows = np.array([a['cell2face'][e, f] * nNf + a['face_perm'][e, f, j] for f in range(nF) for j in range(nNf)])
        la.addValsMatrix(rows, rows, cond['S'][e])
        la.addValsRHS(rows, cond['S0'][e])
    la.assemble()
    rp, c, v, b = la.csr()
    rp0, c0, v0, b0 = sysm.csr()
    assert np.array_equal(rp, rp0) and np.array_equal(c, c0)
    assert np.array_equal(v, v0) and np.array_equal(b, b0)
    assert lenL == nF * nNf


def test_cpp_adapter_compiles_against_the_reference_interface(tmp_path):
    """HipLinAlgebra.h implements the reference's abstract class; compiled and
    driven through its call-ordering checks (no solve: no GPU here), which
    throw the reference's ErrorHandle as PetscInterface's do, with the
    class and member function in its message (ErrorHandle.cpp: "class :
    function : message")."""
    iface = '/root/reference/src/resolution/LinAlgebraInterface.h'
    errh = '/root/reference/src/globals/ErrorHandle.cpp'
    if not os.path.exists(iface) or not os.path.exists(errh):
        pytest.skip('reference interface headers not present')
    if shutil.which('g++') is None:
        pytest.skip('no g++')
    src = tmp_path / 'drive.cpp'
    src.write_text(r'''
#include <cstdio>
#include <string>
#include <vector>
#include "HipLinAlgebra.h"
using namespace hfox;
template <class F> static int throws(F f) { try { f(); } catch (const ErrorHandle&) { return 1; } return 0; }
int main() {
  HipOpts o; o.rtol = 1e-16;
  LinAlgebraInterface* la = new HipLinAlgebra(o);
  std::vector<double> dv = {1}; std::vector<int> iv = {0};
  int bad = 0;
  bad += !throws([&] { la->assemble(); });
  try { la->assemble(); } catch (const ErrorHandle& e) {
    std::printf("%s\n", e.what());
    bad += std::string(e.what()).rfind("HipLinAlgebra : assemble : ", 0) != 0;
  }
  bad += !throws([&] { la->setValMatrix(0, 0, 1.0); });
  bad += !throws([&] { la->configure(); });
  la->initialize();
  bad += !throws([&] { la->allocate(1); });
  la->configure();
  bad += !throws([&] { la->solve(&dv); });
  la->allocate(2);
  la->addValsMatrix(iv, iv, dv.data());
  bad += !throws([&] { la->setValMatrix(1, 1, 2.0); });
  la->assembleFlush();
  la->setValMatrix(1, 1, 2.0);
  la->setValsRHS(iv, dv.data());
  la->assemble();
  bad += la->getNumDofs() != 2;
  la->clearSystem();
  la->destroySystem();
  delete la;
  std::printf("bad=%d\n", bad);
  return bad;
}
''')
    exe = tmp_path / 'drive'
    libdir = os.path.join(ROOT, 'hyperfox_amd')
    cmd = ['g++', '-std=c++17', '-O1', str(src), errh, '-I', os.path.join(ROOT, 'include'), '-I',
           os.path.dirname(iface), '-I', os.path.dirname(errh), '-L', libdir, '-lhfx_hip', '-Wl,-rpath,' + libdir,
           '-o', str(exe)]
    subprocess.run(cmd, check=True, capture_output=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert 'bad=0' in r.stdout
