"""HDF5 mesh and field files (HDF5Io, src/io/HDF5Io.cpp:13-390) on the CPU,
against the reference's own fixtures (tests/golden/h5, copied from
ressources/meshes) and what its tests check (tests/unittests/io/
TestHDF5Io.cpp:162-214: fieldTest.h5 holds NodeField = 0..8 and CellField
= (i, -i)); the meshes against the conversion made with the HDF5 command-line
dumper (tests/golden/meshes); files written here against h5dump's reading."""
import os
import shutil
import subprocess

import numpy as np
import pytest

import hyperfox_amd as hf

h5io = pytest.importorskip('hyperfox_amd.h5io')
if not os.path.exists(h5io.LIB_PATH):
    pytest.skip('libhfx_io.so not built (no HDF5 C library)', allow_module_level=True)

HERE = os.path.dirname(os.path.abspath(__file__))
H5 = os.path.join(HERE, 'golden', 'h5')
NPZ = os.path.join(HERE, 'golden', 'meshes')
H5DUMP = shutil.which('h5dump') or ('/opt/conda/bin/h5dump' if os.path.exists('/opt/conda/bin/h5dump') else None)


def test_reference_field_fixture():
    v, ft = h5io.read_field(os.path.join(H5, 'fieldTest.h5'), 'NodeField')
    assert v.shape == (9, 1, 1) and ft == h5io.FT_NODE
    assert np.array_equal(v.reshape(-1), np.arange(9.0))
    v, ft = h5io.read_field(os.path.join(H5, 'fieldTest.h5'), 'CellField')
    assert v.shape == (8, 2, 1) and ft == h5io.FT_CELL
    i = np.arange(8.0)
    assert np.array_equal(v.reshape(8, 2), np.stack([i, -i], 1))


@pytest.mark.parametrize('name', ['lightTri2', 'regression_dim-3_h-3e-1_ord-2'])
def test_reference_meshes_read_as_the_dumper_reads_them(name):
    nodes, cells = h5io.read_mesh(os.path.join(H5, name + '.h5'))
    z = np.load(os.path.join(NPZ, name + '.npz'))
    assert np.array_equal(nodes, z['nodes']) and np.array_equal(cells, z['cells'])


def test_hdf5io_load_builds_the_mesh_and_registered_fields():
    io = h5io.HDF5Io()
    io.load(os.path.join(H5, 'lightTri2.h5'))
    z = np.load(os.path.join(NPZ, 'lightTri2.npz'))
    a = io.mesh.arrays()
    assert io.mesh.dim == 2 and io.mesh.order == 2
    assert np.array_equal(a['cells'], z['cells']) and np.array_equal(a['coords'], z['nodes'])
    io2 = h5io.HDF5Io()
    io2.set_field('CellField')
    io2.set_field('NodeField')
    io2.load(os.path.join(H5, 'fieldTest.h5'))
    assert io2.mesh is None and io2.fields['NodeField'][1] == h5io.FT_NODE
    assert np.array_equal(io2.fields['NodeField'][0].reshape(-1), np.arange(9.0))


def test_write_then_read_round_trip(tmp_path):
    m = hf.Mesh.structured(3, 2, 2, 0.1, 5)
    a = m.arrays()
    rng = np.random.default_rng(0)
    sol = rng.normal(size=(m.n_cells, m.sizes['nN'], 1))
    flux = rng.normal(size=(m.n_cells, m.sizes['nN'], 3))
    trace = rng.normal(size=(m.n_faces, m.sizes['nNf'], 1))
    io = h5io.HDF5Io(m)
    io.set_field('Solution', sol, h5io.FT_CELL)
    io.set_field('Flux', flux, h5io.FT_CELL)
    io.set_field('Trace', trace, h5io.FT_FACE)
    path = str(tmp_path / 'out.h5')
    io.write(path)
    back = h5io.HDF5Io()
    for n in ('Solution', 'Flux', 'Trace'):
        back.set_field(n)
    back.load(path)
    b = back.mesh.arrays()
    for k in ('coords', 'cells', 'faces', 'face2cell', 'face_perm'):
        assert np.array_equal(a[k], b[k]), k
    for n, (v, ft) in (('Solution', (sol, 3)), ('Flux', (flux, 3)), ('Trace', (trace, 2))):
        assert np.array_equal(back.fields[n][0], v) and back.fields[n][1] == ft
    if H5DUMP:
        hdr = subprocess.run([H5DUMP, '-H', path], capture_output=True, text=True, check=True).stdout
        assert 'GROUP "Mesh"' in hdr and 'GROUP "FieldData"' in hdr
        assert 'DATASET "Nodes"' in hdr and '( %d, 3 )' % m.n_nodes in hdr
        assert 'DATASET "Cells"' in hdr and '( %d, %d )' % (m.n_cells, m.sizes['nN']) in hdr
        assert '( %d, %d, 3 )' % (m.n_cells, m.sizes['nN']) in hdr and 'ATTRIBUTE "ftype"' in hdr
        vals = subprocess.run([H5DUMP, '-d', '/FieldData/Trace', '-a', '/FieldData/Trace/ftype', path],
                              capture_output=True, text=True, check=True).stdout
        assert '(0): 2' in vals


def test_errors_follow_the_reference(tmp_path):
    bad = tmp_path / 'not.h5'
    bad.write_text('plain text')
    with pytest.raises(h5io.HDF5IoError, match='not an HDF5 file'):
        h5io.read_mesh(str(bad))
    with pytest.raises(h5io.HDF5IoError, match='not an HDF5 file'):
        h5io.HDF5Io().load(str(bad))
    with pytest.raises(h5io.HDF5IoError, match='field with name Missing was not found in FieldData'):
        h5io.read_field(os.path.join(H5, 'fieldTest.h5'), 'Missing')
    with pytest.raises(h5io.HDF5IoError, match='no Mesh group'):
        h5io.read_mesh(os.path.join(H5, 'fieldTest.h5'))
    io = h5io.HDF5Io()
    io.set_field('Empty')
    with pytest.raises(h5io.HDF5IoError, match='no values'):
        io.write(str(tmp_path / 'x.h5'))


@pytest.mark.gpu
def test_load_with_gpu_topology_solve_and_write_the_solution(tmp_path):
    path = os.path.join(H5, 'regression_dim-3_h-3e-1_ord-2.h5')
    io = h5io.HDF5Io().load(path, device=0)
    host = h5io.HDF5Io().load(path)
    for k in ('faces', 'cell2face', 'face2cell', 'boundary', 'face_perm'):
        assert np.array_equal(io.mesh.arrays()[k], host.mesh.arrays()[k]), k
    m = io.mesh
    nNf = m.sizes['nNf']
    prob = hf.HDGProblem(3, 2)
    prob.set_mesh(m)
    prob.set_model(1)
    prob.set_fields(np.ones((m.n_faces, nNf)), 1, dirichlet=np.zeros((m.n_faces, nNf)))
    prob.set_source(np.ones((m.n_cells, prob.sizes['nIP'])), None)
    prob.assemble()
    assert prob.solve(pc=hf.PC_BJACOBI, rtol=1e-12, maxits=20000)[1] == 2
    out = str(tmp_path / 'solution.h5')
    w = h5io.HDF5Io(m)
    for n, (v, ft) in h5io.problem_fields(prob).items():
        w.set_field(n, v, ft)
    w.write(out)
    r = h5io.HDF5Io()
    for n in ('Solution', 'Flux', 'Trace'):
        r.set_field(n)
    r.load(out)
    u, q, lam = prob.solution()
    assert np.array_equal(r.fields['Solution'][0].reshape(u.shape), u)
    assert np.array_equal(r.fields['Flux'][0].reshape(q.shape), q)
    assert np.array_equal(r.fields['Trace'][0].reshape(-1), lam)
    assert r.fields['Trace'][1] == h5io.FT_FACE and np.abs(u).max() > 0
