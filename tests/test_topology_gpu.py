"""Mesh topology built on the GPU (faces, cell2face, face2cell, boundary,
face-node permutations; the reference's MOAB computeFaces,
src/mesh/Mesh.cpp:183-537, conventions at :377-537) against the host
builder: the arrays must be identical, on structured meshes, on the
reference's own mesh fixtures (tests/golden/meshes, converted from
ressources/meshes) and on shuffled, renumbered inputs."""
import glob
import os
import time

import numpy as np
import pytest

import hyperfox_amd as hf

pytestmark = pytest.mark.gpu

KEYS = ('faces', 'cell2face', 'face2cell', 'boundary', 'face_perm')


def same(m0, m1):
    a0, a1 = m0.arrays(), m1.arrays()
    for k in KEYS:
        assert np.array_equal(a0[k], a1[k]), k
    assert (m0.n_faces, m0.n_boundary) == (m1.n_faces, m1.n_boundary)


@pytest.mark.parametrize('dim,p,n', [(2, 1, 1), (2, 1, 7), (2, 3, 9), (2, 5, 4), (3, 1, 1), (3, 1, 5), (3, 2, 4),
                                     (3, 3, 6)])
def test_structured_meshes_match_the_host_build(dim, p, n):
    same(hf.Mesh.structured(dim, p, n, 0.1, 3), hf.Mesh.structured(dim, p, n, 0.1, 3, device=0))


def fixtures():
    here = os.path.join(os.path.dirname(__file__), 'golden', 'meshes')
    return sorted(glob.glob(os.path.join(here, '*.npz')))


@pytest.mark.parametrize('path', fixtures(), ids=lambda s: os.path.basename(s))
def test_reference_mesh_fixtures_match_the_host_build(path):
    z = np.load(path)
    nodes, cells = z['nodes'], z['cells']
    dim = nodes.shape[1]
    nn = cells.shape[1]
    p = next(q for q in range(1, 6) if hf.refel_sizes(dim, q)['nN'] == nn)
    same(hf.Mesh.from_cells(dim, p, nodes, cells), hf.Mesh.from_cells(dim, p, nodes, cells, device=0))


@pytest.mark.parametrize('dim,p', [(2, 2), (3, 3)])
def test_shuffled_renumbered_input_matches_the_host_build(dim, p):
    m = hf.Mesh.structured(dim, p, 5, 0.1, 7)
    a = m.arrays()
    rng = np.random.default_rng(1)
    order = rng.permutation(m.n_cells)
    relabel = rng.permutation(m.n_nodes)
    nodes = np.empty_like(a['coords'])
    nodes[relabel] = a['coords']
    cells = relabel[a['cells'][order]]
    same(hf.Mesh.from_cells(dim, p, nodes, cells), hf.Mesh.from_cells(dim, p, nodes, cells, device=0))


def test_a_face_of_three_cells_is_rejected():
    m = hf.Mesh.structured(2, 1, 2, 0.0, 0)
    a = m.arrays()
    cells = np.concatenate([a['cells'], a['cells'][:1]])   # a duplicated cell: its faces have three cells
    with pytest.raises(hf.HfxError, match='more than two cells'):
        hf.Mesh.from_cells(2, 1, a['coords'], cells)
    with pytest.raises(hf.HfxError, match='more than two cells'):
        hf.Mesh.from_cells(2, 1, a['coords'], cells, device=0)


def test_benchmark_mesh_and_timing():
    m = hf.Mesh.structured(3, 3, 24, 0.05, 4321)
    a = m.arrays()
    t0 = time.perf_counter()
    mh = hf.Mesh.from_cells(3, 3, a['coords'], a['cells'])
    t1 = time.perf_counter()
    mg = hf.Mesh.from_cells(3, 3, a['coords'], a['cells'], device=0)
    t2 = time.perf_counter()
    same(mh, mg)
    print('24^3 x 6 p=3 topology: host %.1f ms, GPU %.1f ms' % ((t1 - t0) * 1e3, (t2 - t1) * 1e3))
