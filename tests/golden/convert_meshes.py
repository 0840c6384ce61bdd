"""Convert the reference's small HDF5 mesh fixtures to .npz (data only).

The reference stores meshes as /Mesh/Nodes [nNodes x dim] f8 and
/Mesh/Cells [nCells x nNodesPerCell] i4 (src/io/HDF5Io.cpp:111-190).  This
script reads them with the HDF5 command-line dumper (no Python HDF5 binding
is installed here) and writes nodes/cells arrays to tests/golden/meshes/.

Usage: python convert_meshes.py <reference ressources/meshes dir> [h5dump path]
"""
import os
import subprocess
import sys
import tempfile

import numpy as np

FILES = [
    'lightTri2.h5',
    'regression/regression_dim-2_h-3e-1_ord-1.h5',
    'regression/regression_dim-2_h-3e-1_ord-2.h5',
    'regression/regression_dim-2_h-3e-1_ord-3.h5',
    'regression/regression_dim-2_h-2e-1_ord-1.h5',
    'regression/regression_dim-2_h-2e-1_ord-2.h5',
    'regression/regression_dim-2_h-2e-1_ord-3.h5',
    'regression/regression_dim-2_h-1e-1_ord-1.h5',
    'regression/regression_dim-2_h-1e-1_ord-2.h5',
    'regression/regression_dim-2_h-1e-1_ord-3.h5',
    'regression/regression_dim-3_h-3e-1_ord-1.h5',
    'regression/regression_dim-3_h-3e-1_ord-2.h5',
    'regression/regression_dim-3_h-3e-1_ord-3.h5',
    'regression/regression_dim-3_h-2e-1_ord-1.h5',
    'regression/regression_dim-3_h-2e-1_ord-2.h5',
    'regression/regression_dim-3_h-2e-1_ord-3.h5',
]


def dump(h5dump, path, dset, dtype):
    with tempfile.TemporaryDirectory() as td:
        out = os.path.join(td, 'o.txt')
        # %.17g: every double printed with enough digits to read back exactly
        fmt = ['-m', '%.17g'] if dtype == np.float64 else []
        subprocess.run([h5dump, '-d', dset, '-y', '-w', '0'] + fmt + ['-o', out, path], check=True,
                       stdout=subprocess.DEVNULL)
        hdr = subprocess.run([h5dump, '-H', '-d', dset, path], check=True,
                             capture_output=True, text=True).stdout
        dims = hdr.split('SIMPLE { (')[1].split(')')[0]
        shape = tuple(int(v) for v in dims.split(','))
        txt = open(out).read().replace(',', ' ').split()
        return np.array(txt, dtype=dtype).reshape(shape)


def main():
    root = sys.argv[1]
    h5dump = sys.argv[2] if len(sys.argv) > 2 else 'h5dump'
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'meshes')
    os.makedirs(dst, exist_ok=True)
    for f in FILES:
        nodes = dump(h5dump, os.path.join(root, f), '/Mesh/Nodes', np.float64)
        cells = dump(h5dump, os.path.join(root, f), '/Mesh/Cells', np.int32)
        name = os.path.basename(f).replace('.h5', '.npz')
        np.savez_compressed(os.path.join(dst, name), nodes=nodes, cells=cells)
        print(name, nodes.shape, cells.shape)


if __name__ == '__main__':
    main()
