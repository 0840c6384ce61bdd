"""The bulk HDG path driven from plain C (tests/c_caller/hdg_caller.c, built
with gcc against include/hfx_hip.h and linked to libhfx_hip.so): the call
sequence a HipHDGSolver adapter makes in place of HDGSolver's element loop
(src/solver/HDGSolver.cpp:5-779) - structured mesh, problem, CDRS model,
Tau / DiffusionTensor / Velocity / Dirichlet fields, source and reaction at
the GPU's IP coordinates, assemble, BiCGStab + block Jacobi to 1e-13,
solution.  The caller writes the inputs it chose and the fields it got; the
same problem in the FP64 oracle (dense LU of its trace system) gives the
same lambda, u and q to 1e-10."""
import os
import subprocess

import numpy as np
import pytest

import hyperfox_amd as hf
import oracle as orc
from hdg_helpers import rel

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, 'c_caller', 'hdg_caller')


def test_c_caller_matches_oracle(tmp_path):
    assert os.path.exists(EXE), 'build the library first (make -C hyperfox_amd/csrc builds the C caller)'
    out = tmp_path / 'hdg_caller.bin'
    print('C caller of the bulk path', flush=True)
    r = subprocess.run([EXE, str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    raw = out.read_bytes()
    head = np.frombuffer(raw[:48], dtype=np.int32)
    dim, order, n, seed, nNodes, nCells, nFaces, ntrace, nN, nIP, its, reason = (int(v) for v in head)
    vals = np.frombuffer(raw[48:], dtype=np.float64)
    nNf = ntrace // nFaces
    sizes = [1, 1, nFaces * nNf, nNodes, nNodes * dim, nFaces * nNf, nCells * nIP, nCells * nIP, nCells * nN,
             nCells * nN * dim, ntrace]
    assert vals.size == sum(sizes)
    parts = np.split(vals, np.cumsum(sizes)[:-1])
    jitter, relres, tau, D, b, g, f, r_ip, u, q, lam = parts
    assert reason == 2 and relres[0] <= 1e-13 and its > 0
    mesh = hf.Mesh.structured(dim, order, n, float(jitter[0]), seed)
    a = mesh.arrays()
    assert (mesh.n_nodes, mesh.n_cells, mesh.n_faces) == (nNodes, nCells, nFaces)
    s = orc.System(dim, order, 2, a, tau.reshape(nFaces, nNf), tau_sides=1, D=D, ndD=1, b=b.reshape(nNodes, dim),
                   f_ip=f.reshape(nCells, nIP), r_ip=r_ip.reshape(nCells, nIP), dirichlet=g.reshape(nFaces, nNf))
    lo, _, _, _ = s.solve(method=1)
    uo, qo = s.reconstruct(lo)
    print('C caller vs oracle: lambda %.2e, u %.2e, q %.2e' % (rel(lam, lo), rel(u, uo.ravel()), rel(q, qo.ravel())))
    assert rel(lam, lo) < 1e-10
    assert rel(u, uo.ravel()) < 1e-10
    assert rel(q, qo.ravel()) < 1e-10
