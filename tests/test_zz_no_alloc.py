"""hfx_hdg_solve_noalloc (the allocation-free trace solve, no counterpart in
the reference) against hfx_hdg_solve with BiCGStab and the same
preconditioner: the same kernels in the same order, so the same iteration
count, residual, trace, solution and flux bit for bit.  Two calls in a row,
the second on a smaller system (its workspace still holds the first call's
values), then the same pair again in the other preconditioner; a system
larger than the fixed workspace is refused."""
import numpy as np
import pytest

import hyperfox_amd as hf
import hyperfox_amd.problems as pr

pytestmark = pytest.mark.gpu


def assembled(w):
    prob = hf.HDGProblem(w.dim, w.p)
    pr.setup_gpu(w, prob)
    prob.assemble()
    return prob


def reference(prob, pc):
    prob.set_krylov(hf.KRYLOV_BICGSTAB)
    out = prob.solve(pc=pc, rtol=1e-12, maxits=5000)
    return out, [a.copy() for a in prob.solution()]


@pytest.mark.parametrize('pc', [hf.PC_BJACOBI, hf.PC_JACOBI])
def test_noalloc_solve_equals_solve(pc):
    print('noalloc solve, pc %d' % pc, flush=True)
    big = assembled(pr.convection_diffusion_3d(4, 3))      # 384 cells, block size 10
    small = assembled(pr.diffusion_source_2d(8, 2))        # 128 cells, block size 3
    refs = [reference(p, pc) for p in (big, small)]
    got = []
    for prob in (big, small):                              # two calls in a row, smaller second
        out = prob.solve_noalloc(pc=pc, rtol=1e-12, maxits=5000)
        got.append((out, [a.copy() for a in prob.solution()]))
    for (ro, rf), (go, gf) in zip(refs, got):
        assert ro[1] == 2 and ro[0] > 8, ro
        assert go == ro, (go, ro)
        for a, b in zip(gf, rf):
            assert np.array_equal(a, b)


def test_noalloc_refuses_a_system_above_its_workspace():
    print('noalloc size limit', flush=True)
    prob = assembled(pr.diffusion_source_2d(400, 1))       # 320,000 cells, 961,200 trace DOFs
    with pytest.raises(hf.HfxError, match='above the library'):
        prob.solve_noalloc(pc=hf.PC_BJACOBI)
