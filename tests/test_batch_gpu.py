"""Batched assembly of many 2-D problems of mixed order (one persistent
launch per order, run concurrently): every problem's condensed blocks and
trace system equal, bit for bit,
what its own hfx_hdg_assemble produces (the same device code per cell), for
eager launches, for several replays of the captured HIP graph (with eager
work in between) and after autotuning.  At the workload's full shape (64
problems of 1k-40k cells, orders 1..5) the smallest problem of every order
is also compared with the oracle's trace system."""
import hashlib

import numpy as np
import pytest

import hyperfox_amd as hf
import hyperfox_amd.problems as pr
import oracle as orc
from hdg_helpers import rel

pytestmark = pytest.mark.gpu


def build(specs):
    probs = []
    for s in specs:
        w = pr.batch_workload(s)
        prob = hf.HDGProblem(2, s['p'])
        pr.setup_gpu(w, prob)
        probs.append(prob)
    return probs


def snapshot(probs):
    out = []
    for prob in probs:
        c = prob.condensed()
        rp, cols, vals, rhs = prob.trace_csr()
        out.append((c['S'], c['S0'], c['U'], c['Q'], c['U0'], c['Q0'], vals, rhs))
    return out


def same(a, b):
    return all(np.array_equal(x, y) for pa, pb in zip(a, b) for x, y in zip(pa, pb))


@pytest.fixture(scope='module')
def batch_case():
    specs = pr.batch_specs(count=10, seed=3, cells=(20, 400))
    assert {s['p'] for s in specs} >= {1, 2, 3}
    probs = build(specs)
    for prob in probs:
        prob.assemble()
    ref = snapshot(probs)
    return specs, probs, ref


@pytest.mark.parametrize('waves,tile', [(1, 8), (2, 16), (4, 64)])
def test_batched_equals_per_problem(batch_case, waves, tile):
    specs, probs, ref = batch_case
    b = hf.Batch()
    for prob in probs:
        b.add(prob)
    b.set_config(waves, tile)
    # poison the outputs so stale values cannot pass
    for prob in probs:
        prob.set_kernel(1)
        prob.assemble()
        prob.set_kernel(0)
    b.assemble(use_graph=False)
    assert same(snapshot(probs), ref)


def test_graph_replays_and_tuning(batch_case):
    specs, probs, ref = batch_case
    b = hf.Batch()
    for prob in probs:
        b.add(prob)
    b.set_config(2, 32)
    for rep in range(3):
        for prob in probs:           # eager work between replays
            prob.set_kernel(1)
            prob.assemble()
            prob.set_kernel(0)
        b.assemble(use_graph=True)
        assert same(snapshot(probs), ref), rep
    best = b.tune(reps=2)
    cfg = b.config()
    print('tuned: %s, %.3f ms' % (cfg, best))
    assert best > 0 and all(c['waves'] in (1, 2, 4) and c['items'] > 0 for c in cfg.values())
    assert set(cfg) == {s['p'] for s in specs}
    b.assemble(use_graph=True)
    assert same(snapshot(probs), ref)
    t = b.time(3, warmup=1, reps=2)
    assert np.all(t > 0)


def test_batch_solves_match_single_problem(batch_case):
    specs, probs, ref = batch_case
    b = hf.Batch()
    for prob in probs[:3]:
        b.add(prob)
    b.assemble(use_graph=False)
    for prob in probs[:3]:
        prob.set_krylov(hf.KRYLOV_BICGSTAB)
        its, reason, rel = prob.solve(pc=hf.PC_BJACOBI, rtol=1e-12, maxits=20000)
        assert reason == 2 and rel <= 1e-12


def test_batch_rejects_unsupported_problems():
    m = hf.Mesh.structured(3, 1, 2, 0.0, 1)
    prob = hf.HDGProblem(3, 1)
    prob.set_mesh(m)
    prob.set_model(hf.MODEL_LAPLACE)
    prob.set_fields(np.ones((m.n_faces, m.sizes['nNf'])), 1)
    b = hf.Batch()
    with pytest.raises(hf.HfxError, match='2-D'):
        b.add(prob)


def digest(probs):
    """SHA-1 of every problem's condensed blocks and trace system (the full
    batch holds a few GB of them: compared by digest, not kept)."""
    out = []
    for prob in probs:
        h = hashlib.sha1()
        c = prob.condensed()
        for k in ('S', 'S0', 'U', 'Q', 'U0', 'Q0'):
            h.update(c[k].tobytes())
        del c
        rp, cols, vals, rhs = prob.trace_csr()
        h.update(vals.tobytes())
        h.update(rhs.tobytes())
        out.append(h.hexdigest())
    return out


def test_full_batch_workload_against_oracle():
    """The batched workload as specified: 64 problems, cell counts
    log-uniform in [1k, 40k], orders 1..5, diffusivity log-uniform in
    [1e-2, 1e2], 8-mode Fourier sources.  Every problem of the batched
    assembly (eager and graph replay) is bit-identical to its own
    hfx_hdg_assemble, and the smallest problem of each order has the
    oracle's trace matrix and RHS (relative 1e-12, entries to 1e-11 of the
    largest)."""
    specs = pr.batch_specs()
    assert len(specs) == 64 and {s['p'] for s in specs} == {1, 2, 3, 4, 5}
    ws = [pr.batch_workload(s) for s in specs]
    cells = sum(w.mesh.n_cells for w in ws)
    print('full batch: 64 problems, %d cells' % cells, flush=True)
    assert all(500 <= w.mesh.n_cells <= 45000 for w in ws)
    probs, fs = [], []
    for w in ws:
        prob = hf.HDGProblem(2, w.p)
        fs.append(pr.setup_gpu(w, prob))
        probs.append(prob)
    for prob in probs:
        prob.assemble()
    ref = digest(probs)
    b = hf.Batch()
    for prob in probs:
        b.add(prob)
    for use_graph in (False, True, True):
        for prob in probs:               # other values in the outputs first
            prob.set_kernel(1)
            prob.assemble()
            prob.set_kernel(0)
        b.assemble(use_graph=use_graph)
        assert digest(probs) == ref, use_graph
    for p in range(1, 6):
        k = min((i for i in range(64) if specs[i]['p'] == p), key=lambda i: ws[i].mesh.n_cells)
        w, prob = ws[k], probs[k]
        s = orc.System(2, p, w.kind, w.mesh.arrays(), w.tau, tau_sides=w.tau_sides, D=w.D, ndD=w.ndD, f_ip=fs[k],
                       dirichlet=w.dirichlet)
        rp, cols, vals, rhs = prob.trace_csr()
        rpo, colso, valso, rhso = s.csr()
        print('order %d: problem %d, %d cells, trace matrix rel %.2e, rhs rel %.2e'
              % (p, k, w.mesh.n_cells, rel(vals, valso), rel(rhs, rhso)), flush=True)
        assert np.array_equal(rp, rpo) and np.array_equal(cols, colso)
        assert rel(vals, valso) < 1e-12 and rel(rhs, rhso) < 1e-12
        assert np.abs(vals - valso).max() <= 1e-11 * np.abs(valso).max()
