"""Benchmark: HDG assembly + static condensation + trace scatter on one MI355X.

Workload: 3-D HDG convection-diffusion, unit cube split into 24^3 x 6 = 82,944
jittered tetrahedra, p = 3, D = 1e-2, b = (sin 2pi y, sin 2pi z, sin 2pi x),
manufactured u = e^{x+y+z} sin(pi x) sin(pi y) sin(pi z), upwind
double-valued tau.  One step = one pass of [element assembly + condensation +
Dirichlet rows + BSR scatter] over all elements, inputs resident in HBM.
value = u-DOFs (cells x nodes per cell) per second over the K timed steps.
The trace solve (BiCGStab and FGMRES(30) with point Jacobi, block Jacobi and
block ILU(0), to a true FP64 relative residual of 1e-10) is timed after the
steps and reported beside the main number, with the learned preconditioner
against block Jacobi and ILU(0) on the 2-D GRF-coefficient workload
(256 x 256 x 2 triangles, p = 3).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 24] [--no-cpu]
       python bench.py --copy-calibration MiB   (a plain device copy of known
       size, run under rocprofv3 --pmc to scale FETCH_SIZE / WRITE_SIZE)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'DOF/s assembly+condense; trace-solve time to 1e-10 residual (learned vs plain)'
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# HBM KiB per element-kernel dispatch, (read, write), from rocprofv3 --pmc
# FETCH_SIZE / WRITE_SIZE runs of this script, scaled by a 1 GiB copy kernel
# (FETCH_SIZE x2.00, WRITE_SIZE x1.00); the summary is kept under profiles/.
MEASURED_TRAFFIC_KIB = {(24, 3): (2 * 80524.0, 3901410.0)}


def elem_flops(dim, p, nN, nIP, nNf, nIPf, conv):
    """Algorithmic flops of assembly + condensation per element, crediting
    only the block-diagonal Sqq = I_d (x) M (same count for every code path)."""
    nU, nQ, nL = nN, dim * nN, (dim + 1) * nNf
    c = 1 if conv else 0
    f_asm = 2 * nIP * nN * nN * (1 + 2 * dim + c) + (dim + 1) * 2 * nIPf * nNf * nNf * (1 + dim + c)
    f_cond = (2.0 / 3.0) * nN ** 3 + 2 * dim * nN * nN * (nU + nL) + 2 * nU * nQ * nU + (2.0 / 3.0) * nU ** 3 \
        + 2 * nU * nQ * nL + 2 * nU * nU * nL + 2 * nQ * nU * nL + 2 * nL * nU * nL + 2 * nL * nQ * nL
    return f_asm + f_cond


def elem_bytes(dim, nN, nNf, nIP, nnz_dofs_in):
    """Algorithmic HBM bytes per element of the element kernel: inputs
    (node coords + node fields, tau, source at IPs, connectivity) and FP64
    outputs S, S0, U, Q, U0, Q0."""
    nL = (dim + 1) * nNf
    out = nL * nL + nL + nN * nL + dim * nN * nL + nN + dim * nN
    return 8 * (out + nnz_dofs_in) + 4 * nN


def cpu_system(n, p, threads):
    """The oracle's whole-mesh pipeline on the benchmark workload at n^3 x 6."""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import oracle as orc
    import hyperfox_amd.problems as pr
    w = pr.convection_diffusion_3d(n, p)
    a = w.mesh.arrays()
    X = a['coords'][a['cells']]
    # source at IPs on the CPU (same formula as the GPU leg)
    t = orc.refel(3, p)
    xip = np.einsum('qn,end->eqd', t['phi'], X)
    f = w.source_at(xip)
    s = orc.System(3, p, 2, a, w.tau, tau_sides=2, D=w.D, ndD=1, b=w.b, f_ip=f, dirichlet=w.dirichlet,
                   nthreads=threads)
    return s, w.mesh.n_cells * orc.sizes(3, p)['nN'], w.mesh.n_cells


def host_info(threads):
    """The host the CPU legs ran on: model name (lscpu), logical CPUs,
    OMP_NUM_THREADS and the OpenMP threads the oracle was given."""
    model = None
    try:
        import subprocess
        for line in subprocess.run(['lscpu'], capture_output=True, text=True, timeout=10).stdout.splitlines():
            if line.startswith('Model name:'):
                model = line.split(':', 1)[1].strip()
                break
    except Exception:
        pass
    if model is None:
        try:
            with open('/proc/cpuinfo') as fh:
                model = next((ln.split(':', 1)[1].strip() for ln in fh if ln.startswith('model name')), None)
        except OSError:
            pass
    return {'cpu_model': model, 'logical_cpus': os.cpu_count(), 'OMP_NUM_THREADS': os.environ.get('OMP_NUM_THREADS'),
            'oracle_threads': threads}


def oracle_gmres(s, rtol=1e-10):
    """The reference's trace solve on the oracle's system: GMRES(30) with left
    point Jacobi (PETSc's KSPGMRES + PCJACOBI as PetscInterface configures
    them), zero initial guess, stopping on the preconditioned residual."""
    t = time.perf_counter()
    lam, its, conv, relres = s.solve(method=0, rtol=rtol, maxits=100000, restart=30)
    return lam, {'method': 'GMRES(30) + left Jacobi (PETSc semantics)', 'rtol_preconditioned': rtol,
                 'iters': its, 'converged': conv, 'true_relres': relres, 's': time.perf_counter() - t}


def cpu_baseline(n_cpu, p, threads, n_cpu1=8):
    """Oracle (FP64 C++ restatement of the reference path: assembly, Householder
    condensation, Dirichlet rows, CSR scatter) timed on the host cores: on all
    `threads` OpenMP threads over n_cpu^3 x 6 cells (best of 2), and on one core
    over n_cpu1^3 x 6 cells.  Per-cell cost does not depend on the mesh size.
    The trace solve of the same system by the reference's GMRES(30) + Jacobi
    to 1e-10 is timed beside it."""
    s, dofs, cells = cpu_system(n_cpu, p, threads)
    best = min(s.time_assembly(cells, threads) for _ in range(2))
    _, solve = oracle_gmres(s)
    print('# CPU oracle: assembly %.2f s, GMRES(30)+Jacobi %d its %.1f s' % (best, solve['iters'], solve['s']),
          file=sys.stderr, flush=True)
    del s
    s1, dofs1, cells1 = cpu_system(n_cpu1, p, 1)
    best1 = min(s1.time_assembly(cells1, 1) for _ in range(2))
    return {'value': dofs / best, 'unit': 'DOF/s', 'cores': threads, 'kind': 'port',
            'sample': '%d^3x6 = %d tets, same p/model/fields; assembly+condensation+BC+CSR scatter, best of 2: %.2f s on '
                      '%d threads' % (n_cpu, cells, best, threads),
            'value_1core': dofs1 / best1,
            'sample_1core': '%d^3x6 = %d tets on 1 thread, best of 2: %.2f s' % (n_cpu1, cells1, best1),
            'trace_solve': solve, 'host': host_info(threads)}


def configs_2d(hf, pr, threads, reps=10):
    """The 2-D diffusion-source configurations at full size: 32 x 32 x 2
    triangles at p = 2 and 128 x 128 x 2 at p = 4 (jittered 0.1 h, D = 1,
    u = sin(pi x) sin(pi y)).  Per configuration: GPU assembly + condensation
    + scatter (HIP events, median of `reps` after 3 warm-ups) as u-DOF/s and
    trace-DOF/s; GPU trace solves to a true residual of 1e-10; the solution
    error of the GPU path against the FP64 oracle's own system solved by
    sparse LU (GPU solved to 1e-13), for lambda, u and q, and both L2 errors
    against the exact solution; the oracle's assembly rate on the host
    threads, and on the smaller mesh its GMRES(30) + Jacobi solve."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    sys.path.insert(0, os.path.join(ROOT, 'tests'))
    import oracle as orc
    from hdg_helpers import l2_projection_error
    out = {}
    for name, n, p in (('32x32x2_p2', 32, 2), ('128x128x2_p4', 128, 4)):
        w = pr.diffusion_source_2d(n, p)
        prob = hf.HDGProblem(2, p)
        f = pr.setup_gpu(w, prob)
        prob.assemble()
        cells, nN = w.mesh.n_cells, prob.sizes['nN']
        ms = float(np.median(prob.time(1, warmup=3, reps=reps)))
        kms = float(np.mean(prob.time(0, warmup=1, reps=reps)))
        solves = {}
        for method, pc in (('bicgstab', 'ilu0'), ('bicgstab', 'block-jacobi'), ('fgmres30', 'jacobi')):
            solves['%s+%s' % (method, pc)] = timed_solve(hf, prob, method, pc, 100000)
        s = orc.System(2, p, w.kind, w.mesh.arrays(), w.tau, tau_sides=w.tau_sides, f_ip=f, dirichlet=w.dirichlet,
                       nthreads=threads)
        rp, cols, vals, rhs = s.csr()
        N = s.ndofs
        t = time.perf_counter()
        lo = spla.splu(sp.csr_matrix((vals, cols, rp), shape=(N, N)).tocsc()).solve(rhs)
        lu_s = time.perf_counter() - t
        uo, qo = s.reconstruct(lo)
        prob.set_krylov(hf.KRYLOV_BICGSTAB)
        its, reason, relres = prob.solve(pc=hf.PC_ILU0, rtol=1e-13, maxits=100000)
        u, q, lam = prob.solution()
        a = w.mesh.arrays()
        ue = pr.u_sin2d(a['coords'][a['cells']])

        def rl(x, y):
            return float(np.linalg.norm(x - y) / np.linalg.norm(y))
        row = {'workload': 'HDG 2D diffusion-source p=%d, %d^2x2=%d jittered tris' % (p, n, cells),
               'cells': cells, 'u_dofs': cells * nN, 'trace_dofs': int(w.mesh.n_trace),
               'gpu': {'assembly_ms': ms, 'element_kernel_ms': kms, 'dof_per_s': cells * nN / (ms * 1e-3),
                       'trace_dof_per_s': w.mesh.n_trace / (ms * 1e-3), 'solves_to_1e-10': solves},
               'vs_fp64_oracle': {'oracle_solve': 'sparse LU (SuperLU) of the oracle trace system, %.1f s' % lu_s,
                                  'gpu_solve': 'BiCGStab + ILU(0), %d its, true relres %.1e' % (its, relres),
                                  'lambda_rel_l2': rl(lam, lo), 'u_rel_l2': rl(u, uo), 'q_rel_l2': rl(q, qo)},
               'l2_error_vs_exact': {'gpu': l2_projection_error(2, p, a, u, ue),
                                     'oracle': l2_projection_error(2, p, a, uo, ue)}}
        t_asm = min(s.time_assembly(cells, threads) for _ in range(2))
        row['cpu'] = {'assembly_s': t_asm, 'dof_per_s': cells * nN / t_asm, 'threads': threads}
        if n <= 32:
            _, row['cpu']['trace_solve'] = oracle_gmres(s)
        else:
            row['cpu']['trace_solve'] = ('not run in the bench: GMRES(30) + Jacobi needs about 21,000 iterations '
                                         'on this system (51 s on 8 host threads)')
        del s
        out[name] = row
        print('# %s: GPU %.3f ms/assembly (%.1f M DOF/s); vs oracle lambda %.1e u %.1e q %.1e; L2 err %.3e' % (
            name, ms, cells * nN / (ms * 1e-3) / 1e6, row['vs_fp64_oracle']['lambda_rel_l2'],
            row['vs_fp64_oracle']['u_rel_l2'], row['vs_fp64_oracle']['q_rel_l2'], row['l2_error_vs_exact']['gpu']),
            file=sys.stderr, flush=True)
    return out


PC_NAMES = {'jacobi': 1, 'block-jacobi': 2, 'learned': 3, 'ilu0': 4, 'ilu0-natural': 5, 'ilu0-f32': 6}


def solver_bytes_per_iter(n, nb, bs, nnzb, method, pc='block-jacobi'):
    """Algorithmic HBM bytes of one Krylov iteration: BiCGStab = 2 BSR SpMV
    (values, column ids, x and y once) + 2 preconditioner applies + 21 vector
    passes; FGMRES at Krylov dimension ~15 = 1 SpMV + 1 apply + 2 x (2k + 3)
    vector passes (two Gram-Schmidt sweeps).  Applies: point Jacobi 3 vector
    passes; block Jacobi the inverse diagonal blocks + 2 passes; ILU(0) every
    factor block (FP64) with its column id + the inverse pivot blocks + 4
    passes (ilu0-f32: the factor blocks in FP32); learned the FP32 BSR matrix
    + column ids + 2 passes."""
    spmv = nnzb * bs * bs * 8 + nnzb * 4 + 2 * n * 8
    apply = {'jacobi': 3 * n * 8, 'block-jacobi': nb * bs * bs * 8 + 2 * n * 8,
             'ilu0': nnzb * bs * bs * 8 + nnzb * 8 + nb * bs * bs * 8 + 4 * n * 8,
             'ilu0-f32': nnzb * bs * bs * 4 + nnzb * 8 + nb * bs * bs * 4 + 4 * n * 8,
             'learned': nnzb * bs * bs * 4 + nnzb * 4 + 2 * n * 8}[pc.replace('-natural', '')]
    if method == 'bicgstab':
        return 2 * spmv + 2 * apply + 21 * n * 8
    return spmv + apply + 2 * (2 * 15 + 3) * n * 8


def timed_solve(hf, prob, method, pc, maxits, warm=True):
    """One solve to a true residual of 1e-10 (after an untimed warm-up solve
    that allocates the Krylov work space), wall time in ms."""
    from hyperfox_amd.lpc import pc_setup_ms
    prob.set_krylov(hf.KRYLOV_BICGSTAB if method == 'bicgstab' else hf.KRYLOV_FGMRES)
    if warm:
        prob.solve(pc=PC_NAMES[pc], rtol=1e-10, maxits=min(maxits, 50), restart=30)
    t = time.perf_counter()
    its, reason, relres = prob.solve(pc=PC_NAMES[pc], rtol=1e-10, maxits=maxits, restart=30)
    ms = (time.perf_counter() - t) * 1e3
    return {'method': method, 'pc': pc, 'iters': its, 'reason': reason, 'converged': reason == 2,
            'true_relres': relres, 'ms': ms, 'pc_setup_ms': pc_setup_ms(prob)}


def topology_times(hf, w, reps=3):
    """Mesh topology (faces, cell2face, face2cell, boundary, face-node
    permutations) of the benchmark mesh from its cells, on the host and on the
    GPU (identical arrays), best of `reps`, wall ms including the cell copy."""
    a = w.mesh.arrays()
    out = {}
    for name, dev in (('host_ms', None), ('gpu_ms', 0)):
        best = 1e30
        for _ in range(reps):
            t = time.perf_counter()
            m = hf.Mesh.from_cells(3, w.p, a['coords'], a['cells'], device=dev)
            best = min(best, (time.perf_counter() - t) * 1e3)
        out[name] = best
        out['faces'] = m.n_faces
    print('# topology: host %.1f ms, GPU %.1f ms' % (out['host_ms'], out['gpu_ms']), file=sys.stderr, flush=True)
    return out


def trace_solves(hf, pr, prob, w):
    """Trace solve of the benchmark problem to a true FP64 relative residual
    of 1e-10: BiCGStab with block Jacobi and block ILU(0) (multicolour order),
    and the reference's GMRES(30) (as FGMRES) with point Jacobi (PETSc's
    PCJACOBI), block Jacobi and ILU(0)."""
    n_dofs, nnz, nnzb = prob.trace_sizes()
    bs = prob.sizes['nNf']
    nb = n_dofs // bs
    out = {}
    for method, pc in (('bicgstab', 'block-jacobi'), ('bicgstab', 'ilu0'), ('bicgstab', 'ilu0-f32'),
                       ('fgmres30', 'jacobi'), ('fgmres30', 'block-jacobi'), ('fgmres30', 'ilu0')):
        r = timed_solve(hf, prob, method, pc, 20000)
        by = solver_bytes_per_iter(n_dofs, nb, bs, nnzb, method, pc) * r['iters']
        r.update(algorithmic_GB=by / 1e9, achieved_GBs=by / (r['ms'] * 1e-3) / 1e9)
        out['%s+%s' % (method, pc)] = r
        print('# solve %s + %s: %d its, relres %.2e, %.1f ms' % (method, pc, r['iters'], r['true_relres'], r['ms']),
              file=sys.stderr, flush=True)
    prob.set_krylov(hf.KRYLOV_BICGSTAB)
    prob.solve(pc=hf.PC_BJACOBI, rtol=1e-10, maxits=20000)
    u, q, lam = prob.solution()
    a = w.mesh.arrays()
    ue = pr.u_c3(a['coords'][a['cells']])
    out['nodal_rel_err_vs_exact'] = float(np.linalg.norm(u - ue) / np.linalg.norm(ue))
    return out


def transient(hf, pr, prob, w, steps=3, dt=1e-2):
    """Implicit Euler on the benchmark problem (the reference's Euler with
    isExplicit = false on the u rows): from u(0) = the manufactured solution,
    each step assembles with the previous step's solution field and solves
    (BiCGStab + ILU(0), true residual 1e-10).  Reports per-step times and the
    assembly rate with the scheme on (HIP events, element kernel + scatter)."""
    a = w.mesh.arrays()
    u0 = pr.u_c3(a['coords'][a['cells']])
    prob.set_solution(u0)
    prob.set_time_scheme(hf.TS_BEULER, dt)
    prob.set_krylov(hf.KRYLOV_BICGSTAB)
    rows = []
    for k in range(steps):
        t0 = time.perf_counter()
        prob.assemble()
        t1 = time.perf_counter()
        its, reason, relres = prob.solve(pc=hf.PC_ILU0, rtol=1e-10, maxits=20000)
        t2 = time.perf_counter()
        u = prob.solution()[0]
        rows.append({'step': k + 1, 'assemble_ms': (t1 - t0) * 1e3, 'solve_ms': (t2 - t1) * 1e3, 'iters': its,
                     'reason': reason, 'true_relres': relres, 'max_u': float(np.abs(u).max())})
        print('# implicit Euler step %d: assemble %.2f ms, solve %d its %.1f ms' % (
            k + 1, rows[-1]['assemble_ms'], its, rows[-1]['solve_ms']), file=sys.stderr, flush=True)
    ms = float(np.median(prob.time(1, warmup=1, reps=5)))
    dofs = w.mesh.n_cells * prob.sizes['nN']
    # diagonally implicit Runge-Kutta (ALX2, two stages; the reference's
    # RungeKutta with the auxiliary fields Flux and Trace) from the last
    # solution: per stage assemble, solve, computeStage; then computeSolution
    prob.set_time_scheme(hf.TS_RK, dt)
    prob.set_butcher('ALX2')
    rk = []
    for k in range(2):
        t0 = time.perf_counter()
        infos = prob.rk_step(pc=hf.PC_ILU0, rtol=1e-10, maxits=20000)
        t1 = time.perf_counter()
        rk.append({'step_ms': (t1 - t0) * 1e3, 'stage_iters': [i for i, _, _ in infos],
                   'stage_reasons': [r for _, r, _ in infos], 'max_u': float(np.abs(prob.solution()[0]).max())})
        print('# ALX2 step %d: %.1f ms, stage iterations %s' % (k + 1, rk[-1]['step_ms'], rk[-1]['stage_iters']),
              file=sys.stderr, flush=True)
    prob.rk_start_step()
    ms_rk = float(np.median(prob.time(1, warmup=1, reps=5)))  # stage-0 assembly (local vectors included)
    prob.set_time_scheme(hf.TS_NONE)
    # the reference's strongly explicit HDG mode: S = Sll, S0 from the current
    # solution and flux fields, face-local trace solves
    prob.set_solver_type(hf.HDG_SEXPLICIT)
    ex = []
    for k in range(3):
        t0 = time.perf_counter()
        prob.assemble()
        t1 = time.perf_counter()
        its, reason, relres = prob.solve(pc=hf.PC_BJACOBI, rtol=1e-10, maxits=10)
        t2 = time.perf_counter()
        ex.append({'assemble_ms': (t1 - t0) * 1e3, 'face_solves_and_reconstruction_ms': (t2 - t1) * 1e3,
                   'true_relres': relres})
    prob.set_solver_type(hf.HDG_IMPLICIT)
    print('# strongly explicit HDG step: assemble %.2f ms, face solves + reconstruction %.2f ms' % (
        ex[-1]['assemble_ms'], ex[-1]['face_solves_and_reconstruction_ms']), file=sys.stderr, flush=True)
    return {'scheme': 'implicit Euler (Euler, isExplicit=false)', 'dt': dt, 'steps': rows,
            'assembly_ms_with_scheme': ms, 'assembly_dof_per_s_with_scheme': dofs / (ms * 1e-3),
            'rk_alx2': {'steps': rk, 'stage_assembly_ms': ms_rk, 'stage_assembly_dof_per_s': dofs / (ms_rk * 1e-3)},
            'sexplicit_steps': ex}


def learned_vs_plain(hf, pr, n, fgmres_maxits=2000, seeds=12):
    """Learned-preconditioner workload: 2-D HDG diffusion, n x n x 2 jittered
    triangles, p = 3, k = exp(GRF(ell 0.1, sigma 1)), f = 1, `seeds` held-out
    GRF seeds (disjoint from the training seeds).  BiCGStab to a true residual of 1e-10 with block Jacobi, with the
    trained network's correction of block Jacobi (BF16 MFMA inference; set-up
    time included) and with block ILU(0); FGMRES(30) with the same three
    preconditioners, capped at `fgmres_maxits` iterations (on these
    coefficient contrasts restarted GMRES stagnates far above 1e-10)."""
    path = os.path.join(ROOT, 'hyperfox_amd', 'data', 'lpc_grf2d_p3.npy')
    if not os.path.exists(path):
        return {'skipped': 'no trained network at hyperfox_amd/data/lpc_grf2d_p3.npy'}
    from hyperfox_amd.lpc import LearnedPreconditioner
    net = LearnedPreconditioner(4)
    net.load(path)
    mesh = hf.Mesh.structured(2, 3, n, 0.1, 2024)
    prob = hf.HDGProblem(2, 3)
    net.attach(prob, bf16=True)
    rows, fg = [], []
    for i, seed in enumerate(900000 + n * 1000 + 500 + k for k in range(seeds)):
        w = pr.diffusion_grf_2d(n, 3, seed, mesh=mesh)
        if i == 0:
            pr.setup_gpu(w, prob)
        else:
            prob.set_fields(w.tau, w.tau_sides, w.D, w.ndD, w.b, w.dirichlet)
        prob.assemble()
        runs = {}
        for pc in ('block-jacobi', 'learned', 'ilu0'):
            runs[pc] = timed_solve(hf, prob, 'bicgstab', pc, 50000)
            runs[pc]['lambda'] = prob.solution()[2].copy()
        lam_b = runs['block-jacobi'].pop('lambda')
        row = {'seed': seed}
        for pc in ('learned', 'ilu0'):
            lam = runs[pc].pop('lambda')
            runs[pc]['solution_rel_diff_vs_block_jacobi'] = float(np.abs(lam - lam_b).max() / np.abs(lam_b).max())
        row.update(runs)
        rows.append(row)
        print('# learned vs plain n=%d seed %d: BiCGStab block Jacobi %d its %.1f ms, learned %d its %.1f ms, '
              'ILU(0) %d its %.1f ms' % (n, seed, runs['block-jacobi']['iters'], runs['block-jacobi']['ms'],
                                         runs['learned']['iters'], runs['learned']['ms'], runs['ilu0']['iters'],
                                         runs['ilu0']['ms']), file=sys.stderr, flush=True)
        if i == 0:
            for pc in ('block-jacobi', 'learned', 'ilu0'):
                r = timed_solve(hf, prob, 'fgmres30', pc, fgmres_maxits, warm=False)
                r['seed'] = seed
                fg.append(r)
                print('# FGMRES(30) + %s: %d its, relres %.2e, %.1f ms' % (pc, r['iters'], r['true_relres'], r['ms']),
                      file=sys.stderr, flush=True)

    def ratio(a, b, key):
        return float(np.mean([r[a][key] / r[b][key] for r in rows]))

    def total(pc, key):
        return float(np.sum([r[pc][key] for r in rows]))

    return {'workload': 'HDG 2D diffusion, %dx%dx2=%d tris, p=3, k=exp(GRF ell=0.1 sigma=1), f=1' % (n, n, 2 * n * n),
            'trace_dofs': int(mesh.n_trace), 'method': 'BiCGStab FP64, true relres 1e-10', 'plain_pc': 'block-jacobi',
            'learned_pc': 'block-jacobi + GNN correction (BF16 MFMA inference)', 'runs': rows,
            'iters_ratio': ratio('block-jacobi', 'learned', 'iters'),
            'time_ratio': ratio('block-jacobi', 'learned', 'ms'),
            'ilu0_iters_ratio': ratio('block-jacobi', 'ilu0', 'iters'),
            'ilu0_time_ratio': ratio('block-jacobi', 'ilu0', 'ms'),
            'held_out_seeds': len(rows),
            'total_ms': {pc: total(pc, 'ms') for pc in ('block-jacobi', 'learned', 'ilu0')},
            'total_iters': {pc: total(pc, 'iters') for pc in ('block-jacobi', 'learned', 'ilu0')},
            'learned_vs_ilu0_time': total('ilu0', 'ms') / total('learned', 'ms'),
            'fgmres30_capped': fg}


def batched(hf, pr, count=64, seed=7):
    """Batched workload: `count` independent 2-D diffusion problems (cells
    log-uniform in [1k, 40k], p uniform in 1..5, D log-uniform in
    [1e-2, 1e2], 8-mode Fourier sources) assembled by one persistent launch
    per order (running concurrently) whose (waves, tile) the autotuner picks,
    eager and as a replayed HIP graph, against the same problems launched one
    by one."""
    specs = pr.batch_specs(count=count, seed=seed)
    probs = []
    for sp in specs:
        w = pr.batch_workload(sp)
        prob = hf.HDGProblem(2, sp['p'])
        pr.setup_gpu(w, prob)
        probs.append(prob)
    b = hf.Batch()
    for prob in probs:
        b.add(prob)
    dofs = sum(prob.mesh.n_cells * prob.sizes['nN'] for prob in probs)
    cells = sum(prob.mesh.n_cells for prob in probs)
    best = b.tune(reps=3)
    cfg = {str(k): v for k, v in b.config().items()}
    cfg['concurrent_orders'] = b.concurrent
    cfg['concurrent_orders_in_graph'] = b.graph_concurrent
    out = {'workload': '%d 2-D HDG diffusion problems, %d cells, orders %s' % (
        count, cells, sorted({sp['p'] for sp in specs})), 'u_dofs': dofs, 'tuned': cfg, 'tune_best_kernel_ms': best}
    for name, what in (('kernel_only', 0), ('eager', 1), ('graph', 2), ('one_by_one', 3)):
        ms = float(np.median(b.time(what, warmup=1, reps=5)))
        out[name] = {'ms': ms, 'dof_per_s': dofs / (ms * 1e-3)}
        print('# batched %s: %.3f ms' % (name, ms), file=sys.stderr, flush=True)
    out['graph_vs_one_by_one'] = out['one_by_one']['ms'] / out['graph']['ms']
    return out


def cole_hopf(X, D=1.0, A=0.5, x0=(1.0, 0.0)):
    """Stationary viscous Burgers solution u = -2 D grad(pot) / pot of the
    reference's regression case (A = 0.5, x0 = (1, 0))."""
    x, y = X[:, 0] - x0[0], X[:, 1] - x0[1]
    ep, em = np.exp(A * x), np.exp(-A * x)
    pot = 0.001 * A * np.exp((1 + x0[0]) * A) * (1 + X[:, 0]) + (ep + em) * np.cos(A * y)
    gx = 0.001 * A * np.exp((1 + x0[0]) * A) + A * (ep - em) * np.cos(A * y)
    gy = -A * (ep + em) * np.sin(A * y)
    return -2.0 * D * np.stack([gx, gy], 1) / pot[:, None]


def burgers_flops(p):
    """Dominant FP64 flops of one Burgers cell (2-D, m = 2): the two
    Gauss-Jordan eliminations and the U, Q, S output products."""
    nN, nNf = (p + 1) * (p + 2) // 2, p + 1
    NL = 3 * nNf
    MU, ML, NC1 = 2 * nN, 2 * NL, nN + 2 * nN + 2 * NL
    gj1 = 2.0 * nN * nN * NC1
    gj2 = 2.0 * MU * MU * (MU + ML + 1)
    outs = 2.0 * 2 * MU * (ML + 1) * nN + 2.0 * ML * (ML + 1) * MU
    return gj1 + gj2 + outs


def burgers(hf, n=128, p=3, reps=5, graphs=False):
    """The Burgers path on the reference's stationary regression case scaled
    up: 2-D viscous Burgers (m = 2) on [0, 1]^2 with the Cole-Hopf solution
    as integrated Dirichlet data, D = 1, tau = 2 D / h I, zero initial
    guess, Newton to 1e-6 (BiCGStab + ILU(0) to 1e-10)."""
    mesh = hf.Mesh.structured(2, p, n, 0.0, 0)
    a = mesh.arrays()
    s = mesh.sizes
    h = 1.0 / n
    tau = np.zeros((mesh.n_faces, s['nNf'], 2, 2, 2))
    tau[:] = np.eye(2) * (2.0 / h)
    g = cole_hopf(a['coords'][a['faces'].reshape(-1)]).reshape(mesh.n_faces, s['nNf'], 2)
    prob = hf.BurgersProblem(2, p)
    prob.set_mesh(mesh)
    prob.set_fields(tau, D=np.ones(mesh.n_nodes), dirichlet=g)
    prob.set_krylov(1)
    prob.set_solver_graphs(graphs)
    prob.assemble()  # warm-up of the cell kernel and scatter
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        prob.assemble()
        t.append(1e3 * (time.perf_counter() - t0))
    asm_ms = float(np.median(t))
    z = np.zeros((mesh.n_cells, s['nN'], 2))
    prob.set_state(sol=z, buffer=z, trace=np.zeros((mesh.n_faces, s['nNf'], 2)))
    t0 = time.perf_counter()
    its, res, lin = prob.nonlinear_solve(max_iters=10, res_tol=1e-6, pc=4, rtol=1e-10, maxits=20000)
    nl_ms = 1e3 * (time.perf_counter() - t0)
    sol = prob.state()[0].reshape(mesh.n_cells, -1)
    ana = cole_hopf(a['coords'][a['cells'].reshape(-1)]).reshape(mesh.n_cells, -1)
    dofs = mesh.n_cells * s['nN'] * 2
    out = {'workload': '2-D viscous Burgers p=%d, %d^2x2=%d triangles, Cole-Hopf data' % (p, n, mesh.n_cells),
           'cells': mesh.n_cells, 'u_dofs': dofs, 'trace_dofs': mesh.n_faces * s['nNf'] * 2,
           'assemble_ms': asm_ms, 'assemble_dof_per_s': dofs / (asm_ms * 1e-3),
           'cell_flops': burgers_flops(p), 'newton_iters': its, 'residual': res, 'linear_iters': lin,
           'nonlinear_solve_ms': nl_ms,
           'rel_err_vs_exact': float(np.linalg.norm(sol - ana) / np.linalg.norm(ana))}
    print('# burgers: assemble %.2f ms, %d Newton iterations (%d linear) in %.1f ms, error %.2e' % (
        asm_ms, its, lin, nl_ms, out['rel_err_vs_exact']), file=sys.stderr, flush=True)
    return out


def ngamma(hf, n=128, p=3):
    """HDGnGammaModel on the manufactured (n, Gamma) problem of the tests
    scaled up: n = 1 + 0.3 sin(pi x) sin(pi y), Gamma = 0.2 cos(pi x)
    sin(pi y), b = (cos 0.4, sin 0.4), c_s = 1, D = 0.1, G = 0.15, tau = (2 +
    D n) I, integrated Dirichlet data; Newton from (1, 0) to 1e-10 with
    BiCGStab + ILU(0) to 1e-12."""
    A_N, B_G, CS, DIF, GDIF = 0.3, 0.2, 1.0, 0.1, 0.15
    bdir = np.array([np.cos(0.4), np.sin(0.4)])

    def exact(X):
        return np.stack([1.0 + A_N * np.sin(np.pi * X[:, 0]) * np.sin(np.pi * X[:, 1]),
                         B_G * np.cos(np.pi * X[:, 0]) * np.sin(np.pi * X[:, 1])], 1)

    def source(X):
        s1, c1 = np.sin(np.pi * X[:, 0]), np.cos(np.pi * X[:, 0])
        s2, c2 = np.sin(np.pi * X[:, 1]), np.cos(np.pi * X[:, 1])
        nn, g = 1.0 + A_N * s1 * s2, B_G * c1 * s2
        gn = A_N * np.pi * np.stack([c1 * s2, s1 * c2], 1)
        gg = B_G * np.pi * np.stack([-s1 * s2, c1 * c2], 1)
        gF = (2 * g / nn)[:, None] * gg - (g * g / (nn * nn))[:, None] * gn + CS * CS * gn
        return np.stack([gg @ bdir + DIF * 2 * np.pi ** 2 * A_N * s1 * s2,
                         gF @ bdir + GDIF * 2 * np.pi ** 2 * g], 1)

    mesh = hf.Mesh.structured(2, p, n, 0.0, 0)
    a = mesh.arrays()
    s = mesh.sizes
    tabs = hf.refel_tables(2, p)
    xq = np.einsum('qi,eic->eqc', tabs['phi'], a['coords'][a['cells']])
    f = source(xq.reshape(-1, 2)).reshape(mesh.n_cells, s['nIP'], 2)
    tau = np.zeros((mesh.n_faces, s['nNf'], 2, 2, 2))
    tau[:] = np.eye(2) * (2.0 + DIF * n)
    g = exact(a['coords'][a['faces'].reshape(-1)]).reshape(mesh.n_faces, s['nNf'], 2)
    nn = mesh.n_nodes
    prob = hf.NGammaProblem(p)
    prob.set_mesh(mesh)
    prob.set_fields(tau, D=np.full(nn, DIF), G=np.full(nn, GDIF), b=np.tile(bdir, (nn, 1)), sound_speed=CS,
                    dirichlet=g, source=f)
    prob.set_krylov(1)
    init = np.tile([1.0, 0.0], (mesh.n_cells, s['nN']))
    prob.set_state(sol=init, buffer=init)
    prob.assemble()
    t = []
    for _ in range(5):
        t0 = time.perf_counter()
        prob.assemble()
        t.append(1e3 * (time.perf_counter() - t0))
    asm_ms = float(np.median(t))
    t0 = time.perf_counter()
    its, res, lin = prob.nonlinear_solve(max_iters=20, res_tol=1e-10, pc=4, rtol=1e-12, maxits=20000)
    nl_ms = 1e3 * (time.perf_counter() - t0)
    sol = prob.state()[0].reshape(mesh.n_cells, -1)
    ex = exact(a['coords'][a['cells'].reshape(-1)]).reshape(mesh.n_cells, -1)
    dofs = mesh.n_cells * s['nN'] * 2
    out = {'workload': '2-D nGamma (n, Gamma) p=%d, %d^2x2=%d triangles, manufactured' % (p, n, mesh.n_cells),
           'cells': mesh.n_cells, 'u_dofs': dofs, 'assemble_ms': asm_ms,
           'assemble_dof_per_s': dofs / (asm_ms * 1e-3), 'newton_iters': its, 'residual': res,
           'linear_iters': lin, 'nonlinear_solve_ms': nl_ms,
           'rel_err_vs_exact': float(np.linalg.norm(sol - ex) / np.linalg.norm(ex))}
    print('# ngamma: assemble %.2f ms, %d Newton iterations (%d linear) in %.1f ms, error %.2e' % (
        asm_ms, its, lin, nl_ms, out['rel_err_vs_exact']), file=sys.stderr, flush=True)
    return out


def copy_calibration(mib):
    """Copy `mib` MiB device-to-device 5 times with torch (one plain copy
    kernel per call): the PMC reading of FETCH_SIZE / WRITE_SIZE for these
    dispatches against the known 1 x mib read + 1 x mib write gives the scale
    factor applied to the element kernel's counters."""
    import torch
    n = mib * (1 << 20) // 8
    x = torch.arange(n, dtype=torch.float64, device='cuda')
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({'copy_calibration_MiB': mib, 'bytes_read': n * 8, 'bytes_written': n * 8, 'dispatches': 5}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--n', type=int, default=24)
    ap.add_argument('--p', type=int, default=3)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-solve', action='store_true')
    ap.add_argument('--cpu-n', type=int, default=24)
    ap.add_argument('--traffic', type=float, nargs=2, metavar=('FETCH_KIB', 'WRITE_KIB'), default=None,
                    help='HBM KiB per element-kernel dispatch from rocprofv3 --pmc runs, already scaled')
    ap.add_argument('--copy-calibration', type=int, default=0, metavar='MiB')
    ap.add_argument('--no-learned', action='store_true')
    ap.add_argument('--no-batch', action='store_true')
    ap.add_argument('--learned-n', type=int, default=256)
    ap.add_argument('--no-burgers', action='store_true')
    ap.add_argument('--no-2d', action='store_true', help='skip the 2-D diffusion-source configurations')
    args = ap.parse_args()
    if args.gpus != 1:
        print('bench.py: this build runs on exactly one GPU (multi-GPU sharding is designed, not built); '
              'use --gpus 1', file=sys.stderr)
        sys.exit(2)

    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return

    import hyperfox_amd as hf
    import hyperfox_amd.problems as pr

    t0 = time.time()
    w = pr.convection_diffusion_3d(args.n, args.p)
    prob = hf.HDGProblem(3, args.p, device=0)
    pr.setup_gpu(w, prob)
    setup_s = time.time() - t0
    sz = prob.sizes
    nEl = w.mesh.n_cells
    dofs = nEl * sz['nN']
    print('# setup %.1f s: %d cells, %d trace dofs' % (setup_s, nEl, w.mesh.n_trace), file=sys.stderr, flush=True)

    for _ in range(args.warmup):
        prob.assemble()
    t1 = time.perf_counter()
    for _ in range(args.steps):
        prob.assemble()
    wall = time.perf_counter() - t1
    ms_per_step = 1e3 * wall / args.steps
    value = dofs / (wall / args.steps)
    print('# assembly: %.3f ms/step' % ms_per_step, file=sys.stderr, flush=True)

    # dominant kernel: element assembly + condensation, HIP events on its stream
    k_ms = prob.time(0, warmup=1, reps=max(3, args.steps))
    k_avg = float(np.mean(k_ms))
    fl = elem_flops(3, args.p, sz['nN'], sz['nIP'], sz['nNf'], sz['nIPf'], True) * nEl
    nin = sz['nN'] * (3 + 1 + 3) + sz['nF'] * sz['nNf'] * 2 + sz['nIP']
    by = elem_bytes(3, sz['nN'], sz['nNf'], sz['nIP'], nin) * nEl
    achieved = fl / (k_avg * 1e-3) / 1e12
    steps_ms = prob.time(1, warmup=0, reps=max(3, args.steps))
    info = prob.kernel_info()
    kname = ('hdg_affine_kernel<3,%d,W=%d>' % (args.p, info['threads'] // 64) if info['kernel'] == 'affine'
             else 'hdg_element_kernel<3,%d,NT=%d>' % (args.p, info['threads']))
    if args.traffic is not None:
        traffic = 1024.0 * (args.traffic[0] + args.traffic[1])
    elif (args.n, args.p) in MEASURED_TRAFFIC_KIB:
        traffic = 1024.0 * sum(MEASURED_TRAFFIC_KIB[(args.n, args.p)])
    else:
        traffic = None

    res = {
        'metric': METRIC, 'value': value, 'unit': 'DOF/s', 'n_gpus': 1, 'steps': args.steps,
        'warmup': args.warmup, 'ms_per_step': ms_per_step, 'higher_is_better': True, 'scaling': 'weak',
        'vs_baseline': None, 'dtype': 'f64', 'data': 'synthetic',
        'config': {'workload': 'HDG 3D convection-diffusion p=%d, %d^3x6=%d jittered tets, D=1e-2' % (args.p, args.n, nEl),
                   'cells': nEl, 'p': args.p, 'u_dofs': dofs, 'trace_dofs': w.mesh.n_trace},
        'roofline': {'bound': 'mfma', 'achieved': achieved, 'peak': FP64_PEAK_TFLOPS, 'unit': 'TFLOP/s',
                     'frac': achieved / FP64_PEAK_TFLOPS, 'traffic': traffic,
                     'kernel': kname, 'kernel_ms': k_avg, 'threads_per_wg': info['threads'],
                     'lds_bytes_per_wg': info['lds_bytes'],
                     'algorithmic_flops_per_cell': fl / nEl, 'algorithmic_bytes_per_cell': by / nEl,
                     'hbm_GBs': by / (k_avg * 1e-3) / 1e9},
        'event_ms_per_step': float(np.median(steps_ms)),
    }
    if not args.no_solve:
        res['trace_solve'] = trace_solves(hf, pr, prob, w)
        res['transient'] = transient(hf, pr, prob, w)
        if not args.no_learned:
            res['learned_vs_plain'] = learned_vs_plain(hf, pr, args.learned_n)
    res['topology'] = topology_times(hf, w)
    if not args.no_burgers:
        res['burgers'] = burgers(hf)
        res['ngamma'] = ngamma(hf)
    if not args.no_batch:
        res['batched'] = batched(hf, pr)
    threads = min(16, os.cpu_count() or 1)
    if not args.no_2d:
        res['configs_2d'] = configs_2d(hf, pr, threads)
    if not args.no_cpu:
        res['cpu_baseline'] = cpu_baseline(args.cpu_n, args.p, threads)
    print(json.dumps(res))


if __name__ == '__main__':
    main()
