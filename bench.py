"""Benchmark: HDG assembly + static condensation + trace scatter on one MI355X.

Workload: 3-D HDG convection-diffusion, unit cube split into 24^3 x 6 = 82,944
jittered tetrahedra, p = 3, D = 1e-2, b = (sin 2pi y, sin 2pi z, sin 2pi x),
manufactured u = e^{x+y+z} sin(pi x) sin(pi y) sin(pi z), upwind
double-valued tau.  One step = one pass of [element assembly + condensation +
Dirichlet rows + BSR scatter] over all elements, inputs resident in HBM.
value = u-DOFs (cells x nodes per cell) per second over the K timed steps.
The trace solve (FGMRES + block Jacobi to a true FP64 relative residual of
1e-10) is timed once after the steps and reported beside the main number.

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
MEASURED_TRAFFIC_KIB = {(24, 3): (2 * 105748.1, 3191374.5)}


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


def cpu_baseline(n_cpu, p, threads):
    """Oracle (FP64 C++ restatement of the reference path) on a smaller mesh of
    the same workload: per-element cost does not depend on the mesh size."""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import oracle as orc
    import hyperfox_amd.problems as pr
    w = pr.convection_diffusion_3d(n_cpu, p)
    a = w.mesh.arrays()
    X = a['coords'][a['cells']]
    s0 = orc.sizes(3, p)
    # source at IPs on the CPU (same formula as the GPU leg)
    t = orc.refel(3, p)
    xip = np.einsum('qn,end->eqd', t['phi'], X)
    f = w.source_at(xip)
    s = orc.System(3, p, 2, a, w.tau, tau_sides=2, D=w.D, ndD=1, b=w.b, f_ip=f, dirichlet=w.dirichlet,
                   nthreads=threads)
    best = min(s.time_assembly(w.mesh.n_cells, threads) for _ in range(2))
    dofs = w.mesh.n_cells * s0['nN']
    return {'value': dofs / best, 'unit': 'DOF/s', 'cores': threads, 'kind': 'port',
            'sample': '%d^3x6 = %d tets, same p/model/fields; assembly+condensation+BC+CSR scatter, %.2f s'
                      % (n_cpu, w.mesh.n_cells, best)}


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
    ap.add_argument('--cpu-n', type=int, default=12)
    ap.add_argument('--traffic', type=float, nargs=2, metavar=('FETCH_KIB', 'WRITE_KIB'), default=None,
                    help='HBM KiB per element-kernel dispatch from rocprofv3 --pmc runs, already scaled')
    ap.add_argument('--copy-calibration', type=int, default=0, metavar='MiB')
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
        t2 = time.perf_counter()
        its, reason, relres = prob.solve(pc=hf.PC_BJACOBI, rtol=1e-10, maxits=20000, restart=30)
        solve_s = time.perf_counter() - t2
        res['trace_solve'] = {'pc': 'block-jacobi', 'method': 'FGMRES(30) FP64', 'iters': its, 'reason': reason,
                              'true_relres': relres, 'ms': solve_s * 1e3}
        u, q, lam = prob.solution()
        a = w.mesh.arrays()
        ue = pr.u_c3(a['coords'][a['cells']])
        res['trace_solve']['nodal_rel_err_vs_exact'] = float(np.linalg.norm(u - ue) / np.linalg.norm(ue))
        print('# solve: %d its, relres %.2e, %.1f ms' % (its, relres, solve_s * 1e3), file=sys.stderr, flush=True)
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        res['cpu_baseline'] = cpu_baseline(args.cpu_n, args.p, threads)
    print(json.dumps(res))


if __name__ == '__main__':
    main()
