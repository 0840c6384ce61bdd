"""Generate trace systems, train the learned preconditioner on the GPU, and
compare it with block Jacobi on held-out systems.

Workload: 2-D HDG diffusion on the unit square (n x n squares split in two,
vertices jittered +-0.1h, mesh seed 2024), p = 3, coefficient k = exp(g) with
g a squared-exponential Gaussian random field (correlation length 0.1,
sigma 1), source 1, u = 0 on the boundary.  Every sample is assembled and
solved (BiCGStab + block Jacobi, true residual 1e-10) by this library; its
trace matrix and solution stay on the GPU as a training sample.  Held-out
systems are solved with BiCGStab and both preconditioners to the same true
residual.

Usage:
  python -m hyperfox_amd.train_lpc --out W.npy [--summary S.json]
      [--train 64:1400,128:500,256:100] [--heldout 64:100,128:60,256:20]
      [--epochs 8] [--decay 0.7] [--lr 1e-3] [--nrand 2]
Training seeds of size n are n * 100000 + i; held-out seeds are
900000 + n * 1000 + i, outside every training range.
"""
import argparse
import json
import sys
import time

import numpy as np

import hyperfox_amd as hf
import hyperfox_amd.problems as pr
from hyperfox_amd.lpc import LearnedPreconditioner, pc_setup_ms

P = 3
MESH_SEED = 2024


def parse_mix(s):
    out = []
    for part in s.split(','):
        if part:
            n, c = part.split(':')
            out.append((int(n), int(c)))
    return out


class Problems:
    """One mesh and one HDG problem handle per size, reused for every seed."""

    def __init__(self):
        self.by_n = {}

    def get(self, n):
        if n not in self.by_n:
            mesh = hf.Mesh.structured(2, P, n, 0.1, MESH_SEED)
            prob = hf.HDGProblem(2, P)
            w = pr.diffusion_grf_2d(n, P, 0, mesh=mesh)
            pr.setup_gpu(w, prob)
            prob.set_krylov(hf.KRYLOV_BICGSTAB)
            self.by_n[n] = (mesh, prob)
        return self.by_n[n]

    def system(self, n, seed):
        """Assemble the system of GRF `seed` on the size-n mesh."""
        mesh, prob = self.get(n)
        w = pr.diffusion_grf_2d(n, P, seed, mesh=mesh)
        prob.set_fields(w.tau, w.tau_sides, w.D, w.ndD, w.b, w.dirichlet)
        prob.assemble()
        return prob


def log(msg):
    print(msg, flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--out', required=True)
    ap.add_argument('--summary', default=None)
    ap.add_argument('--train', default='64:1400,128:500,256:100')
    ap.add_argument('--heldout', default='64:100,128:60,256:20')
    ap.add_argument('--epochs', type=int, default=8)
    ap.add_argument('--decay', type=float, default=0.7)
    ap.add_argument('--lr', type=float, default=1e-3)
    ap.add_argument('--nrand', type=int, default=2)
    ap.add_argument('--seed', type=int, default=1)
    args = ap.parse_args(argv)

    probs = Problems()
    net = LearnedPreconditioner(4)
    net.init_params(args.seed)
    summary = {'train': {}, 'heldout': {}, 'epochs': []}

    t0 = time.time()
    for n, count in parse_mix(args.train):
        its = []
        for i in range(count):
            prob = probs.system(n, n * 100000 + i)
            it, reason, rel = prob.solve(pc=hf.PC_BJACOBI, rtol=1e-10, maxits=50000)
            if reason != 2:
                raise RuntimeError('training system %d/%d did not converge (reason %d)' % (n, i, reason))
            net.add_sample(prob)
            its.append(it)
            if (i + 1) % 100 == 0 or i + 1 == count:
                log('# data n=%d: %d/%d systems, block-Jacobi its mean %.1f, %.0f s' % (n, i + 1, count, np.mean(its),
                                                                                       time.time() - t0))
        summary['train'][str(n)] = {'systems': count, 'bjacobi_its_mean': float(np.mean(its))}
    gen_s = time.time() - t0

    ns = net.num_samples()
    rng = np.random.default_rng(args.seed)
    t1 = time.time()
    step = 0
    for ep in range(args.epochs):
        order = rng.permutation(ns).astype(np.int32)
        lr = args.lr * (args.decay ** ep)
        losses = []
        for c0 in range(0, ns, 250):
            chunk = order[c0:c0 + 250]
            losses.append(net.train(chunk, lr=lr, n_random=args.nrand, seed=1000003 * (ep + 1) + c0))
            step += chunk.size
        losses = np.concatenate(losses)
        log('# epoch %d: lr %.1e, mean loss %.4e (first 100 %.4e, last 100 %.4e), %.0f s' % (
            ep, lr, losses.mean(), losses[:100].mean(), losses[-100:].mean(), time.time() - t1))
        summary['epochs'].append({'lr': lr, 'mean_loss': float(losses.mean())})
    train_s = time.time() - t1
    net.save(args.out)

    for n, count in parse_mix(args.heldout):
        rows = []
        for i in range(count):
            prob = probs.system(n, 900000 + n * 1000 + i)
            tb = time.perf_counter()
            ib, rb, relb = prob.solve(pc=hf.PC_BJACOBI, rtol=1e-10, maxits=50000)
            tb = time.perf_counter() - tb
            net.attach(prob, bf16=True)
            tl = time.perf_counter()
            il, rl, rell = prob.solve(pc=hf.PC_LEARNED, rtol=1e-10, maxits=50000)
            tl = time.perf_counter() - tl
            rows.append((ib, il, tb, tl, pc_setup_ms(prob), rb, rl, relb, rell))
        r = np.array(rows, dtype=float)
        summary['heldout'][str(n)] = {
            'systems': count, 'bjacobi_its_mean': float(r[:, 0].mean()), 'learned_its_mean': float(r[:, 1].mean()),
            'its_ratio_mean': float((r[:, 0] / r[:, 1]).mean()), 'its_ratio_min': float((r[:, 0] / r[:, 1]).min()),
            'bjacobi_ms_mean': 1e3 * float(r[:, 2].mean()), 'learned_ms_mean': 1e3 * float(r[:, 3].mean()),
            'learned_setup_ms_mean': float(r[:, 4].mean()), 'all_converged': bool(np.all(r[:, 5:7] == 2)),
            'max_true_relres': float(r[:, 7:9].max())}
        log('# held-out n=%d: %s' % (n, json.dumps(summary['heldout'][str(n)])))
    summary.update({'generation_s': gen_s, 'training_s': train_s, 'steps': step, 'params': net.n_params,
                    'lr': args.lr, 'nrand': args.nrand})
    if args.summary:
        with open(args.summary, 'w') as fh:
            json.dump(summary, fh, indent=1)
    log(json.dumps(summary))


if __name__ == '__main__':
    sys.exit(main())
