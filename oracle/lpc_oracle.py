"""FP64 numpy oracle of the learned trace preconditioner.

TEST INFRASTRUCTURE ONLY: the checker for the GPU forward pass, loss and
backward pass of hyperfox_amd/csrc/lpc.hip.  Only the tests load it.

It restates the network defined in hyperfox_amd/csrc/lpc.h (there is no
counterpart in the reference): diagonal normalisation of the block trace
matrix, block-Jacobi inverse, a three-layer message-passing network of width
64 emitting one correction block per matrix block, the probe loss
||z - Mhat Ahat z||^2 / ||z||^2, and its exact gradient by reverse-mode
differentiation written out by hand.  The gradient is pinned by central
finite differences in tests/test_lpc_oracle.py.
"""
import numpy as np

H = 64
LAYERS = 3


def layout(b):
    NI, NE, O = b * b + 1, b * b, b * b
    off, o = {}, 0
    for name, n in [('Wn', NI * H), ('bn', H), ('We', NE * H), ('be', H)]:
        off[name] = (o, n)
        o += n
    for l in range(LAYERS):
        for name, n in [('Wm%d' % l, 3 * H * H), ('bm%d' % l, H), ('Wu%d' % l, 2 * H * H), ('bu%d' % l, H)]:
            off[name] = (o, n)
            o += n
    for name, n in [('Wdn', H * O), ('bdn', O), ('Wde', H * O), ('bde', O)]:
        off[name] = (o, n)
        o += n
    off['total'] = o
    return off


def unpack(p, b):
    L = layout(b)
    NI, NE, O = b * b + 1, b * b, b * b
    shapes = {'Wn': (NI, H), 'We': (NE, H), 'Wdn': (H, O), 'Wde': (H, O)}
    for l in range(LAYERS):
        shapes['Wm%d' % l] = (3 * H, H)
        shapes['Wu%d' % l] = (2 * H, H)
    out = {}
    for k, (o, n) in ((k, v) for k, v in L.items() if k != 'total'):
        out[k] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shapes.get(k, (n,)))
    return out


def pack(d, b):
    L = layout(b)
    p = np.zeros(L['total'])
    for k, (o, n) in ((k, v) for k, v in L.items() if k != 'total'):
        p[o:o + n] = d[k].ravel()
    return p


class Graph:
    """Block pattern (rowptr, col over blocks, sorted per row), its edges
    (off-diagonal blocks in row order) and the Dirichlet rows."""

    def __init__(self, rowptr, col, dirichlet):
        self.rowptr = np.asarray(rowptr, dtype=np.int64)
        self.col = np.asarray(col, dtype=np.int64)
        self.nb = self.rowptr.size - 1
        self.rowf = np.repeat(np.arange(self.nb), np.diff(self.rowptr))
        self.diag = np.array([self.rowptr[F] + np.nonzero(self.col[self.rowptr[F]:self.rowptr[F + 1]] == F)[0][0]
                              for F in range(self.nb)])
        off = self.col != self.rowf
        self.eblk = np.nonzero(off)[0]
        self.esrc = self.rowf[self.eblk]
        self.edst = self.col[self.eblk]
        self.E = self.eblk.size
        self.blk2edge = -np.ones(self.col.size, dtype=np.int64)
        self.blk2edge[self.eblk] = np.arange(self.E)
        self.dir = np.asarray(dirichlet, dtype=bool)


def graph_from_scalar_csr(rowptr, cols, b, dirichlet):
    """Block pattern of a scalar CSR whose nonzero structure is made of
    dense b x b blocks (as the trace matrix is)."""
    nb = (len(rowptr) - 1) // b
    brp, bcol = [0], []
    for F in range(nb):
        r = F * b
        cs = np.asarray(cols[rowptr[r]:rowptr[r + 1]])
        g = np.unique(cs // b)
        bcol.extend(g.tolist())
        brp.append(len(bcol))
    return Graph(brp, bcol, dirichlet)


def blocks_from_scalar_csr(g, rowptr, cols, vals, b):
    A = np.zeros((g.col.size, b, b))
    for F in range(g.nb):
        for i in range(b):
            r = F * b + i
            cs = np.asarray(cols[rowptr[r]:rowptr[r + 1]])
            vs = np.asarray(vals[rowptr[r]:rowptr[r + 1]])
            for c, v in zip(cs, vs):
                G = c // b
                k = g.rowptr[F] + np.searchsorted(g.col[g.rowptr[F]:g.rowptr[F + 1]], G)
                A[k, i, c % b] = v
    return A


def normalise(g, A):
    """s_F = mean |diag(A_FF)|; Ahat = S^-1/2 A S^-1/2; dinv = inv(Ahat_FF)."""
    b = A.shape[1]
    d = np.abs(np.diagonal(A[g.diag], axis1=1, axis2=2)).mean(axis=1)
    sinv = np.where(d > 0, 1.0 / np.sqrt(np.where(d > 0, d, 1.0)), 1.0)
    ahat = A * (sinv[g.rowf] * sinv[g.col])[:, None, None]
    dinv = np.linalg.inv(ahat[g.diag])
    return sinv, ahat, dinv


def relu(x):
    return np.maximum(x, 0.0)


def forward(P, g, ahat):
    b = ahat.shape[1]
    O = b * b
    dirf = g.dir.astype(np.float64)
    Xn = np.concatenate([ahat[g.diag].reshape(g.nb, O), dirf[:, None]], axis=1)
    Xe = ahat[g.eblk].reshape(g.E, O)
    acts = dict(Xn=Xn, Xe=Xe, h=[relu(Xn @ P['Wn'] + P['bn'])], e=[relu(Xe @ P['We'] + P['be'])], m=[], a=[], u=[])
    for l in range(LAYERS):
        h, e = acts['h'][-1], acts['e'][-1]
        Xm = np.concatenate([h[g.esrc], h[g.edst], e], axis=1)
        m = relu(Xm @ P['Wm%d' % l] + P['bm%d' % l])
        a = np.zeros((g.nb, H))
        np.add.at(a, g.esrc, m)
        u = relu(np.concatenate([h, a], axis=1) @ P['Wu%d' % l] + P['bu%d' % l])
        acts['m'].append(m)
        acts['a'].append(a)
        acts['u'].append(u)
        acts['e'].append(e + m)
        acts['h'].append(h + u)
    cn = acts['h'][-1] @ P['Wdn'] + P['bdn']
    cn[g.dir] = 0.0
    ce = acts['e'][-1] @ P['Wde'] + P['bde']
    ce[g.dir[g.esrc]] = 0.0
    acts['cn'], acts['ce'] = cn, ce
    return acts


def mhat_blocks(g, dinv, cn, ce):
    b = dinv.shape[1]
    M = np.zeros((g.col.size, b, b))
    M[g.diag] = dinv + cn.reshape(-1, b, b)
    M[g.eblk] = ce.reshape(-1, b, b)
    return M


def bsr_mv(g, blocks, x):
    b = blocks.shape[1]
    xb = x.reshape(g.nb, b)
    y = np.zeros((g.nb, b))
    np.add.at(y, g.rowf, np.einsum('kij,kj->ki', blocks, xb[g.col]))
    return y.ravel()


def loss_and_dc(g, dinv, acts, probes):
    """probes: list of (z, scale); loss = sum scale * ||z - Mhat Ahat z||^2,
    with v = Ahat z supplied as the third tuple entry."""
    b = dinv.shape[1]
    M = mhat_blocks(g, dinv, acts['cn'], acts['ce'])
    loss = 0.0
    dM = np.zeros_like(M)
    for z, v, scale in probes:
        r = z - bsr_mv(g, M, v)
        loss += scale * np.dot(r, r)
        rb, vb = r.reshape(g.nb, b), v.reshape(g.nb, b)
        dM += -2.0 * scale * np.einsum('ki,kj->kij', rb[g.rowf], vb[g.col])
    dM[g.dir[g.rowf]] = 0.0
    dcn = dM[g.diag].reshape(g.nb, b * b)
    dce = dM[g.eblk].reshape(g.E, b * b)
    return loss, dcn, dce


def backward(P, g, acts, dcn, dce):
    G = {}
    hL, eL = acts['h'][-1], acts['e'][-1]
    G['Wdn'] = hL.T @ dcn
    G['bdn'] = dcn.sum(0)
    G['Wde'] = eL.T @ dce
    G['bde'] = dce.sum(0)
    dh = dcn @ P['Wdn'].T
    de = dce @ P['Wde'].T
    for l in reversed(range(LAYERS)):
        h, e, m, a, u = acts['h'][l], acts['e'][l], acts['m'][l], acts['a'][l], acts['u'][l]
        dPu = dh * (u > 0)
        Xu = np.concatenate([h, a], axis=1)
        G['Wu%d' % l] = Xu.T @ dPu
        G['bu%d' % l] = dPu.sum(0)
        dXu = dPu @ P['Wu%d' % l].T
        dh_new = dh + dXu[:, :H]
        da = dXu[:, H:]
        dm = de + da[g.esrc]
        dPm = dm * (m > 0)
        Xm = np.concatenate([h[g.esrc], h[g.edst], e], axis=1)
        G['Wm%d' % l] = Xm.T @ dPm
        G['bm%d' % l] = dPm.sum(0)
        dXm = dPm @ P['Wm%d' % l].T
        np.add.at(dh_new, g.esrc, dXm[:, :H])
        np.add.at(dh_new, g.edst, dXm[:, H:2 * H])
        de = de + dXm[:, 2 * H:]
        dh = dh_new
    dPe = de * (acts['e'][0] > 0)
    G['We'] = acts['Xe'].T @ dPe
    G['be'] = dPe.sum(0)
    dPh = dh * (acts['h'][0] > 0)
    G['Wn'] = acts['Xn'].T @ dPh
    G['bn'] = dPh.sum(0)
    return G


def rademacher(n, seed):
    """The GPU's counter-based Rademacher stream (splitmix64 finaliser)."""
    M = (1 << 64) - 1
    out = np.empty(n)
    for t in range(n):
        x = (seed + 0x9E3779B97F4A7C15 * (t + 1)) & M
        x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M
        x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M
        x ^= x >> 31
        out[t] = 1.0 if (x >> 63) else -1.0
    return out


def probe_seed(seed, q):
    """Seed of random probe q of a loss evaluation with `seed` (as on the GPU)."""
    M = (1 << 64) - 1
    return (seed * 1000003 + q * 7919 + 17) & M


def loss_grad(p, b, g, A, lam, n_random=0, seed=0):
    """Loss and flat gradient for one system (A blocks FP64, lam = its solution)."""
    P = unpack(p, b)
    sinv, ahat, dinv = normalise(g, A)
    z = lam / np.repeat(sinv, b)
    z = z / np.linalg.norm(z)
    probes = [(z, bsr_mv(g, ahat, z), 1.0)]
    n = g.nb * b
    for q in range(n_random):
        zr = rademacher(n, probe_seed(seed, q))
        probes.append((zr, bsr_mv(g, ahat, zr), 1.0 / n))
    acts = forward(P, g, ahat)
    loss, dcn, dce = loss_and_dc(g, dinv, acts, probes)
    G = backward(P, g, acts, dcn, dce)
    return loss, pack(G, b), acts
