"""Learned trace preconditioner: Python handle over the hfx_lpc_* C ABI.

The network, its loss and its parameter layout are defined in
hyperfox_amd/csrc/lpc.h; training, inference and the preconditioned solve
all run in libhfx_hip.so on the GPU.  There is no counterpart in the
reference (its trace solve is PetscInterface::solve with PCJACOBI,
src/resolution/PetscInterface.cpp:222-247).
"""
import ctypes

import numpy as np

from ._lib import check, lib

H = 64
LAYERS = 3


def layout(b):
    """Offsets of the flat FP32 parameter vector (lpc.h LpcLayout)."""
    NI, NE, O = b * b + 1, b * b, b * b
    off = {}
    o = 0

    def take(name, n):
        nonlocal o
        off[name] = (o, n)
        o += n
    take('Wn', NI * H); take('bn', H)
    take('We', NE * H); take('be', H)
    for l in range(LAYERS):
        take('Wm%d' % l, 3 * H * H); take('bm%d' % l, H)
        take('Wu%d' % l, 2 * H * H); take('bu%d' % l, H)
    take('Wdn', H * O); take('bdn', O)
    take('Wde', H * O); take('bde', O)
    off['total'] = o
    return off


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class LearnedPreconditioner:
    def __init__(self, block_size, device=0):
        self.b = block_size
        h = ctypes.c_void_p()
        check(lib.hfx_lpc_create(device, block_size, ctypes.byref(h)))
        self._h = h
        n = ctypes.c_longlong()
        check(lib.hfx_lpc_num_params(self._h, ctypes.byref(n)))
        self.n_params = n.value

    def __del__(self):
        if getattr(self, '_h', None):
            lib.hfx_lpc_free(self._h)
            self._h = None

    def init_params(self, seed=0):
        check(lib.hfx_lpc_init_params(self._h, ctypes.c_ulonglong(seed)))

    def params(self):
        p = np.zeros(self.n_params, dtype=np.float32)
        check(lib.hfx_lpc_get_params(self._h, _p(p), self.n_params))
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.n_params:
            raise ValueError('expected %d parameters, got %d' % (self.n_params, p.size))
        check(lib.hfx_lpc_set_params(self._h, _p(p), self.n_params))

    def save(self, path):
        np.save(path, self.params())

    def load(self, path):
        self.set_params(np.load(path, allow_pickle=False))

    def add_sample(self, prob):
        check(lib.hfx_lpc_add_sample(self._h, prob._h))

    def num_samples(self):
        n = ctypes.c_int()
        check(lib.hfx_lpc_num_samples(self._h, ctypes.byref(n)))
        return n.value

    def clear_samples(self):
        check(lib.hfx_lpc_clear_samples(self._h))

    def graph_sizes(self, sample=0):
        out = np.zeros(4, dtype=np.int64)
        check(lib.hfx_lpc_graph_sizes(self._h, sample, _p(out)))
        return dict(nodes=int(out[0]), blocks=int(out[1]), edges=int(out[2]), bs=int(out[3]))

    def loss_grad(self, sample, n_random=0, seed=0, grad=True):
        loss = np.zeros(1)
        g = np.zeros(self.n_params, dtype=np.float32) if grad else None
        check(lib.hfx_lpc_loss_grad(self._h, sample, n_random, ctypes.c_ulonglong(seed), _p(loss),
                                    None if g is None else _p(g)))
        return float(loss[0]), g

    def train(self, order, lr=1e-3, n_random=1, seed=0):
        order = np.ascontiguousarray(order, dtype=np.int32)
        losses = np.zeros(max(order.size, 1))
        check(lib.hfx_lpc_train(self._h, _p(order), order.size, ctypes.c_float(lr), n_random,
                                ctypes.c_ulonglong(seed), _p(losses)))
        return losses[:order.size]

    def outputs(self, sample, bf16=False):
        sz = self.graph_sizes(sample)
        O = self.b * self.b
        cn = np.zeros((sz['nodes'], O), dtype=np.float32)
        ce = np.zeros((max(sz['edges'], 1), O), dtype=np.float32)
        check(lib.hfx_lpc_outputs(self._h, sample, 1 if bf16 else 0, _p(cn), _p(ce)))
        return cn, ce[:sz['edges']]

    def attach(self, prob, bf16=True):
        """Use this network for PC_LEARNED solves of `prob`."""
        check(lib.hfx_hdg_set_lpc(prob._h, self._h, 1 if bf16 else 0))
        prob._lpc = self  # keep the network alive while attached


def pc_setup_ms(prob):
    v = ctypes.c_double()
    check(lib.hfx_hdg_pc_setup_ms(prob._h, ctypes.byref(v)))
    return v.value
