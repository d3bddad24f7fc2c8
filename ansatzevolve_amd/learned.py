"""Learned warm start for the TUPS optimiser.

A 3-layer FP32 network (width 256, tanh) maps integral features to initial
angles.  Training and inference run on the GPU through the library's
mlp* entry points (hand-written HIP kernels); this module prepares
features and drives them.

Features of an n-orbital Hamiltonian, padded to nmax orbitals:
  * h, n x n;
  * the leading four factors of the two-electron supermatrix
    G[(pq),(rs)] = g_pqrs (symmetric positive semi-definite for the seeded
    integrals): eigenvectors of the four largest eigenvalues times the
    square root of the eigenvalue, each reshaped to n x n, sign fixed so that
    its largest-magnitude entry is positive.
Angles are predicted per parameter and padded to the largest parameter
count; a mask marks the real ones.
"""
import ctypes

import numpy as np

from . import load_library

_fp = ctypes.POINTER(ctypes.c_float)


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_fp)


def features(h, g, nmax, nfactors=4):
    n = h.shape[0]
    G = np.asarray(g).reshape(n * n, n * n)
    G = 0.5 * (G + G.T)
    w, V = np.linalg.eigh(G)
    order = np.argsort(w)[::-1][:nfactors]
    out = [np.pad(h, ((0, nmax - n), (0, nmax - n)))]
    for i in order:
        f = V[:, i] * np.sqrt(max(w[i], 0.0))
        if f[np.argmax(np.abs(f))] < 0:
            f = -f
        out.append(np.pad(f.reshape(n, n), ((0, nmax - n), (0, nmax - n))))
    return np.concatenate([o.reshape(-1) for o in out])


class WarmStartMLP:
    """The network on the GPU (mlpCreate / mlpForward / mlpGradient / mlpTrainStep)."""

    def __init__(self, nin, nout, hidden=256, seed=0):
        self.lib = load_library()
        self.h = self.lib.mlpCreate(nin, hidden, nout, seed)
        if not self.h:
            raise RuntimeError("mlpCreate failed")
        self.nin, self.nout, self.hidden = nin, nout, hidden

    def close(self):
        if self.h:
            self.lib.mlpDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def nparams(self):
        return self.lib.mlpParameterCount(self.h)

    def get_params(self):
        p = np.zeros(self.nparams(), dtype=np.float32)
        assert self.lib.mlpGetParameters(self.h, p.ctypes.data_as(_fp)) == 0
        return p

    def set_params(self, p):
        p, pp = _f(p)
        assert self.lib.mlpSetParameters(self.h, pp) == 0

    def unpack(self, p=None):
        """(W0, b0, W1, b1, W2, b2) as numpy arrays."""
        p = self.get_params() if p is None else p
        i, hd, o = self.nin, self.hidden, self.nout
        sizes = [(i, hd), (hd,), (hd, hd), (hd,), (hd, o), (o,)]
        out, off = [], 0
        for s in sizes:
            cnt = int(np.prod(s))
            out.append(p[off:off + cnt].reshape(s))
            off += cnt
        return out

    def forward(self, X):
        X, xp = _f(X)
        Y = np.zeros((X.shape[0], self.nout), dtype=np.float32)
        assert self.lib.mlpForward(self.h, X.shape[0], xp, Y.ctypes.data_as(_fp)) == 0
        return Y

    def gradient(self, X, T, M):
        X, xp = _f(X)
        T, tp = _f(T)
        M, mp = _f(M)
        loss = ctypes.c_float()
        g = np.zeros(self.nparams(), dtype=np.float32)
        assert self.lib.mlpGradient(self.h, X.shape[0], xp, tp, mp, ctypes.byref(loss), g.ctypes.data_as(_fp)) == 0
        return loss.value, g

    def train_step(self, X, T, M, lr):
        X, xp = _f(X)
        T, tp = _f(T)
        M, mp = _f(M)
        loss = ctypes.c_float()
        assert self.lib.mlpTrainStep(self.h, X.shape[0], xp, tp, mp, ctypes.c_float(lr), ctypes.byref(loss)) == 0
        return loss.value
