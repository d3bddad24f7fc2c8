"""Seeded synthetic workloads: integrals, the TUPS operator list and angles.

Integrals (orbital sites x_i = s*i + U(-0.2, 0.2)):
  h_pq = -exp(-|x_p - x_q|) for p != q,  h_pp = -1 + 0.1 N(0,1)
  L_k[p,q] = N(0, 0.3) exp(-|x_p - x_q| / 2), symmetric, k < 2n
  g_pqrs = sum_k L_k[p,q] L_k[r,s]            (8-fold symmetric)
g is laid out as the reference's _twoEInts.bin ([p][q][r][s] row-major,
src/hamiltonianmatrix.cpp:448-453); with 8-fold symmetry its index
convention is immaterial.

Spin layout: bits 0..n-1 spin-down, n..2n-1 spin-up (src/operatorpool.cpp:221-222).
The "Hartree-Fock" determinant occupies spatial orbitals 0, 2, 4, ... in both
spins, the alternating pattern of the reference's H-chain inputs
(Hams/H4/L1/H4_Initial.dat:1).

TUPS layer (one-based spin-orbital indices as in the reference's operator
files, Hams/H4/L1/H4_Operators.dat): units on pairs (p, p+1), first
p = 1, 3, 5, ... then p = 2, 4, ...; each unit is
  single  "p p+1 0 0" and "p+n p+1+n 0 0"  (shared angle)
  double  "p p+n p+1 p+1+n"
  single  "p p+1 0 0" and "p+n p+1+n 0 0"  (shared angle)
i.e. 3 parameters and 5 operators per unit, all with scale 1.
"""
import math

import numpy as np

SEED_BASE = 20260000


def seed_for(config, instance=0):
    return SEED_BASE + 100 * config + instance


def synthetic_integrals(n, seed, spacing=1.5, return_factors=False):
    rng = np.random.default_rng(seed)
    x = spacing * np.arange(n) + rng.uniform(-0.2, 0.2, size=n)
    d = np.abs(x[:, None] - x[None, :])
    h = -np.exp(-d)
    h[np.diag_indices(n)] = -1.0 + 0.1 * rng.standard_normal(n)
    nk = 2 * n
    L = np.zeros((nk, n, n))
    iu = np.triu_indices(n)
    for k in range(nk):
        vals = rng.normal(0.0, 0.3, size=len(iu[0])) * np.exp(-d[iu] / 2.0)
        L[k][iu] = vals
        L[k].T[iu] = vals
    g = np.einsum("kpq,krs->pqrs", L, L, optimize=True)
    if return_factors:
        return h, g, L
    return h, g


def hf_index(n):
    """Zero-based full-space index of the alternating determinant."""
    idx = 0
    for i in range(0, n, 2):
        idx |= (1 << i) | (1 << (i + n))
    return idx


def tups_operators(n, layers):
    """Returns (ops (M,4) zero-based, -1 absent; order (M,) zero-based; scale)."""
    ops, order = [], []
    par = 0
    for _ in range(layers):
        for start in (1, 2):
            for p in range(start, n, 2):
                q = p + 1
                single = [(p, q, 0, 0), (p + n, q + n, 0, 0)]
                double = [(p, p + n, q, q + n)]
                for grp in (single, double, single):
                    for o in grp:
                        ops.append([x - 1 for x in o])
                        order.append(par)
                    par += 1
    ops = np.array(ops, dtype=np.int32)
    return ops, np.array(order, dtype=np.int32), np.ones(len(ops))


def seeded_angles(nparams, seed, count=None):
    rng = np.random.default_rng(seed)
    if count is None:
        return rng.uniform(-math.pi, math.pi, size=nparams)
    return rng.uniform(-math.pi, math.pi, size=(count, nparams))


def sector_dim(n, nup, ndown):
    return math.comb(n, nup) * math.comb(n, ndown)


# The workloads: n spatial orbitals at half filling, TUPS with `layers` layers.
CONFIGS = {
    1: dict(n=4, layers=2),
    2: dict(n=10, layers=5),
    3: dict(n=12, layers=6),
    4: dict(n=14, layers=7),
}


def make_workload(config, instance=0):
    c = CONFIGS[config]
    n, layers = c["n"], c["layers"]
    seed = seed_for(config, instance)
    h, g = synthetic_integrals(n, seed)
    ops, order, scale = tups_operators(n, layers)
    P = int(order.max()) + 1
    theta = seeded_angles(P, seed + 1)
    return dict(n=n, nq=2 * n, layers=layers, h=h, g=g, ops=ops, order=order, scale=scale,
                start_idx=hf_index(n), theta=theta, nup=n // 2, ndown=n // 2,
                N=sector_dim(n, n // 2, n // 2), P=P, M=len(ops), seed=seed)
