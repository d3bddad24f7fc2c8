"""Subspace diagonalisation (the reference's doSubspaceDiagonalisation) and
the properties table (writeProperties) through the C interface, against
restatements on the oracle's states with numpy's dense linear algebra."""
import numpy as np
import pytest

from ansatzevolve_amd import from_workload
from ansatzevolve_amd.workloads import hf_index, seeded_angles, synthetic_integrals, tups_operators
from oracle.oracle import OracleEngine

pytestmark = pytest.mark.gpu


def _case(n, layers, seed):
    h, g = synthetic_integrals(n, seed)
    ops, order, scale = tups_operators(n, layers)
    w = dict(n=n, nq=2 * n, h=h, g=g, ops=ops, order=order, scale=scale, start_idx=hf_index(n))
    o = OracleEngine(2 * n, [hf_index(n)], [1.0], ops, order, scale)
    o.set_integrals(h, g, explicit=True)
    return w, o, int(order.max()) + 1


def _dense_h(o):
    N = len(o.states())
    return np.array([o.ham_apply(e) for e in np.eye(N)]).T


def _lowdin(H, S):
    s, V = np.linalg.eigh(S)
    keep = s >= 1e-14
    inv = np.where(keep, 1.0 / np.sqrt(np.where(keep, s, 1.0)), 0.0)
    X = (V * inv) @ V.T
    return np.linalg.eigvalsh(X @ H @ X), int((~keep).sum())


@pytest.mark.parametrize("n,layers", [(4, 2), (6, 2)])
def test_subspace_against_oracle(gpu, n, layers):
    w, o, P = _case(n, layers, 5100 + n)
    thetas = seeded_angles(P, 61 + n, count=5)
    states = np.array([o.evolve(t) for t in thetas])
    hs = np.array([o.ham_apply(s) for s in states])
    H0, S0 = states @ hs.T, states @ states.T
    H0 = 0.5 * (H0 + H0.T)
    ev0, nr0 = _lowdin(H0, S0)
    with from_workload(w) as a:
        ev, H, S, nr = a.subspace_diagonalise(thetas)
        assert np.max(np.abs(H - H0)) <= 1e-12 and np.max(np.abs(S - S0)) <= 1e-12
        assert nr == nr0 == 0
        assert np.max(np.abs(ev - ev0)) <= 1e-10
        # variational: no subspace eigenvalue below the sector's ground state
        assert ev.min() >= np.linalg.eigvalsh(_dense_h(o)).min() - 1e-10
        # one state: its energy; a repeated state: a dropped direction
        e1, _, _, nr1 = a.subspace_diagonalise(thetas[:1])
        assert abs(e1[0] - o.energy(thetas[0])) <= 1e-12 and nr1 == 0
        e2, _, _, nr2 = a.subspace_diagonalise(np.stack([thetas[1], thetas[1], thetas[2]]))
        assert nr2 == 1
        ev3, _ = _lowdin(H0[1:3, 1:3], S0[1:3, 1:3])
        # the dropped direction contributes a zero eigenvalue
        assert np.max(np.abs(np.sort(e2) - np.sort(np.concatenate([[0.0], ev3])))) <= 1e-9


@pytest.mark.parametrize("n,layers", [(4, 2), (6, 2)])
def test_properties_against_oracle(gpu, n, layers):
    w, o, P = _case(n, layers, 5200 + n)
    thetas = np.concatenate([np.zeros((1, P)), seeded_angles(P, 71 + n, count=2)])
    Hd = _dense_h(o)
    lam, vec = np.linalg.eigh(Hd)
    with from_workload(w) as a:
        rows, Hs, Gs, low = a.properties(thetas, noverlaps=2)
        assert np.max(np.abs(low - lam[:2])) <= 1e-9
        for th, r, Hm, Gm in zip(thetas, rows, Hs, Gs):
            E, g = a.energy_gradient(th)
            assert abs(r["energy"] - E) <= 1e-12 and abs(r["energy_with_nuclear"] - E) <= 1e-12
            assert abs(r["grad_norm"] - np.linalg.norm(g)) <= 1e-12
            assert np.max(np.abs(Hm - a.hessian(th))) <= 1e-12
            # metric against the oracle's tangents by central differences
            J = np.array([(o.evolve(th + 1e-5 * e) - o.evolve(th - 1e-5 * e)) / 2e-5 for e in np.eye(P)])
            assert np.max(np.abs(Gm - J @ J.T)) <= 1e-8
            he, ge = np.linalg.eigvalsh(Hm), np.linalg.eigvalsh(Gm)

            def clear(v):  # no value so close to the 1e-10 threshold that rounding decides its class
                return not np.any((np.abs(v) > 1e-11) & (np.abs(v) < 1e-9))
            if clear(he):
                assert (r["hess_neg"], r["hess_zero"], r["hess_pos"]) == \
                       ((he < -1e-10).sum(), (np.abs(he) <= 1e-10).sum(), (he > 1e-10).sum())
            assert r["hess_neg"] + r["hess_zero"] + r["hess_pos"] == P
            d = np.diag(Hm)
            if clear(d):
                assert (r["hess_diag_neg"], r["hess_diag_zero"], r["hess_diag_pos"]) == \
                       ((d < -1e-10).sum(), (np.abs(d) <= 1e-10).sum(), (d > 1e-10).sum())
            if clear(ge):
                assert (r["metric_pos"], r["metric_zero"]) == ((ge > 1e-10).sum(), (ge <= 1e-10).sum())
            if clear(np.diag(Gm)):
                assert r["metric_diag_zero"] == (np.abs(np.diag(Gm)) <= 1e-10).sum()
            psi = o.evolve(th)
            for t in range(2):
                gap = min(abs(lam[t] - lam[u]) for u in range(len(lam)) if u != t)
                if gap > 1e-6:  # a non-degenerate eigenvector, unique up to sign
                    assert abs(abs(r["overlaps"][t]) - abs(psi @ vec[:, t])) <= 1e-8
        # the zero angles: the metric of the first half-layer's gauge
        # directions is singular (three angles per unit, two independent)
        assert rows[0]["metric_zero"] >= 1
        rows2, _, _, _ = a.properties(thetas, noverlaps=0, no_hessian=True)
        assert rows2[1]["hess_pos"] == 0 and abs(rows2[1]["energy"] - rows[1]["energy"]) <= 1e-12
