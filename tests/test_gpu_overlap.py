"""The projected derivative (getOverlapGradient, the reference's
FusedEvolve::evolveDerivativeProj): overlap <p|psi(theta)> and its
per-parameter gradient by the adjoint sweep with p in place of H psi,
against the final state's dot product, central differences, and the
closed form at p = psi(theta_0) (overlap 1, gradient 0 for a real
normalised state)."""
import numpy as np
import pytest

from ansatzevolve_amd import AnsatzError, from_workload
from ansatzevolve_amd.workloads import hf_index, seeded_angles, synthetic_integrals, tups_operators

pytestmark = pytest.mark.gpu


def _case(n, layers, seed):
    h, g = synthetic_integrals(n, seed)
    ops, order, scale = tups_operators(n, layers)
    return dict(n=n, nq=2 * n, h=h, g=g, ops=ops, order=order, scale=scale, start_idx=hf_index(n))


@pytest.mark.parametrize("n,layers", [(4, 2), (6, 2)])
def test_overlap_gradient(gpu, n, layers):
    w = _case(n, layers, 3100 + n)
    P = int(w["order"].max()) + 1
    th0 = seeded_angles(P, 77 + n)
    rng = np.random.default_rng(5 + n)
    with from_workload(w) as a:
        psi0 = a.final_state(th0, w["nq"])
        ov, g = a.overlap_gradient(th0, psi0)
        assert abs(ov - 1.0) <= 1e-12 and np.max(np.abs(g)) <= 1e-12
        # a generic projection vector, nonzero also outside the sector
        p = rng.standard_normal(1 << w["nq"])
        th = th0 + 0.3 * rng.standard_normal(P)
        ov, g = a.overlap_gradient(th, p)
        assert abs(ov - p @ a.final_state(th, w["nq"])) <= 1e-12
        h = 1e-5
        fd = np.array([(p @ a.final_state(th + h * e, w["nq"]) - p @ a.final_state(th - h * e, w["nq"])) / (2 * h)
                       for e in np.eye(P)])
        assert np.max(np.abs(g - fd)) <= 1e-7 * max(1.0, np.max(np.abs(fd)))
        # the call leaves psi(th) as the current state: energy and gradient there agree with a fresh context
        E1, g1 = a.energy_gradient(th)
    with from_workload(w) as b:
        E2, g2 = b.energy_gradient(th)
    assert abs(E1 - E2) <= 1e-12 and np.max(np.abs(g1 - g2)) <= 1e-12


def test_overlap_gradient_errors(gpu):
    w = _case(4, 1, 3200)
    P = int(w["order"].max()) + 1
    with from_workload(w) as a:
        with pytest.raises(AnsatzError) as e:
            a.overlap_gradient(np.zeros(P + 1), np.zeros(1 << w["nq"]))
        assert e.value.code == -6
        with pytest.raises(AnsatzError) as e:
            a.overlap_gradient(np.zeros(P), np.zeros(7))
        assert e.value.code == 5
