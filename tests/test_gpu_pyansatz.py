"""The reference's Python module mirror on the device, used the way the
reference's Python users drive stateAnsatzManager: the H4 L1 set-up from
its files, the recorded minimum, the gradients per operator and per
parameter (against central differences and each other), the Hessian, the
batch, the optimiser and the set-up lock after construction."""
import os

import numpy as np
import pytest

from ansatzevolve_amd import AnsatzSynth
from ansatzevolve_amd import PyAnsatzEvolve as pae

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
H4 = os.path.join(HERE, "golden", "runfiles", "H4", "L1", "H4")


def _manager():
    m = pae.stateAnsatzManager()
    state = pae.readCsvState(H4 + "_Initial.dat")
    idx = np.flatnonzero(state)
    assert m.storeInitial(8, idx.tolist(), np.asarray(state)[idx].tolist())
    assert m.storeOperators(pae.loadOperators(H4 + "_Operators.dat"))
    paths, order, P = pae.loadParameters(H4 + "_Order.dat", H4 + "_Parameters.dat", pae.loadPath(H4 + "_Operators.dat"))
    assert m.storeParameterDependencies(order)
    assert m.storeRunPath(H4)
    return m, paths, order, P


def test_manager_minimum_and_derivatives(gpu):
    m, paths, order, P_file = _manager()
    # the manager counts parameters as the highest order entry + 1 (7 here);
    # loadParameters reports the larger parameter-file count (9)
    P = max(p for p, _ in order) + 1
    assert P_file == 9 and P == 7
    E_rec = -2.838345804245581  # H4_Parameters.dat, second set
    x = np.array([a for _, a in paths[2]])  # per operator
    assert abs(m.getExpectationValue(x) - E_rec) <= 1e-10
    th = m.getAngles()
    assert len(th) == P and abs(m.getExpectationValue(th) - E_rec) <= 1e-10
    assert not m.storeInitial(8, [0], [1.0])  # set-up is frozen after construction
    # per-operator gradient at a generic point: central differences, and its
    # compression (sum of scale x per-operator derivative) = the per-parameter one
    x1 = x + 0.1 * np.sin(np.arange(len(x)) + 1.0)
    g_op = m.getGradient(x1)
    h = 1e-5
    fd = np.array([(m.getExpectationValue(x1 + h * e) - m.getExpectationValue(x1 - h * e)) / (2 * h)
                   for e in np.eye(len(x1))])
    assert np.max(np.abs(g_op - fd)) < 1e-7
    th1 = np.array([x1[[k for k, (p, _) in enumerate(order) if p == q][0]] / [s for p, s in order if p == q][0]
                    for q in range(P)])
    xc = np.array([th1[p] * s for p, s in order])
    g_op = m.getGradient(xc)
    comp = np.zeros(P)
    for k, (p, s) in enumerate(order):
        comp[p] += s * g_op[k]
    assert np.max(np.abs(comp - m.getGradientComp(th1))) < 1e-12
    with AnsatzSynth() as a:
        a.load_run_files(H4)
        assert np.max(np.abs(m.getHessianComp(th1) - a.hessian(th1))) < 1e-12
        E = m.getExpectationValues(np.stack([th1, th]))
        assert np.max(np.abs(E - [a.energy(th1), a.energy(th)])) < 1e-12
    with pytest.raises(RuntimeError):
        m.getHessian(th1)
    psi = m.getFinalState(th1)
    assert len(psi) == 256 and abs(np.linalg.norm(psi) - 1.0) < 1e-12


def test_manager_optimise(gpu):
    m, paths, order, _ = _manager()
    P = max(p for p, _ in order) + 1
    x = np.array([a for _, a in paths[2]])
    assert m.setAngles(x)
    th0 = m.getAngles()
    assert m.setAngles(th0 + 0.05 * np.cos(np.arange(P)))
    assert m.optimise()
    assert abs(m.getExpectationValue() - (-2.838345804245581)) <= 1e-10
    assert np.linalg.norm(m.getGradientComp(m.getAngles())) < 1e-9
