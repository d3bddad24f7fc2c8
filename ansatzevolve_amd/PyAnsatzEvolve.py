"""The reference's Python module (src/PyAnsatzEvolve.cpp:17-148, the
pybind11 binding of stateAnsatzManager, src/AnsatzManager.h) over this
library: the same class, method and function names, argument meanings and
return values, so that `import PyAnsatzEvolve` code runs after
`from ansatzevolve_amd import PyAnsatzEvolve`.

Indices follow the C++ methods the binding exposes: storeInitial and
storeHamiltonian take zero-based full-space indices, storeParameterDependencies
zero-based (parameter, scale) pairs, storeOperators and loadOperators the
one-based four-index lists of the operator files.  Set-up methods return
True or False like the reference's; the getters return numpy arrays.
"""
import os

import numpy as np

from . import AnsatzError, AnsatzSynth, load_library, read_parameter_sets

__all__ = ["stateAnsatzManager", "loadParameters", "loadPath", "loadNuclearEnergy", "readCsvState",
           "loadOperators"]


class stateAnsatzManager:
    """stateAnsatzManager (src/AnsatzManager.h:119-300).  The set-up is
    stored until the first evaluation, which constructs the device engine;
    later set-up calls fail, as in the reference."""

    def __init__(self):
        self._a = AnsatzSynth()
        self._ops = None      # one-based (M, 4)
        self._deps = None     # zero-based parameter, scale
        self._pushed = False
        self._angles = None   # the angles of the latest evolution, as given
        self._constructed = False

    def _ok(self, fn):
        try:
            fn()
            return True
        except AnsatzError:
            return False

    # ---- set-up (src/PyAnsatzEvolve.cpp:21-45)
    def storeOperators(self, operators):
        ops = [list(op) for op in operators]
        for op in ops:
            if len(op) != 4:
                raise ValueError("Operator must be 4 long")
            if max(op) - 1 > 127:
                raise OverflowError("Operator index too big, < 127")
        if self._constructed or not ops:
            return False
        self._ops = np.array(ops, dtype=np.int64)
        return self._push()

    def storeParameterDependencies(self, parameterDependency):
        deps = [(int(p), float(s)) for p, s in parameterDependency]
        if self._constructed or not deps:
            return False
        self._deps = deps
        return self._push()

    def _push(self):
        # the C ABI takes operators and their parameter map together
        if self._ops is None or self._deps is None or len(self._ops) != len(self._deps):
            return True
        order = np.array([p + 1 for p, _ in self._deps])
        scale = np.array([s for _, s in self._deps])
        self._pushed = self._ok(lambda: self._a.set_excitation(self._ops, order, scale))
        return self._pushed

    def storeInitial(self, numberOfQubits, indexes, coeffs):
        if self._constructed:
            return False
        return self._ok(lambda: self._a.set_initial_state(int(numberOfQubits), np.asarray(indexes) + 1,
                                                          np.asarray(coeffs, dtype=np.float64)))

    def storeHamiltonian(self, iIndexes, jIndexes, Coeffs):
        if self._constructed:
            return False
        return self._ok(lambda: self._a.set_hamiltonian(np.asarray(iIndexes) + 1, np.asarray(jIndexes) + 1,
                                                        np.asarray(Coeffs, dtype=np.float64)))

    def storeNuclearEnergy(self, NuclearEnergy):
        if self._constructed:
            return False
        return self._ok(lambda: self._a.set_nuclear_energy(float(NuclearEnergy)))

    def storeRunPath(self, runPath):
        if self._constructed:
            return False
        return self._ok(lambda: self._a.set_run_path(runPath))

    # ---- angles (src/AnsatzManager.cpp:467-523)
    def _dims(self):
        self._constructed = True
        return self._a.dimensions()

    def _use(self, angles):
        if angles is None:
            if self._angles is None:
                raise RuntimeError("no angles have been set")
            return self._angles
        return np.asarray(angles, dtype=np.float64)

    def setAngles(self, angles):
        a = np.asarray(angles, dtype=np.float64)
        try:
            _, M, P = self._dims()
            if len(a) not in (M, P):
                return False
            self._a.energy(a)  # evolves, as the reference does
        except AnsatzError:
            return False
        self._angles = a
        return True

    def getAngles(self):
        """The current angles per parameter: the first operator of each
        parameter's angle divided by its scale (the reference's normCompress)."""
        a = self._use(None)
        _, M, P = self._dims()
        if len(a) == P:
            return a.copy()
        out = np.zeros(P)
        seen = np.zeros(P, dtype=bool)
        for k, (p, s) in enumerate(self._deps):
            if not seen[p]:
                out[p] = a[k] / s if s != 0 else 0.0
                seen[p] = True
        return out

    # ---- evaluations (src/PyAnsatzEvolve.cpp:50-94)
    def getExpectationValue(self, angles=None):
        a = self._use(angles)
        self._dims()
        E = self._a.energy(a)
        self._angles = a
        return E

    def getFinalState(self, angles=None):
        a = self._use(angles)
        self._dims()
        nq = self._a.num_qubits
        psi = self._a.final_state(a, nq)
        self._angles = a
        return psi

    def getGradient(self, angles=None):
        """Per operator (decompressed)."""
        a = self._use(angles)
        self._dims()
        g = self._a.gradient_operators(a)
        self._angles = a
        return g

    def getGradientComp(self, angles=None):
        """Per parameter (compressed)."""
        a = self._use(angles)
        _, M, P = self._dims()
        if len(a) == M:
            a = self._compress(a)
        g = self._a.gradient(a)
        self._angles = a
        return g

    def getHessian(self, angles=None):
        # the reference's fused build stops here (src/AnsatzManager.h:229-236: "DO NOT USE,
        # use getHessianComp(); crashes unless compiled without USE_FUSED")
        raise RuntimeError("getHessian (per operator) is not available; use getHessianComp")

    def getHessianComp(self, angles=None):
        a = self._use(angles)
        _, M, P = self._dims()
        if len(a) == M:
            a = self._compress(a)
        H = self._a.hessian(a)
        self._angles = a
        return H

    def _compress(self, a):
        self._angles = a
        return self.getAngles()

    def getExpectationValues(self, angles):
        """Energies of the rows of `angles` (per operator or per parameter)."""
        A = np.atleast_2d(np.asarray(angles, dtype=np.float64))
        self._dims()
        return self._a.energies(A)

    def optimise(self):
        """The plain Newton optimiser from the current angles; the result is
        read with getExpectationValue() and getAngles() (src/AnsatzManager.cpp:756-773)."""
        a = self._use(None)
        self._dims()
        try:
            th, E, _ = self._a.optimise(a)
        except AnsatzError:
            return False
        self._angles = th
        return True


# ---- module functions (src/PyAnsatzEvolve.cpp:97-147)
def loadParameters(orderFilePath, parameterFilepath, templatePath):
    """(rotationPaths, order, numberOfUniqueParameters): path 0 is the
    template, path i + 1 the template with the angles of parameter set i
    (parameter value times scale per operator); order holds zero-based
    (parameter, scale) pairs (src/TUPSLoadingUtils.cpp:176-296)."""
    order = []
    for line in open(orderFilePath):
        line = line.strip()
        if not line:
            continue
        parts = [x for x in line.replace(",", " ").split()]
        order.append((int(parts[0]) - 1, float(parts[1]) if len(parts) > 1 else 1.0))
    if os.path.exists(parameterFilepath):
        _, sets = read_parameter_sets(parameterFilepath)
        count = sets.shape[1] if len(sets) else 0
    else:
        sets, count = np.zeros((0, 0)), 0
    template = [tuple(r) for r in templatePath]
    paths = [list(template)]
    for s in sets:
        path = list(template)
        for k, (p, sc) in enumerate(order):
            if k < len(path) and p < len(s):
                path[k] = (path[k][0], float(s[p] * sc))
        paths.append(path)
    nparams = max(max((p for p, _ in order), default=-1) + 1, count)
    return paths, order, nparams


def loadPath(filepath):
    """One (0, 0.0) rotation element per operator line of an _Operators.dat file."""
    return [(0, 0.0) for _ in loadOperators(filepath)]


def loadNuclearEnergy(filepath):
    """<filepath>_Nuclear_Energy.dat, 0 when missing (src/TUPSLoadingUtils.cpp:156-174)."""
    try:
        return float(open(filepath + "_Nuclear_Energy.dat").read().split()[0])
    except (OSError, IndexError, ValueError):
        return 0.0


def readCsvState(filepath):
    """Dense coefficients over 2^(bitstring length) basis states from an
    _Initial.dat file (rightmost character = bit 0)."""
    coeffs = []
    for line in open(filepath):
        line = line.strip()
        if not line:
            continue
        parts = [x.strip() for x in line.split(",")]
        bits = parts[0]
        idx = int(bits, 2)
        c = float(parts[1]) if len(parts) > 1 and parts[1] else 1.0
        if len(coeffs) < (1 << len(bits)):
            coeffs.extend([0.0] * ((1 << len(bits)) - len(coeffs)))
        coeffs[idx] = c
    return coeffs


def loadOperators(filepath):
    """The one-based four-index lists of an _Operators.dat file."""
    out = []
    for line in open(filepath):
        parts = line.split()
        if len(parts) == 4:
            out.append([int(x) for x in parts])
    return out
