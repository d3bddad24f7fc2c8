"""MI355X statevector engine for AnsatzEvolve-style TUPS ansatze.

The product is ``libansatzsynth.so`` (C ABI in include/ansatzsynth.h, a
drop-in for the reference's AnsatzSynthInterface).  This module is the
Python host mirror of that interface: ``AnsatzSynth`` wraps one ctx and
calls the C entry points by their reference names.  There is no CPU
fallback: if the library is missing, importing the bindings fails.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libansatzsynth.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p

# (name, restype, argtypes) of every C entry point in include/ansatzsynth.h
C_API = [
    ("cleanup", ctypes.c_int, [_vp]),
    ("getEnergy", ctypes.c_int, [ctypes.c_int, _dp, _dp, _vp]),
    ("getFinalState", ctypes.c_int, [ctypes.c_int, _dp, ctypes.c_int, _dp, _vp]),
    ("getFinalStateComplex", ctypes.c_int, [ctypes.c_int, _dp, ctypes.c_int, _vp, _vp]),
    ("getGradient_COMP", ctypes.c_int, [ctypes.c_int, _dp, _dp, _vp]),
    ("getHessian_COMP", ctypes.c_int, [ctypes.c_int, _dp, _dp, _vp]),
    ("init", _vp, []),
    ("setExcitation", ctypes.c_int, [ctypes.c_int, _ip, _ip, _vp]),
    ("setExcitationScale", ctypes.c_int, [ctypes.c_int, _ip, _ip, _dp, _vp]),
    ("setHamiltonian", ctypes.c_int, [ctypes.c_int, _ip, _ip, _dp, _vp]),
    ("setInitialState", ctypes.c_int, [ctypes.c_int, ctypes.c_int, _ip, _dp, _vp]),
    ("setInitialStateComplex", ctypes.c_int, [ctypes.c_int, ctypes.c_int, _ip, _vp, _vp]),
    ("setTraceInterfaceCalls", None, [ctypes.c_int]),
    # extensions
    ("setHamiltonianIntegrals", ctypes.c_int, [ctypes.c_int, _dp, _dp, _vp]),
    ("getGradientOperators", ctypes.c_int, [ctypes.c_int, _dp, _dp, _vp]),
    ("getOverlapGradient", ctypes.c_int, [ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp, _vp]),
    ("getHessianProj", ctypes.c_int, [ctypes.c_int, _dp, ctypes.c_int, _dp, _dp, _dp, _dp, _vp]),
    ("setNuclearEnergy", ctypes.c_int, [ctypes.c_double, _vp]),
    ("setHamiltonianStorage", ctypes.c_int, [ctypes.c_int, _vp]),
    ("getEnergyGradient", ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, _vp]),
    ("getEnergies", ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp, _vp]),
    ("getEnergyGradients", ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _vp]),
    ("getDimensions", ctypes.c_int, [ctypes.POINTER(ctypes.c_longlong), _ip, _ip, _vp]),
    ("uploadAngleBatch", ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _vp]),
    ("runAngleBatch", ctypes.c_int, [ctypes.c_int, _vp]),
    ("downloadAngleBatch", ctypes.c_int, [_dp, _dp, _vp]),
    ("setKernelTiming", ctypes.c_int, [ctypes.c_int, _vp]),
    ("getKernelTiming", ctypes.c_int, [ctypes.c_int, _dp, ctypes.POINTER(ctypes.c_longlong), _vp]),
    ("optimise", ctypes.c_int, [ctypes.c_int, _dp, _dp, ctypes.c_double, ctypes.c_int, _dp, _ip, _vp]),
    ("iterativeOptimise", ctypes.c_int, [ctypes.c_int, _dp, _dp, ctypes.c_int, ctypes.c_int, _dp, _ip, _vp]),
    ("subspaceDiagonalise", ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _ip, _vp]),
    ("getEnergyGradientNoAlloc", ctypes.c_int, [ctypes.c_int, _dp, _dp, _dp, _vp]),
    ("getProperties", ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                      _dp, _vp]),
    ("multiBatchCreate", _vp, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    ("multiBatchUpload", ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]),
    ("multiBatchRun", ctypes.c_int, [_vp, ctypes.c_int]),
    ("multiBatchDownload", ctypes.c_int, [_vp, ctypes.c_int, _dp, _dp]),
    ("multiBatchDestroy", ctypes.c_int, [_vp]),
    ("setEngineOption", ctypes.c_int, [ctypes.c_int, ctypes.c_int, _vp]),
    ("getEngineOption", ctypes.c_int, [ctypes.c_int, _ip, _vp]),
    ("autotune", ctypes.c_int, [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p,
                                ctypes.c_int, _vp]),
    # run files
    ("setRunPath", ctypes.c_int, [ctypes.c_char_p, _vp]),
    ("loadRunFiles", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, _vp]),
    ("readParameterSets", ctypes.c_int, [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, _ip, _ip, _dp, _dp]),
    # learned warm start
    ("mlpCreate", _vp, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong]),
    ("mlpDestroy", ctypes.c_int, [_vp]),
    ("mlpParameterCount", ctypes.c_longlong, [_vp]),
    ("mlpGetParameters", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float)]),
    ("mlpSetParameters", ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_float)]),
    ("mlpForward", ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float)]),
    ("mlpGradient", ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                   ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                   ctypes.POINTER(ctypes.c_float)]),
    ("mlpTrainStep", ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_float),
                                    ctypes.POINTER(ctypes.c_float), ctypes.c_float, ctypes.POINTER(ctypes.c_float)]),
]

REFERENCE_SYMBOLS = [n for n, _, _ in C_API[:13]]

_LIB = None


def load_library(path=LIBPATH):
    """Loads the HIP library; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise ImportError("libansatzsynth.so is not built (python -m ansatzevolve_amd.build); "
                          "there is no CPU fallback")
    lib = ctypes.CDLL(path)
    for name, res, args in C_API:
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _LIB = lib
    return lib


class AnsatzError(RuntimeError):
    def __init__(self, fn, code):
        super().__init__("%s returned %d" % (fn, code))
        self.fn = fn
        self.code = code


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


class AnsatzSynth:
    """One ctx of the C interface.  Index conventions are the reference's:
    one-based basis indices (bit string + 1) and one-based operator
    components (0 = absent).  Methods raise AnsatzError on a non-zero code;
    ``call`` returns raw codes."""

    def __init__(self):
        self.lib = load_library()
        self.ctx = ctypes.c_void_p(self.lib.init())
        self.nparams = None

    def close(self):
        if self.ctx is not None and self.ctx.value:
            self.lib.cleanup(ctypes.byref(self.ctx))
        self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, fn, code):
        if code != 0:
            raise AnsatzError(fn, code)

    # ---- set-up (reference C interface)
    def set_initial_state(self, num_qubits, indexes_1based, coeffs):
        i, ip = _i(indexes_1based)
        c, cp = _d(coeffs)
        self._chk("setInitialState", self.lib.setInitialState(num_qubits, len(i), ip, cp, self.ctx))
        self.num_qubits = int(num_qubits)

    def set_initial_state_complex(self, num_qubits, indexes_1based, coeffs):
        """Complex start state (setInitialStateComplex); the context then
        evaluates complex amplitudes."""
        i, ip = _i(indexes_1based)
        c = np.ascontiguousarray(np.asarray(coeffs, dtype=np.complex128).reshape(-1))
        self._chk("setInitialStateComplex",
                  self.lib.setInitialStateComplex(num_qubits, len(i), ip, c.ctypes.data_as(_vp), self.ctx))
        self.num_qubits = int(num_qubits)

    def set_hamiltonian(self, i_1based, j_1based, coeffs):
        a, ap = _i(i_1based)
        b, bp = _i(j_1based)
        c, cp = _d(coeffs)
        self._chk("setHamiltonian", self.lib.setHamiltonian(len(a), ap, bp, cp, self.ctx))

    def set_excitation(self, operators_1based, order_1based, scale=None):
        """operators_1based: (M, 4) one-based components, 0 = absent."""
        ops = np.asarray(operators_1based, dtype=np.int32).reshape(-1, 4)
        M = len(ops)
        flat, fp = _i(ops.T.reshape(-1))  # operators[M*j + i] = component j of operator i
        o, op = _i(order_1based)
        if scale is None:
            code = self.lib.setExcitation(M, fp, op, self.ctx)
        else:
            s, sp = _d(scale)
            code = self.lib.setExcitationScale(M, fp, op, sp, self.ctx)
        self._chk("setExcitation", code)
        self.nparams = int(np.max(order_1based))

    # ---- extensions
    def set_integrals(self, h, g):
        h = np.asarray(h, dtype=np.float64)
        n = h.shape[0]
        a, ap = _d(h.reshape(-1))
        b, bp = _d(np.asarray(g, dtype=np.float64).reshape(-1))
        self._chk("setHamiltonianIntegrals", self.lib.setHamiltonianIntegrals(n, ap, bp, self.ctx))

    def set_nuclear_energy(self, e):
        self._chk("setNuclearEnergy", self.lib.setNuclearEnergy(float(e), self.ctx))

    def set_storage(self, mode):
        self._chk("setHamiltonianStorage", self.lib.setHamiltonianStorage(int(mode), self.ctx))

    def dimensions(self):
        d = ctypes.c_longlong()
        no = ctypes.c_int()
        npar = ctypes.c_int()
        self._chk("getDimensions", self.lib.getDimensions(ctypes.byref(d), ctypes.byref(no), ctypes.byref(npar),
                                                          self.ctx))
        return d.value, no.value, npar.value

    # ---- evaluations
    def energy(self, angles):
        a, ap = _d(angles)
        E = ctypes.c_double()
        self._chk("getEnergy", self.lib.getEnergy(len(a), ap, ctypes.byref(E), self.ctx))
        return E.value

    def gradient(self, angles):
        a, ap = _d(angles)
        g = np.zeros(len(a))
        self._chk("getGradient_COMP", self.lib.getGradient_COMP(len(a), ap, g.ctypes.data_as(_dp), self.ctx))
        return g

    def gradient_operators(self, angles):
        """Gradient per operator (dE/dx_k for operator k's angle)."""
        a, ap = _d(angles)
        M = self.dimensions()[1]
        g = np.zeros(M)
        self._chk("getGradientOperators", self.lib.getGradientOperators(len(a), ap, g.ctypes.data_as(_dp), self.ctx))
        return g

    def overlap_gradient(self, angles, proj_full):
        """(<proj|psi(angles)>, its per-parameter gradient); proj over the
        full space of 2^qubits basis states."""
        a, ap = _d(angles)
        pr, pp = _d(proj_full)
        g = np.zeros(len(a))
        ov = ctypes.c_double(0)
        self._chk("getOverlapGradient", self.lib.getOverlapGradient(len(a), ap, len(pr), pp, ctypes.byref(ov),
                                                                    g.ctypes.data_as(_dp), self.ctx))
        return ov.value, g

    def hessian_proj(self, angles, proj_full):
        """(Hessian, gradient, value) of the overlap <proj|psi(angles)>; proj
        over the full space of 2^qubits basis states."""
        a, ap = _d(angles)
        pr, pp = _d(proj_full)
        P = len(a)
        H = np.zeros(P * P)
        g = np.zeros(P)
        ov = ctypes.c_double(0)
        self._chk("getHessianProj", self.lib.getHessianProj(P, ap, len(pr), pp, H.ctypes.data_as(_dp),
                                                            g.ctypes.data_as(_dp), ctypes.byref(ov), self.ctx))
        return H.reshape(P, P).T.copy(), g, ov.value

    def hessian(self, angles):
        a, ap = _d(angles)
        P = len(a)
        H = np.zeros(P * P)
        self._chk("getHessian_COMP", self.lib.getHessian_COMP(P, ap, H.ctypes.data_as(_dp), self.ctx))
        return H.reshape(P, P).T.copy()  # column-major -> H[i, j]

    def final_state(self, angles, num_qubits):
        a, ap = _d(angles)
        n = 1 << num_qubits
        out = np.zeros(n)
        self._chk("getFinalState", self.lib.getFinalState(len(a), ap, n, out.ctypes.data_as(_dp), self.ctx))
        return out

    def final_state_complex(self, angles, num_qubits):
        """The final state over the 2^num_qubits basis states as complex
        amplitudes (getFinalStateComplex)."""
        a, ap = _d(angles)
        n = 1 << num_qubits
        out = np.zeros(n, dtype=np.complex128)
        self._chk("getFinalStateComplex",
                  self.lib.getFinalStateComplex(len(a), ap, n, out.ctypes.data_as(_vp), self.ctx))
        return out

    def energy_gradient(self, angles):
        a, ap = _d(angles)
        g = np.zeros(len(a))
        E = ctypes.c_double()
        self._chk("getEnergyGradient", self.lib.getEnergyGradient(len(a), ap, ctypes.byref(E),
                                                                  g.ctypes.data_as(_dp), self.ctx))
        return E.value, g

    def energies(self, angles2d):
        A = np.ascontiguousarray(angles2d, dtype=np.float64)
        nvec, na = A.shape
        E = np.zeros(nvec)
        self._chk("getEnergies", self.lib.getEnergies(nvec, na, A.ctypes.data_as(_dp), E.ctypes.data_as(_dp),
                                                      self.ctx))
        return E

    def energy_gradients(self, angles2d):
        A = np.ascontiguousarray(angles2d, dtype=np.float64)
        nvec, na = A.shape
        E = np.zeros(nvec)
        G = np.zeros((nvec, na))
        self._chk("getEnergyGradients", self.lib.getEnergyGradients(nvec, na, A.ctypes.data_as(_dp),
                                                                    E.ctypes.data_as(_dp),
                                                                    G.ctypes.data_as(_dp), self.ctx))
        return E, G


    # ---- device-resident batch
    def upload_batch(self, angles2d):
        A = np.ascontiguousarray(angles2d, dtype=np.float64)
        self._batch_shape = A.shape
        self._chk("uploadAngleBatch", self.lib.uploadAngleBatch(A.shape[0], A.shape[1], A.ctypes.data_as(_dp),
                                                                self.ctx))

    def run_batch(self, with_gradient=True):
        self._chk("runAngleBatch", self.lib.runAngleBatch(1 if with_gradient else 0, self.ctx))

    def download_batch(self, with_gradient=True):
        nvec, na = self._batch_shape
        E = np.zeros(nvec)
        G = np.zeros((nvec, na)) if with_gradient else None
        self._chk("downloadAngleBatch", self.lib.downloadAngleBatch(
            E.ctypes.data_as(_dp), G.ctypes.data_as(_dp) if with_gradient else None, self.ctx))
        return E, G

    def set_kernel_timing(self, on):
        self._chk("setKernelTiming", self.lib.setKernelTiming(1 if on else 0, self.ctx))

    def kernel_timing(self, cls):
        ms = ctypes.c_double()
        n = ctypes.c_longlong()
        self._chk("getKernelTiming", self.lib.getKernelTiming(cls, ctypes.byref(ms), ctypes.byref(n), self.ctx))
        return ms.value, n.value

    # engine option keys (setEngineOption)
    (OPT_SWEEP, OPT_LANE_MIN, OPT_GRAPH, OPT_FUSE_DEPTH, OPT_LDS_BUDGET, OPT_HAM_KIND, OPT_SEGMENTS,
     OPT_LANE_GROUPS, OPT_LANE_SEGMENTS, OPT_LANE_GROUPS_FWD, OPT_CHECKPOINT, OPT_CHECKPOINT_ACTIVE,
     OPT_SIGMA_HALVES, OPT_SIGMA_HALVES_AVAILABLE, OPT_SIGMA_GROUPED, OPT_SIGMA_GROUPED_AVAILABLE,
     OPT_GEMM_SOLUTION_BETA, OPT_GEMM_SOLUTION_ALPHA, OPT_REACHABLE, OPT_SWEEP_AMP_FORWARD,
     OPT_SWEEP_AMP_ADJOINT_LAMBDA, OPT_SWEEP_AMP_ADJOINT_STATE, OPT_SWEEP_TWO_GROUP_AVAILABLE, OPT_SIGMA1_ROWS, OPT_VECTOR_SEGMENTS,
     OPT_WINDOW, OPT_WINDOW_WAVES, OPT_WINDOW_PASSES, OPT_WINDOW_REGISTERS, OPT_WINDOW_VECTOR,
     OPT_WINDOW_ENERGY, OPT_WINDOW_HESSIAN, OPT_WINDOW_STAGED, OPT_SIGMA_EXCHANGE) = range(34)

    def set_option(self, key, value):
        self._chk("setEngineOption", self.lib.setEngineOption(int(key), int(value), self.ctx))

    def get_option(self, key):
        v = ctypes.c_int()
        self._chk("getEngineOption", self.lib.getEngineOption(int(key), ctypes.byref(v), self.ctx))
        return v.value

    def autotune(self, thetas, cache=None, reps=3):
        """Runs the batch autotuner on the angle batch `thetas` (nvec x
        angles) and returns its report (dict); the choice stays applied."""
        import json
        t = np.ascontiguousarray(np.atleast_2d(thetas), dtype=np.float64)
        buf = ctypes.create_string_buffer(1 << 16)
        self._chk("autotune", self.lib.autotune(t.shape[0], t.shape[1], t.ctypes.data_as(_dp),
                                                cache.encode() if cache else None, int(reps), buf, len(buf),
                                                self.ctx))
        return json.loads(buf.value.decode())


    # ---- run files (<prefix>_Initial.dat, _Operators.dat, _Order.dat, ...)
    def set_run_path(self, prefix):
        """The Hamiltonian is read from <prefix>_Ham_Coeff/_Ham_Index.dat or
        <prefix>_oneEInts/_twoEInts.bin at construction when none was set."""
        self._chk("setRunPath", self.lib.setRunPath(os.fsencode(prefix), self.ctx))

    def load_run_files(self, prefix, with_nuclear=False):
        """The whole set-up from the run files of `prefix`; with_nuclear adds
        <prefix>_Nuclear_Energy.dat to every energy."""
        self._chk("loadRunFiles", self.lib.loadRunFiles(os.fsencode(prefix), 1 if with_nuclear else 0, self.ctx))
        with open(os.fsdecode(prefix) + "_Initial.dat") as f:
            self.num_qubits = len(f.read().split(",")[0].strip())

    def optimise(self, angles, stop_energy=float("nan"), max_trace=100000):
        """Plain Newton optimisation from `angles`; returns (angles, E, trace)
        with trace rows (seconds, energy, |g|, Hessian evals, energy evals)."""
        a = np.array(angles, dtype=np.float64)
        E = ctypes.c_double()
        tr = np.zeros((max_trace, 5))
        nt = ctypes.c_int()
        self._chk("optimise", self.lib.optimise(len(a), a.ctypes.data_as(_dp), ctypes.byref(E), float(stop_energy),
                                                max_trace, tr.ctypes.data_as(_dp), ctypes.byref(nt), self.ctx))
        return a, E.value, tr[:nt.value].copy()

    def subspace_diagonalise(self, angles2d):
        """(eigenvalues, H, S, null rank) of the Hamiltonian in the span of
        the states of the angle vectors (rows)."""
        A = np.ascontiguousarray(np.atleast_2d(angles2d), dtype=np.float64)
        n = A.shape[0]
        ev, H, S = np.zeros(n), np.zeros((n, n)), np.zeros((n, n))
        nr = ctypes.c_int()
        self._chk("subspaceDiagonalise", self.lib.subspaceDiagonalise(
            n, A.shape[1], A.ctypes.data_as(_dp), ev.ctypes.data_as(_dp), H.ctypes.data_as(_dp),
            S.ctypes.data_as(_dp), ctypes.byref(nr), self.ctx))
        return ev, H, S, nr.value

    PROPERTY_COLUMNS = ["energy", "energy_with_nuclear", "hess_neg", "hess_zero", "hess_pos", "hess_diag_neg",
                        "hess_diag_zero", "hess_diag_pos", "metric_pos", "metric_zero", "metric_diag_zero",
                        "grad_norm"]

    def properties(self, angles2d, noverlaps=1, no_hessian=False):
        """writeProperties: (rows as dicts, Hessians, metrics, lowest
        eigenvalues of H)."""
        A = np.ascontiguousarray(np.atleast_2d(angles2d), dtype=np.float64)
        n = A.shape[0]
        P = self.dimensions()[2]
        cols = 12 + 2 * noverlaps
        R = np.zeros((n, cols))
        H = np.zeros((n, P, P))
        G = np.zeros((n, P, P))
        ev = np.zeros(max(noverlaps, 1))
        self._chk("getProperties", self.lib.getProperties(
            n, A.shape[1], A.ctypes.data_as(_dp), int(noverlaps), 1 if no_hessian else 0, R.ctypes.data_as(_dp),
            H.ctypes.data_as(_dp), G.ctypes.data_as(_dp), ev.ctypes.data_as(_dp), self.ctx))
        rows = []
        for r in R:
            d = {k: float(v) for k, v in zip(self.PROPERTY_COLUMNS, r[:12])}
            d["overlaps"] = [float(r[12 + 2 * t]) for t in range(noverlaps)]
            rows.append(d)
        return rows, H, G, ev[:noverlaps]

    def iterative_optimise(self, angles, ediis=False, max_trace=10000):
        """The iterative projected optimiser (iterativeTups) from `angles`;
        returns (angles, E, trace) with one trace row per outer iteration."""
        a = np.array(angles, dtype=np.float64)
        E = ctypes.c_double()
        tr = np.zeros((max_trace, 5))
        nt = ctypes.c_int()
        self._chk("iterativeOptimise", self.lib.iterativeOptimise(len(a), a.ctypes.data_as(_dp), ctypes.byref(E),
                                                                  1 if ediis else 0, max_trace,
                                                                  tr.ctypes.data_as(_dp), ctypes.byref(nt),
                                                                  self.ctx))
        return a, E.value, tr[:nt.value].copy()


class MultiBatch:
    """Angle batches of several AnsatzSynth contexts (uneven Hamiltonians)
    evaluated by one graph launch (multiBatch* entry points)."""

    def __init__(self, contexts):
        self.lib = load_library()
        self.contexts = list(contexts)
        arr = (ctypes.c_void_p * len(self.contexts))(*[c.ctx.value for c in self.contexts])
        self.h = self.lib.multiBatchCreate(len(self.contexts), arr)
        if not self.h:
            raise AnsatzError("multiBatchCreate", 5)
        self.nvec = [0] * len(self.contexts)
        self.nparam = [c.dimensions()[2] for c in self.contexts]

    def close(self):
        if self.h:
            self.lib.multiBatchDestroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def upload(self, i, thetas):
        t = np.ascontiguousarray(np.atleast_2d(thetas), dtype=np.float64)
        r = self.lib.multiBatchUpload(self.h, i, t.shape[0], t.shape[1], t.ctypes.data_as(_dp))
        if r != 0:
            raise AnsatzError("multiBatchUpload", r)
        self.nvec[i] = t.shape[0]

    def run(self, with_gradient=True):
        r = self.lib.multiBatchRun(self.h, 1 if with_gradient else 0)
        if r != 0:
            raise AnsatzError("multiBatchRun", r)

    def download(self, i, with_gradient=True):
        E = np.zeros(self.nvec[i])
        G = np.zeros((self.nvec[i], self.nparam[i])) if with_gradient else None
        r = self.lib.multiBatchDownload(self.h, i, E.ctypes.data_as(_dp), G.ctypes.data_as(_dp) if G is not None
                                        else None)
        if r != 0:
            raise AnsatzError("multiBatchDownload", r)
        return (E, G) if with_gradient else E


def read_parameter_sets(path, lib=None):
    """(energies, sets) of a _Parameters.dat file: the recorded energy of
    every set and its values (nsets x nparams, zero padded)."""
    lib = lib or load_library()
    ns, npar = ctypes.c_int(0), ctypes.c_int(0)
    code = lib.readParameterSets(os.fsencode(path), 0, 0, ctypes.byref(ns), ctypes.byref(npar), None, None)
    if code != 0:
        raise AnsatzError("readParameterSets", code)
    E = np.zeros(ns.value)
    P = np.zeros((ns.value, npar.value))
    if ns.value:
        code = lib.readParameterSets(os.fsencode(path), ns.value, npar.value, ctypes.byref(ns), ctypes.byref(npar),
                                     E.ctypes.data_as(_dp), P.ctypes.data_as(_dp))
        if code != 0:
            raise AnsatzError("readParameterSets", code)
    return E, P


def from_workload(w, storage=0):
    """An AnsatzSynth set up from ansatzevolve_amd.workloads.make_workload()."""
    a = AnsatzSynth()
    a.set_initial_state(w["nq"], [w["start_idx"] + 1], [1.0])
    a.set_integrals(w["h"], w["g"])
    if storage:
        a.set_storage(storage)
    ops = np.asarray(w["ops"]) + 1  # -1 (absent) -> 0
    a.set_excitation(ops, np.asarray(w["order"]) + 1, w.get("scale"))
    return a
