"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
import this module, and only as the checker.  ``OracleEngine`` mirrors the
reference's FusedEvolve entry points (src/fusedevolve.h:61-75) over
liboracle.so (oracle/oracle.c); ``RefEngine`` drives the reference's own CPU
build (oracle/_ref/libref_ansatz.so, built by ``make -C oracle ref``).
"""
import ctypes
import os
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBORACLE = os.path.join(HERE, "liboracle.so")
LIBREF = os.path.join(HERE, "_ref", "libref_ansatz.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build_oracle():
    subprocess.check_call(["make", "-s", "-C", HERE], stdout=subprocess.DEVNULL)


def _lib():
    if not os.path.exists(LIBORACLE):
        build_oracle()
    L = ctypes.CDLL(LIBORACLE)
    L.or_engine_create.restype = ctypes.c_void_p
    L.or_engine_create.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _dp, ctypes.c_int, _ip, _ip, _dp]
    for name in ("or_engine_free",):
        getattr(L, name).argtypes = [ctypes.c_void_p]
    L.or_engine_mode.argtypes = [ctypes.c_void_p]
    L.or_engine_dim.argtypes = [ctypes.c_void_p]
    L.or_engine_dim.restype = ctypes.c_long
    L.or_engine_nparams.argtypes = [ctypes.c_void_p]
    L.or_engine_nnz.argtypes = [ctypes.c_void_p]
    L.or_engine_nnz.restype = ctypes.c_long
    L.or_engine_states.argtypes = [ctypes.c_void_p]
    L.or_engine_states.restype = ctypes.POINTER(ctypes.c_uint32)
    L.or_engine_set_triplets.argtypes = [ctypes.c_void_p, ctypes.c_long, _ip, _ip, _dp]
    L.or_engine_set_integrals.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_int]
    L.or_ham_apply.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.or_evolve.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.or_energy.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.or_energy_gradient.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
    L.or_hessian.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
    L.or_overlap_gradient.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
    L.or_engine_set_start_complex.argtypes = [ctypes.c_void_p, ctypes.c_int, _ip, _dp, _dp]
    L.or_evolve_complex.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.or_energy_gradient_complex.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
    L.or_apply_exc_complex.argtypes = [ctypes.c_uint32, ctypes.POINTER(ctypes.c_int8), ctypes.POINTER(ctypes.c_uint32)]
    L.or_hessian_proj.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp, _dp]
    L.or_decompress_angles.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.or_apply_exc.argtypes = [ctypes.c_uint32, ctypes.POINTER(ctypes.c_int8), ctypes.POINTER(ctypes.c_uint32)]
    return L


_L = None


def lib():
    global _L
    if _L is None:
        _L = _lib()
    return _L


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(_dp)


def _i(a):
    a = np.ascontiguousarray(a, dtype=np.int32)
    return a, a.ctypes.data_as(_ip)


def apply_exc(state, op):
    """applyExcToBasisState_ restated: returns (partner, sign)."""
    a = (ctypes.c_int8 * 4)(*[int(x) for x in op])
    out = ctypes.c_uint32()
    s = lib().or_apply_exc(ctypes.c_uint32(state), a, ctypes.byref(out))
    return out.value, s


class OracleEngine:
    """ops: (M,4) zero-based spin-orbital indices (-1 absent); order: zero-based."""

    def __init__(self, nq, start_idx, start_coeff, ops, order, scale=None):
        L = lib()
        self._keep = []
        si, sip = _i(start_idx)
        sc, scp = _d(start_coeff)
        ops = np.asarray(ops, dtype=np.int32).reshape(-1, 4)
        o, op = _i(ops.reshape(-1))
        od, odp = _i(order)
        if scale is None:
            scp2 = None
        else:
            s2, scp2 = _d(scale)
            self._keep.append(s2)
        self.h = L.or_engine_create(nq, len(si), sip, scp, len(ops), op, odp, scp2)
        if not self.h:
            raise ValueError("oracle rejected the set-up")
        self.nops = len(ops)
        self.N = L.or_engine_dim(self.h)
        self.P = L.or_engine_nparams(self.h)
        self.mode = L.or_engine_mode(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_engine_free(self.h)
            self.h = None

    def states(self):
        p = lib().or_engine_states(self.h)
        return np.ctypeslib.as_array(p, shape=(self.N,)).copy()

    def set_triplets(self, i0, j0, v):
        a, ap = _i(i0)
        b, bp = _i(j0)
        c, cp = _d(v)
        assert lib().or_engine_set_triplets(self.h, len(a), ap, bp, cp) == 0

    def set_integrals(self, h, g, explicit=True):
        n = int(round(len(np.ravel(h)) ** 0.5))
        a, ap = _d(np.ravel(h))
        b, bp = _d(np.ravel(g))
        assert lib().or_engine_set_integrals(self.h, n, ap, bp, 1 if explicit else 0) == 0

    def nnz(self):
        return lib().or_engine_nnz(self.h)

    def decompress(self, theta):
        t, tp = _d(theta)
        out = np.zeros(self.nops)
        lib().or_decompress_angles(self.h, tp, out.ctypes.data_as(_dp))
        return out

    def _ops_angles(self, angles):
        angles = np.asarray(angles, dtype=np.float64)
        if len(angles) == self.nops:
            return angles
        if len(angles) == self.P:
            return self.decompress(angles)
        raise ValueError("wrong angle count")

    def ham_apply(self, x):
        x, xp = _d(x)
        y = np.zeros(self.N)
        assert lib().or_ham_apply(self.h, xp, y.ctypes.data_as(_dp)) == 0
        return y

    def evolve(self, angles):
        a, ap = _d(self._ops_angles(angles))
        psi = np.zeros(self.N)
        rc = lib().or_evolve(self.h, ap, psi.ctypes.data_as(_dp))
        if rc:
            raise RuntimeError("oracle evolve failed %d" % rc)
        return psi

    def energy(self, angles):
        a, ap = _d(self._ops_angles(angles))
        E = ctypes.c_double()
        assert lib().or_energy(self.h, ap, ctypes.byref(E)) == 0
        return E.value

    def energy_gradient(self, angles):
        a, ap = _d(self._ops_angles(angles))
        E = ctypes.c_double()
        g = np.zeros(self.P)
        assert lib().or_energy_gradient(self.h, ap, ctypes.byref(E), g.ctypes.data_as(_dp)) == 0
        return E.value, g

    def hessian(self, angles):
        a, ap = _d(self._ops_angles(angles))
        E = ctypes.c_double()
        g = np.zeros(self.P)
        H = np.zeros(self.P * self.P)
        assert lib().or_hessian(self.h, ap, H.ctypes.data_as(_dp), g.ctypes.data_as(_dp), ctypes.byref(E)) == 0
        return H.reshape(self.P, self.P).T.copy(), g, E.value

    def set_start_complex(self, idx0, coeffs):
        """Complex start state (zero-based full-space indices)."""
        i, ip = _i(idx0)
        c = np.asarray(coeffs, dtype=complex)
        re, rp = _d(c.real.copy())
        im, imp = _d(c.imag.copy())
        assert lib().or_engine_set_start_complex(self.h, len(i), ip, rp, imp) == 0

    def evolve_complex(self, angles):
        """psi(angles) in complex mode, as a complex array over the sector."""
        a, ap = _d(self._ops_angles(angles))
        psi = np.zeros(2 * self.N)
        assert lib().or_evolve_complex(self.h, ap, psi.ctypes.data_as(_dp)) == 0
        return psi[0::2] + 1j * psi[1::2]

    def energy_gradient_complex(self, angles):
        a, ap = _d(self._ops_angles(angles))
        E = ctypes.c_double()
        g = np.zeros(self.P)
        assert lib().or_energy_gradient_complex(self.h, ap, ctypes.byref(E), g.ctypes.data_as(_dp)) == 0
        return E.value, g

    def overlap_gradient(self, angles, proj):
        """(proj . psi, its per-parameter gradient); proj over the sector."""
        a, ap = _d(self._ops_angles(angles))
        pr, pp = _d(proj)
        ov = ctypes.c_double()
        g = np.zeros(self.P)
        assert lib().or_overlap_gradient(self.h, ap, pp, ctypes.byref(ov), g.ctypes.data_as(_dp)) == 0
        return ov.value, g

    def hessian_proj(self, angles, proj):
        """(Hessian, gradient, value) of proj . psi; proj over the sector."""
        a, ap = _d(self._ops_angles(angles))
        pr, pp = _d(proj)
        ov = ctypes.c_double()
        g = np.zeros(self.P)
        H = np.zeros(self.P * self.P)
        assert lib().or_hessian_proj(self.h, ap, pp, H.ctypes.data_as(_dp), g.ctypes.data_as(_dp),
                                     ctypes.byref(ov)) == 0
        return H.reshape(self.P, self.P).T.copy(), g, ov.value


# ---------------------------------------------------------------- reference
def ref_available():
    return os.path.exists(LIBREF)


def _reflib():
    L = ctypes.CDLL(LIBREF)
    L.refd_create.restype = ctypes.c_void_p
    L.refd_create.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _dp, ctypes.c_int, _ip, _ip, _dp,
                              ctypes.c_char_p, ctypes.c_int, _ip, _ip, _dp]
    L.refd_destroy.argtypes = [ctypes.c_void_p]
    L.refd_energy.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
    L.refd_gradient.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_int]
    L.refd_hessian.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_int]
    L.refd_state.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_long]
    L.refd_fe_create.restype = ctypes.c_void_p
    L.refd_fe_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _dp,
                                 ctypes.c_int, _ip, _ip, _dp, ctypes.c_int, ctypes.c_char_p]
    L.refd_fe_destroy.argtypes = [ctypes.c_void_p]
    L.refd_fe_egrad.restype = ctypes.c_double
    L.refd_fe_egrad.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, _dp, _dp]
    L.refd_fe_dim.restype = ctypes.c_long
    L.refd_fe_dim.argtypes = [ctypes.c_void_p]
    L.refd_num_threads.restype = ctypes.c_int
    L.refd_fe_hessian_proj.argtypes = [ctypes.c_void_p, _dp, _dp, ctypes.c_int, _dp, _dp, _dp]
    L.refd_fe_iterative.restype = ctypes.c_double
    L.refd_fe_iterative.argtypes = [ctypes.c_void_p, _dp]
    return L


def write_integral_files(prefix, h, g):
    np.ascontiguousarray(h, dtype=np.float64).tofile(prefix + "_oneEInts.bin")
    np.ascontiguousarray(g, dtype=np.float64).tofile(prefix + "_twoEInts.bin")


class RefEngine:
    """The reference's stateAnsatzManager (through oracle/ref_driver.cpp)."""

    def __init__(self, nq, start_idx, start_coeff, ops, order, scale=None,
                 h=None, g=None, triplets=None, workdir=None):
        self.L = _reflib()
        self._tmp = None
        ops = np.asarray(ops, dtype=np.int32).reshape(-1, 4)
        self.nops = len(ops)
        self.P = int(np.max(order)) + 1
        si, sip = _i(start_idx)
        sc, scp = _d(start_coeff)
        o, op = _i(ops.reshape(-1))
        od, odp = _i(order)
        s2, scp2 = (None, None) if scale is None else _d(scale)
        if triplets is not None:
            ti, tip = _i(triplets[0])
            tj, tjp = _i(triplets[1])
            tv, tvp = _d(triplets[2])
            prefix = b""
            ntrip = len(ti)
        else:
            self._tmp = tempfile.TemporaryDirectory(dir=workdir)
            pre = os.path.join(self._tmp.name, "ham")
            write_integral_files(pre, h, g)
            prefix = pre.encode()
            ntrip, tip, tjp, tvp = 0, None, None, None
        self.h = self.L.refd_create(nq, len(si), sip, scp, len(ops), op, odp, scp2,
                                    prefix, ntrip, tip, tjp, tvp)
        if not self.h:
            raise ValueError("reference rejected the set-up")

    def close(self):
        if self.h:
            self.L.refd_destroy(self.h)
            self.h = None
        if self._tmp:
            self._tmp.cleanup()
            self._tmp = None

    def __del__(self):
        self.close()

    def energy(self, angles):
        a, ap = _d(angles)
        E = ctypes.c_double()
        assert self.L.refd_energy(self.h, len(a), ap, ctypes.byref(E)) == 0
        return E.value

    def gradient(self, angles):
        a, ap = _d(angles)
        g = np.zeros(self.P)
        assert self.L.refd_gradient(self.h, len(a), ap, g.ctypes.data_as(_dp), self.P) == 0
        return g

    def hessian(self, angles):
        a, ap = _d(angles)
        H = np.zeros(self.P * self.P)
        assert self.L.refd_hessian(self.h, len(a), ap, H.ctypes.data_as(_dp), self.P) == 0
        return H.reshape(self.P, self.P).T.copy()

    def state(self, angles, nfull):
        a, ap = _d(angles)
        out = np.zeros(nfull)
        assert self.L.refd_state(self.h, len(a), ap, out.ctypes.data_as(_dp), nfull) == 0
        return out


class RefFusedEvolve:
    """The reference's FusedEvolve built directly, as its CLI benchmark does
    (src/main.cpp:404-435); integral Hamiltonian, Sz sector."""

    def __init__(self, nq, nup, ndown, start_idx, start_coeff, ops, order, scale, h, g, workdir=None):
        self.L = _reflib()
        self.nq = nq
        ops = np.asarray(ops, dtype=np.int32).reshape(-1, 4)
        self.nops = len(ops)
        self.P = int(np.max(order)) + 1
        self._tmp = tempfile.TemporaryDirectory(dir=workdir)
        pre = os.path.join(self._tmp.name, "ham")
        write_integral_files(pre, h, g)
        si, sip = _i(start_idx)
        sc, scp = _d(start_coeff)
        o, op = _i(ops.reshape(-1))
        od, odp = _i(order)
        s2, scp2 = (None, None) if scale is None else _d(scale)
        self.h = self.L.refd_fe_create(nq, nup, ndown, len(si), sip, scp, len(ops), op, odp, scp2,
                                       self.P, pre.encode())
        if not self.h:
            raise ValueError("reference rejected the set-up")

    def dim(self):
        return self.L.refd_fe_dim(self.h)

    def threads(self):
        return self.L.refd_num_threads()

    def egrad(self, angles_ops, reps=1):
        a, ap = _d(angles_ops)
        E = ctypes.c_double()
        g = np.zeros(self.P)
        t = self.L.refd_fe_egrad(self.h, ap, reps, ctypes.byref(E), g.ctypes.data_as(_dp))
        return E.value, g, t

    def hessian_proj(self, angles_ops, proj_full):
        """FusedEvolve::evolveHessianProj: (Hessian, gradient, value)."""
        a, ap = _d(angles_ops)
        pr, pp = _d(proj_full)
        H = np.zeros(self.P * self.P)
        g = np.zeros(self.P)
        ov = ctypes.c_double()
        rc = self.L.refd_fe_hessian_proj(self.h, ap, pp, self.nq, H.ctypes.data_as(_dp), g.ctypes.data_as(_dp),
                                         ctypes.byref(ov))
        assert rc == 0, rc
        return H.reshape(self.P, self.P).T.copy(), g, ov.value

    def iterative(self, angles_ops):
        """TUPSQuantities::iterativeTups from per-operator angles: (final
        per-operator angles, energy)."""
        a = np.array(angles_ops, dtype=np.float64)
        E = self.L.refd_fe_iterative(self.h, a.ctypes.data_as(_dp))
        return a, E

    def close(self):
        if self.h:
            self.L.refd_fe_destroy(self.h)
            self.h = None
        if self._tmp:
            self._tmp.cleanup()
            self._tmp = None

    def __del__(self):
        self.close()


# ------------------------------------------------------- reference, complex
LIBREF_COMPLEX = os.path.join(HERE, "_ref", "libref_ansatz_complex.so")


def ref_complex_available():
    return os.path.exists(LIBREF_COMPLEX)


class RefComplex:
    """The reference's stateAnsatzManager compiled in its complex mode
    (through oracle/ref_driver_complex.cpp): complex start, triplet
    Hamiltonian, operators (number operators allowed)."""

    def __init__(self, nq, start_idx, start_coeff, ops, order, scale, ti, tj, tv):
        L = ctypes.CDLL(LIBREF_COMPLEX)
        L.refc_create.restype = ctypes.c_void_p
        L.refc_create.argtypes = [ctypes.c_int, ctypes.c_int, _ip, _dp, _dp, ctypes.c_int, _ip, _ip, _dp,
                                  ctypes.c_int, _ip, _ip, _dp]
        L.refc_destroy.argtypes = [ctypes.c_void_p]
        L.refc_energy.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp]
        L.refc_gradient.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, ctypes.c_int]
        L.refc_state.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, _dp, _dp, ctypes.c_long]
        self.L, self.nq = L, nq
        c = np.asarray(start_coeff, dtype=complex)
        si, sip = _i(start_idx)
        re, rp = _d(c.real.copy())
        im, imp = _d(c.imag.copy())
        ops = np.asarray(ops, dtype=np.int32).reshape(-1, 4)
        o, op = _i(ops.reshape(-1))
        od, odp = _i(order)
        sc, scp = _d(np.ones(len(ops)) if scale is None else scale)
        a, ap = _i(ti)
        b, bp = _i(tj)
        v, vp = _d(tv)
        self.P = int(np.max(order)) + 1
        self.h = L.refc_create(nq, len(si), sip, rp, imp, len(ops), op, odp, scp, len(a), ap, bp, vp)
        if not self.h:
            raise ValueError("reference rejected the set-up")

    def energy(self, angles):
        a, ap = _d(angles)
        E = ctypes.c_double()
        assert self.L.refc_energy(self.h, len(a), ap, ctypes.byref(E)) == 0
        return E.value

    def gradient(self, angles):
        a, ap = _d(angles)
        g = np.zeros(self.P)
        assert self.L.refc_gradient(self.h, len(a), ap, g.ctypes.data_as(_dp), self.P) == 0
        return g

    def state(self, angles):
        a, ap = _d(angles)
        n = 1 << self.nq
        re, im = np.zeros(n), np.zeros(n)
        assert self.L.refc_state(self.h, len(a), ap, re.ctypes.data_as(_dp), im.ctypes.data_as(_dp), n) == 0
        return re + 1j * im

    def close(self):
        if self.h:
            self.L.refc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
