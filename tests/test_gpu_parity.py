"""GPU parity: the HIP library through its C ABI against the CPU oracle.

Tolerances are those of the target: |dE| <= 1e-10, gradient <= 1e-9,
Hessian <= 1e-8 (absolute, FP64).
"""
import json
import os

import numpy as np
import pytest

from ansatzevolve_amd import AnsatzSynth, AnsatzError, from_workload
from ansatzevolve_amd.workloads import (hf_index, make_workload, seeded_angles, synthetic_integrals,
                                        tups_operators)
from oracle.oracle import OracleEngine

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
E_TOL, G_TOL, H_TOL = 1e-10, 1e-9, 1e-8


def kat():
    with open(os.path.join(HERE, "golden", "testf90_kat.json")) as f:
        return json.load(f)


def kat_engines():
    k = kat()
    a = AnsatzSynth()
    a.set_initial_state(k["num_qubits"], [k["initial_index_1based"]], [k["initial_coeff"]])
    a.set_hamiltonian(k["ham_i_1based"], k["ham_j_1based"], k["ham_coeffs"])
    a.set_excitation(k["operators_1based"], k["order_1based"])
    o = OracleEngine(k["num_qubits"], [k["initial_index_1based"] - 1], [k["initial_coeff"]],
                     np.array(k["operators_1based"]) - 1, np.array(k["order_1based"]) - 1)
    o.set_triplets(np.array(k["ham_i_1based"]) - 1, np.array(k["ham_j_1based"]) - 1, k["ham_coeffs"])
    return k, a, o


def test_kat_testf90(gpu):
    """The reference's test.F90 set-up re-driven through the C ABI."""
    k, a, o = kat_engines()
    with a:
        for name in ("hf", "min"):
            ang = np.array(k["angles_" + name])
            E = a.energy(ang)
            assert abs(E - k["energy_" + name]) <= E_TOL, (name, E)
            g = a.gradient(ang)
            assert abs(np.linalg.norm(g) - k["grad_norm_" + name]) <= G_TOL
            Eo, go = o.energy_gradient(ang)
            assert abs(E - Eo) <= E_TOL
            assert np.max(np.abs(g - go)) <= G_TOL
            H = a.hessian(ang)
            Ho, _, _ = o.hessian(ang)
            assert np.max(np.abs(H - Ho)) <= H_TOL
            st = a.final_state(ang, k["num_qubits"])
            assert abs(np.linalg.norm(st) - 1.0) < 1e-12
            so = np.zeros(256)
            so[o.states()] = o.evolve(ang)
            assert np.max(np.abs(st - so)) <= 1e-12


def _integral_case(n, layers, seed, nparam_sets=2):
    h, g = synthetic_integrals(n, seed)
    ops, order, scale = tups_operators(n, layers)
    w = dict(n=n, nq=2 * n, h=h, g=g, ops=ops, order=order, scale=scale, start_idx=hf_index(n))
    o = OracleEngine(2 * n, [w["start_idx"]], [1.0], ops, order, scale)
    o.set_integrals(h, g, explicit=True)
    thetas = seeded_angles(int(order.max()) + 1, seed + 7, count=nparam_sets)
    return w, o, thetas


@pytest.mark.parametrize("storage", [1, 2], ids=["rows", "matrixfree"])
@pytest.mark.parametrize("n,layers", [(4, 2), (6, 2), (8, 2)])
def test_integrals_energy_gradient_hessian(gpu, n, layers, storage):
    w, o, thetas = _integral_case(n, layers, 1000 + n)
    with from_workload(w, storage=storage) as a:
        for th in thetas:
            E, g = a.energy_gradient(th)
            Eo, go = o.energy_gradient(th)
            assert abs(E - Eo) <= E_TOL
            assert np.max(np.abs(g - go)) <= G_TOL
            assert abs(a.energy(th) - Eo) <= E_TOL
            np.testing.assert_allclose(a.gradient(th), go, rtol=0, atol=G_TOL)
            H = a.hessian(th)
            Ho, _, _ = o.hessian(th)
            assert np.max(np.abs(H - Ho)) <= H_TOL
            assert np.max(np.abs(H - H.T)) <= 1e-10


def test_batch_matches_single(gpu):
    w, o, thetas = _integral_case(6, 3, 77, nparam_sets=5)
    with from_workload(w) as a:
        E, G = a.energy_gradients(thetas)
        Eb = a.energies(thetas)
        for v, th in enumerate(thetas):
            Es, gs = a.energy_gradient(th)
            assert abs(E[v] - Es) <= 1e-12
            assert abs(Eb[v] - Es) <= 1e-12
            assert np.max(np.abs(G[v] - gs)) <= 1e-12


def test_per_operator_angles_and_cache(gpu):
    w, o, thetas = _integral_case(4, 2, 5)
    with from_workload(w) as a:
        th = thetas[0]
        x = o.decompress(th)
        assert abs(a.energy(x) - a.energy(th)) <= 1e-13
        with pytest.raises(AnsatzError) as ei:
            a.gradient(x)  # per-operator length: the reference answers -6
        assert ei.value.code == -6
        with pytest.raises(AnsatzError) as ei:
            a.energy(th[:-1])
        assert ei.value.code == 3


@pytest.mark.parametrize("storage", [1, 2], ids=["rows", "matrixfree"])
def test_config2_energy_gradient(gpu, storage):
    """10 orbitals, 63,504 determinants, TUPS L=5 against the reference's
    own CPU build (golden vectors)."""
    gz = np.load(os.path.join(HERE, "golden", "reference_vectors.npz"), allow_pickle=False)
    w = make_workload(2)
    with from_workload(w, storage=storage) as a:
        d, M, P = a.dimensions()
        assert (d, M, P) == (63504, 225, 135)
        for i, th in enumerate(gz["cfg2_theta"]):
            E, g = a.energy_gradient(th)
            assert abs(E - gz["cfg2_E"][i]) <= E_TOL
            assert np.max(np.abs(g - gz["cfg2_grad"][i])) <= G_TOL
        H = a.hessian(gz["cfg2_theta"][0])
        assert np.max(np.abs(H - gz["cfg2_hess"][0])) <= H_TOL


@pytest.mark.parametrize("storage", [2, 1], ids=["matrixfree", "rows"])
def test_config3_energy_gradient(gpu, storage):
    """12 orbitals, 853,776 determinants, TUPS L=6 against the reference's
    own CPU build (its matrix-free path), and batched against single."""
    gz = np.load(os.path.join(HERE, "golden", "reference_vectors.npz"), allow_pickle=False)
    if "cfg3_E" not in gz.files:
        pytest.skip("no 12-orbital golden vector")
    w = make_workload(3)
    assert np.array_equal(gz["cfg3_theta"][0], w["theta"])
    with from_workload(w, storage=storage) as a:
        E, g = a.energy_gradient(w["theta"])
        assert abs(E - gz["cfg3_E"][0]) <= E_TOL
        assert np.max(np.abs(g - gz["cfg3_grad"][0])) <= G_TOL
        th = np.stack([w["theta"], w["theta"] + 1e-2])
        Eb, Gb = a.energy_gradients(th)
        assert abs(Eb[0] - E) <= 1e-12 and np.max(np.abs(Gb[0] - g)) <= 1e-12


def test_config3_benchmarked_batch_path(gpu):
    """The path the throughput benchmark times: 12 orbitals, 853,776
    determinants, TUPS L=6, a 64-vector batch after the autotuner (lane
    layout, window passes with the tiles in registers, stored forward states,
    reachable tiles, grouped-target sigma, tuned DGEMM solutions), replayed
    from its hipGraph.  Vector 0 is theta_0 and is checked against the
    reference CPU build's golden vector; other vectors against single-vector
    calls."""
    gz = np.load(os.path.join(HERE, "golden", "reference_vectors.npz"), allow_pickle=False)
    if "cfg3_E" not in gz.files:
        pytest.skip("no 12-orbital golden vector")
    w = make_workload(3)
    assert np.array_equal(gz["cfg3_theta"][0], w["theta"])
    rng = np.random.default_rng(w["seed"] + 3)
    thetas = w["theta"][None, :] + 1e-2 * rng.uniform(-np.pi, np.pi, size=(64, w["P"]))
    thetas[0] = w["theta"]
    with from_workload(w) as a:
        rep = a.autotune(thetas, reps=1)
        assert rep["rejected"] == 0
        # the benchmarked choice whatever the timing said: lane layout,
        # automatic window passes, stored states (the tuned DGEMMs stay)
        for key, val in ((AnsatzSynth.OPT_LANE_MIN, 1), (AnsatzSynth.OPT_WINDOW, -1),
                         (AnsatzSynth.OPT_CHECKPOINT, 1)):
            a.set_option(key, val)
        a.upload_batch(thetas)
        assert a.get_option(AnsatzSynth.OPT_WINDOW_PASSES) == 16
        assert a.get_option(AnsatzSynth.OPT_WINDOW_REGISTERS) == 2
        assert a.get_option(AnsatzSynth.OPT_HAM_KIND) == 3
        assert a.get_option(AnsatzSynth.OPT_CHECKPOINT_ACTIVE) == 1
        assert a.get_option(AnsatzSynth.OPT_REACHABLE) == 1
        assert a.get_option(AnsatzSynth.OPT_SIGMA_GROUPED_AVAILABLE) == 1
        assert a.get_option(AnsatzSynth.OPT_SIGMA_GROUPED) == 1
        # TUPS units from the Hartree-Fock determinant: exchange-symmetric
        assert a.get_option(AnsatzSynth.OPT_SIGMA_EXCHANGE) == 1
        for _ in range(3):  # capture, then replays
            a.run_batch(True)
        E, G = a.download_batch(True)
        assert abs(E[0] - gz["cfg3_E"][0]) <= E_TOL
        assert np.max(np.abs(G[0] - gz["cfg3_grad"][0])) <= G_TOL
        for v in (1, 17, 40, 63):
            Es, gs = a.energy_gradient(thetas[v])
            assert abs(E[v] - Es) <= 1e-12, v
            assert np.max(np.abs(G[v] - gs)) <= 1e-11, v
        # the same batch with both same-spin DGEMMs
        a.set_option(AnsatzSynth.OPT_SIGMA_EXCHANGE, 0)
        assert a.get_option(AnsatzSynth.OPT_SIGMA_EXCHANGE) == 0
        for _ in range(2):
            a.run_batch(True)
        E0, G0 = a.download_batch(True)
        assert np.max(np.abs(E - E0)) <= 1e-12
        assert np.max(np.abs(G - G0)) <= 1e-11
        a.set_option(AnsatzSynth.OPT_SIGMA_EXCHANGE, -1)
        a.upload_batch(thetas)
        assert a.get_option(AnsatzSynth.OPT_SIGMA_EXCHANGE) == 1


@pytest.mark.parametrize("sym", [True, False], ids=["symmetric", "asymmetric"])
def test_config3_exchange_start_states(gpu, sym):
    """The exchange-symmetric sigma applies only to batches whose states are
    symmetric under the exchange of the alpha and beta strings: a start
    vector with x[a][b] == x[b][a] keeps it, one without falls back to both
    DGEMMs (option 33 reads 2); either way the 64-vector batch equals
    single-vector calls (vector-major sigma with both DGEMMs)."""
    w = make_workload(3)
    n = w["n"]
    hf = w["start_idx"]
    # move one electron p -> q in the lower-bit string, and its mirror in
    # the upper-bit string
    p, q = 0, 1
    lo_move = hf ^ (1 << p) ^ (1 << q)
    hi_move = hf ^ (1 << (p + n)) ^ (1 << (q + n))
    assert bin(lo_move).count("1") == bin(hf).count("1")
    idx = [hf, lo_move, hi_move]
    coef = [0.9, 0.3, 0.3 if sym else -0.1]
    nrm = float(np.sqrt(np.sum(np.square(coef))))
    coef = [c / nrm for c in coef]
    rng = np.random.default_rng(w["seed"] + 5)
    thetas = w["theta"][None, :] + 1e-2 * rng.uniform(-np.pi, np.pi, size=(64, w["P"]))
    with from_workload(w) as a:
        a.set_initial_state(w["nq"], [i + 1 for i in idx], coef)
        a.set_option(AnsatzSynth.OPT_LANE_MIN, 1)
        a.upload_batch(thetas)
        assert a.get_option(AnsatzSynth.OPT_SIGMA_EXCHANGE) == (1 if sym else 2)
        for _ in range(2):
            a.run_batch(True)
        E, G = a.download_batch(True)
        for v in (0, 31, 63):
            Es, gs = a.energy_gradient(thetas[v])
            assert abs(E[v] - Es) <= 1e-12, v
            assert np.max(np.abs(G[v] - gs)) <= 1e-11, v
        # single-vector calls: half the sigma targets (symmetric start) or all
        # of them, against option 33 off (every target); twice each, so the
        # captured call sequence is replayed and recaptured when it changes
        for _ in range(2):
            E1, g1 = a.energy_gradient(thetas[5])
            a.set_option(AnsatzSynth.OPT_SIGMA_EXCHANGE, 0)
            E0, g0 = a.energy_gradient(thetas[5])
            a.set_option(AnsatzSynth.OPT_SIGMA_EXCHANGE, -1)
            assert abs(E1 - E0) <= 1e-12
            assert np.max(np.abs(g1 - g0)) <= 1e-11
        # a vector-major batch (below the lane threshold) takes the same
        # choice per upload
        a.set_option(AnsatzSynth.OPT_LANE_MIN, 1000)
        a.upload_batch(thetas[:4])
        assert a.get_option(AnsatzSynth.OPT_SIGMA_EXCHANGE) == (1 if sym else 2)
        for _ in range(2):
            a.run_batch(True)
        Ev, Gv = a.download_batch(True)
        for v in range(4):
            assert abs(Ev[v] - E[v]) <= 1e-12, v
            assert np.max(np.abs(Gv[v] - G[v])) <= 1e-11, v


def test_config4_energy_gradient(gpu):
    """14 orbitals, 11,778,624 determinants, TUPS L=7 (273 parameters, 455
    operators), matrix-free sigma: energy and gradient at theta_0 and at a
    second seeded angle vector (far from any stationary point) against the
    reference's own CPU build."""
    gz = np.load(os.path.join(HERE, "golden", "reference_vectors.npz"), allow_pickle=False)
    if "cfg4_E" not in gz.files:
        pytest.skip("no 14-orbital golden vector")
    w = make_workload(4)
    assert np.array_equal(gz["cfg4_theta"][0], w["theta"])
    with from_workload(w) as a:
        d, M, P = a.dimensions()
        assert (d, M, P) == (11778624, 455, 273)
        assert a.get_option(AnsatzSynth.OPT_HAM_KIND) == 3
        for i, th in enumerate(gz["cfg4_theta"]):
            E, g = a.energy_gradient(th)
            assert abs(E - gz["cfg4_E"][i]) <= E_TOL, (i, E)
            assert np.max(np.abs(g - gz["cfg4_grad"][i])) <= G_TOL, i
            assert np.linalg.norm(g) > 1e-2  # not a stationary point


HAMS_KEYS = ["H4_L1_", "H4_L5_", "H6_L1_", "H6_L5_", "H8_L1_", "H10_L1_", "H10_L5_", "H14_L5_"]


@pytest.mark.parametrize("key", HAMS_KEYS)
def test_hams_minima(gpu, key):
    """The reference's H-chain minima (Hams/*/L*/*_Parameters.dat, electronic
    energy) through the C ABI; H14 L5 has 11,778,624 determinants and runs
    the matrix-free sigma build."""
    d = np.load(os.path.join(HERE, "golden", "hams.npz"), allow_pickle=False)
    nq = int(d[key + "nq"])
    with AnsatzSynth() as a:
        a.set_initial_state(nq, d[key + "init_idx0"] + 1, d[key + "init_coeff"])
        a.set_integrals(d[key + "h"].reshape(nq // 2, nq // 2), d[key + "g"])
        a.set_excitation(d[key + "ops1"], d[key + "order0"] + 1, d[key + "scale"])
        for E_ref, p in zip(d[key + "energies"], d[key + "params"]):
            x = d[key + "scale"] * p[d[key + "order0"]]
            E = a.energy(x)
            assert abs(E - E_ref) <= E_TOL, (key, E, E_ref)
            P = int(d[key + "order0"].max()) + 1
            E2, g = a.energy_gradient(p[:P])
            assert abs(E2 - E_ref) <= E_TOL
            assert np.linalg.norm(g) < 1e-5


def test_newton_reaches_reference_minimum(gpu):
    """The plain optimiser from a perturbation of the reference's H4 L1
    minimum returns to it (Hams/H4/L1/H4_Parameters.dat: -2.838345804245581)
    with a stationary gradient and a positive semi-definite Hessian."""
    d = np.load(os.path.join(HERE, "golden", "hams.npz"), allow_pickle=False)
    key = "H4_L1_"
    P = int(d[key + "order0"].max()) + 1
    with AnsatzSynth() as a:
        a.set_initial_state(8, d[key + "init_idx0"] + 1, d[key + "init_coeff"])
        a.set_integrals(d[key + "h"].reshape(4, 4), d[key + "g"])
        a.set_excitation(d[key + "ops1"], d[key + "order0"] + 1, d[key + "scale"])
        start = d[key + "params"][0][:P] + 0.05 * np.cos(np.arange(P))
        th, E, trace = a.optimise(start)
        assert abs(E - d[key + "energies"][0]) <= 1e-10
        assert np.linalg.norm(a.gradient(th)) < 1e-9
        assert np.linalg.eigvalsh(a.hessian(th)).min() > -1e-8
        assert len(trace) >= 2 and trace[-1, 1] <= trace[0, 1] + 1e-12


def test_newton_synthetic_converges(gpu):
    """From the Hartree-Fock angles of a synthetic n=6 TUPS, the optimiser
    ends at a stationary point below the start, matching the oracle there."""
    w, o, thetas = _integral_case(6, 2, 4242)
    with from_workload(w) as a:
        P = int(w["order"].max()) + 1
        th, E, trace = a.optimise(0.1 * np.ones(P))
        Eo, go = o.energy_gradient(th)
        assert abs(E - Eo) <= E_TOL
        assert np.linalg.norm(go) < 1e-8
        assert E < o.energy(0.1 * np.ones(P))


def test_graph_replay_matches_eager(gpu):
    """getEnergyGradient at new angles runs the fused sequence from a
    hipGraph; every replay (interleaved with eager calls) must equal the
    eager adjoint sweep at the same angles (getGradient_COMP on the cached
    state) and the oracle."""
    w, o, thetas = _integral_case(6, 3, 99, nparam_sets=4)
    with from_workload(w) as a:
        for rep in range(2):
            for th in thetas:
                E, g = a.energy_gradient(th)       # fused, graph
                g_eager = a.gradient(th)           # cached psi, eager sweep
                E_eager = a.energy(th)
                assert abs(E - E_eager) <= 1e-13
                assert np.max(np.abs(g - g_eager)) <= 1e-13
                Eo, go = o.energy_gradient(th)
                assert abs(E - Eo) <= E_TOL and np.max(np.abs(g - go)) <= G_TOL
                a.energies(thetas[:2])            # other device work between replays


def _triplets_from_dense(states, Hd, thr=0.0):
    ii, jj = np.nonzero(np.abs(Hd) > thr)
    return np.array([states[i] for i in ii]), np.array([states[j] for j in jj]), Hd[ii, jj]


@pytest.mark.parametrize("case", ["superposition_sz", "full_space", "one_parameter"])
def test_sector_modes_and_start_states(gpu, case):
    """Start states and sectors beyond a single determinant: a normalised
    superposition within one S_z sector; a start that mixes particle numbers
    (the full Fock space, as the reference's compressor-free path); and an
    ansatz with a single parameter shared by every operator (scales 1, -0.5,
    2).  Energy, gradient, Hessian and the final state against the oracle."""
    nq = 8
    rng = np.random.default_rng({"superposition_sz": 1, "full_space": 2, "one_parameter": 3}[case])
    if case == "superposition_sz":
        start = [0b00110011, 0b01010101, 0b10011001]  # 2 up + 2 down electrons each
        coeff = np.array([0.6, -0.48, 0.64])
        states = [b for b in range(256) if bin(b & 0xF).count("1") == 2 and bin(b >> 4).count("1") == 2]
        ops = np.array([[0, 1, -1, -1], [4, 5, -1, -1], [0, 4, 2, 6], [1, 3, -1, -1], [5, 7, -1, -1],
                        [2, 6, 3, 7]])
        order = np.array([0, 0, 1, 2, 2, 3])
        scale = None
    elif case == "full_space":
        start = [0b00000011, 0b00000111]  # two and three electrons
        coeff = np.array([0.8, 0.6])
        states = list(range(256))
        ops = np.array([[2, 0, -1, -1], [4, 1, -1, -1], [3, 5, 0, 1], [6, 2, -1, -1], [7, 3, 4, 5]])
        order = np.array([0, 1, 2, 3, 4])
        scale = None
    else:
        start = [0b00110011]
        coeff = np.array([1.0])
        states = [b for b in range(256) if bin(b & 0xF).count("1") == 2 and bin(b >> 4).count("1") == 2]
        ops = np.array([[0, 2, -1, -1], [4, 6, -1, -1], [1, 5, 2, 6]])
        order = np.array([0, 0, 0])
        scale = np.array([1.0, -0.5, 2.0])
    Hd = rng.normal(size=(len(states), len(states)))
    Hd = 0.5 * (Hd + Hd.T)
    ti, tj, tv = _triplets_from_dense(states, Hd, 0.3)
    P = int(order.max()) + 1
    o = OracleEngine(nq, start, coeff, ops, order, scale)
    o.set_triplets(ti, tj, tv)
    with AnsatzSynth() as a:
        a.set_initial_state(nq, [s + 1 for s in start], coeff)
        a.set_hamiltonian(ti + 1, tj + 1, tv)
        a.set_excitation(ops + 1, order + 1, scale)
        for th in seeded_angles(P, 70, count=3):
            E, g = a.energy_gradient(th)
            Eo, go = o.energy_gradient(th)
            assert abs(E - Eo) <= E_TOL
            assert np.max(np.abs(g - go)) <= G_TOL
            H = a.hessian(th)
            Ho, _, _ = o.hessian(th)
            assert np.max(np.abs(H - Ho)) <= H_TOL
            st = a.final_state(th, nq)
            so = np.zeros(256)
            so[o.states()] = o.evolve(th)
            assert np.max(np.abs(st - so)) <= 1e-12
        Eb = a.energies(seeded_angles(P, 71, count=20))
        for v, th in enumerate(seeded_angles(P, 71, count=20)):
            assert abs(Eb[v] - o.energy(th)) <= E_TOL


def test_config3_hessian_exchange_symmetry(gpu):
    """The Hessian's H B at 853,776 determinants: TUPS tangents of a state
    from the Hartree-Fock determinant are exchange-symmetric (mirror
    operators share their parameter), so H B takes one DGEMM and one
    triangle of sigma targets; against option 33 off (both DGEMMs, every
    target)."""
    w = make_workload(3)
    with from_workload(w) as a:
        H1 = a.hessian(w["theta"])
        a.set_option(AnsatzSynth.OPT_SIGMA_EXCHANGE, 0)
        H0 = a.hessian(w["theta"])
        a.set_option(AnsatzSynth.OPT_SIGMA_EXCHANGE, -1)
    scale = max(1.0, float(np.max(np.abs(H0))))
    assert np.max(np.abs(H1 - H0)) <= 1e-12 * scale
    assert np.max(np.abs(H1 - H1.T)) <= 1e-10 * scale
