"""Window passes of the lane-layout sweeps (engine option 25): every run of
operators inside a window of consecutive spatial orbitals applied to
LDS-staged tiles in one read and write of the state.  Checked against the
orbit-segment sweeps, single-vector calls, the oracle and the reference
build's golden vector at 853,776 determinants."""
import os

import numpy as np
import pytest

from ansatzevolve_amd import AnsatzSynth, from_workload
from ansatzevolve_amd.workloads import hf_index, make_workload, seeded_angles, synthetic_integrals, tups_operators
from oracle.oracle import OracleEngine

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
E_TOL, G_TOL = 1e-10, 1e-9
SAME = 1e-12  # the two sweep forms differ only in summation order


def _tups_case(n, layers, seed, nvec):
    h, g = synthetic_integrals(n, seed)
    ops, order, scale = tups_operators(n, layers)
    w = dict(n=n, nq=2 * n, h=h, g=g, ops=ops, order=order, scale=scale, start_idx=hf_index(n))
    thetas = seeded_angles(int(order.max()) + 1, seed + 7, count=nvec)
    return w, thetas


def _window(a, width, waves=4, ck=1):
    a.set_option(AnsatzSynth.OPT_LANE_MIN, 1)
    a.set_option(AnsatzSynth.OPT_CHECKPOINT, ck)
    a.set_option(AnsatzSynth.OPT_WINDOW_WAVES, waves)
    a.set_option(AnsatzSynth.OPT_WINDOW, width)


@pytest.mark.parametrize("n,layers,width,waves,ck", [
    (8, 4, 4, 4, 1), (8, 4, 5, 8, 1), (8, 4, 4, 4, 0), (10, 3, 5, 4, 1), (10, 3, 5, 8, 0), (4, 2, 6, 4, 1),
])
def test_window_passes_match_orbit_segments(gpu, n, layers, width, waves, ck):
    """TUPS batches of 70 vectors (a full and a partial 64-vector chunk) with
    window passes against the orbit-segment sweeps, single-vector calls and
    the oracle; stored states (ck 1) and in-place sweeps (ck 0); tiles in LDS
    (five orbitals) and in registers (four).  A window wider than the
    orbital count is one window over everything."""
    w, thetas = _tups_case(n, layers, 4300 + n, 70)
    with from_workload(w, storage=2) as a:
        _window(a, 0, waves, ck)
        E0, G0 = a.energy_gradients(thetas)
        assert a.get_option(AnsatzSynth.OPT_WINDOW_PASSES) == 0
        _window(a, width, waves, ck)
        E1, G1 = a.energy_gradients(thetas)
        passes = a.get_option(AnsatzSynth.OPT_WINDOW_PASSES)
        reg = a.get_option(AnsatzSynth.OPT_WINDOW_REGISTERS)
        print("n=%d L=%d w=%d waves=%d ck=%d: %d passes, registers %d" % (n, layers, width, waves, ck, passes, reg))
        assert passes > 0
        # four-orbital windows of TUPS units keep their tiles in registers
        assert reg == (2 if min(width, n) == 4 else 0)
        if width >= n:
            assert passes == 1
        assert a.get_option(AnsatzSynth.OPT_CHECKPOINT_ACTIVE) == ck
        assert np.max(np.abs(E1 - E0)) <= SAME and np.max(np.abs(G1 - G0)) <= SAME
        for v in (0, 63, 69):
            Es, gs = a.energy_gradient(thetas[v])
            assert abs(E1[v] - Es) <= SAME and np.max(np.abs(G1[v] - gs)) <= SAME
    if n <= 8:
        o = OracleEngine(2 * n, [w["start_idx"]], [1.0], w["ops"], w["order"], w["scale"])
        o.set_integrals(w["h"], w["g"], explicit=True)
        Eo, go = o.energy_gradient(thetas[5])
        assert abs(E1[5] - Eo) <= E_TOL and np.max(np.abs(G1[5] - go)) <= G_TOL


@pytest.mark.parametrize("width", [4, 5], ids=["registers", "lds"])
@pytest.mark.parametrize("ck", [0, 1], ids=["inplace", "stored"])
def test_window_reachable_tiles(gpu, ck, width):
    """Passes that process only the tiles the start state and the passes
    before them can reach, against passes over every tile: the skipped tiles
    hold exact zeros, so the energies are the same bit for bit and the
    gradients up to the order of the bracket sums.  From one and from two
    start determinants."""
    n = 10
    w, thetas = _tups_case(n, 3, 5252, 70)
    hf = hf_index(n)
    other = hf ^ (1 << (n - 2)) ^ (1 << (n - 1))  # an alpha electron moved up, same sector
    for idx, cf in (([hf + 1], [1.0]), ([hf + 1, other + 1], [0.8, 0.6])):
        with from_workload(w, storage=2) as a:
            a.set_initial_state(w["nq"], idx, cf)
            _window(a, width, 4, ck)
            assert a.get_option(AnsatzSynth.OPT_REACHABLE) == 1
            assert a.get_option(AnsatzSynth.OPT_WINDOW_REGISTERS) == (2 if width == 4 else 0)
            E1, G1 = a.energy_gradients(thetas)
            amp1 = a.get_option(AnsatzSynth.OPT_SWEEP_AMP_FORWARD)
            a.set_option(AnsatzSynth.OPT_REACHABLE, 0)
            E0, G0 = a.energy_gradients(thetas)
            amp0 = a.get_option(AnsatzSynth.OPT_SWEEP_AMP_FORWARD)
            assert amp0 == a.dimensions()[0] * a.get_option(AnsatzSynth.OPT_WINDOW_PASSES)
            assert amp1 < amp0
            a.set_option(AnsatzSynth.OPT_REACHABLE, 1)
            E2, G2 = a.energy_gradients(thetas)
            assert np.array_equal(E1, E0) and np.max(np.abs(G1 - G0)) <= SAME
            assert np.array_equal(E2, E1) and np.array_equal(G2, G1)
            Es, gs = a.energy_gradient(thetas[69])
            assert abs(E1[69] - Es) <= SAME and np.max(np.abs(G1[69] - gs)) <= SAME


def _local_sz_operators(n, count, seed, span):
    """Random S_z-conserving singles and mixed-spin doubles whose orbitals lie
    within `span` consecutive spatial orbitals (so that they fit a window)."""
    rng = np.random.default_rng(seed)
    ops = []
    for _ in range(count):
        lo = int(rng.integers(0, n - span + 1))
        kind = int(rng.integers(0, 3))
        p, q = rng.choice(np.arange(lo, lo + span), 2, replace=False)
        r, s = rng.choice(np.arange(lo, lo + span), 2, replace=False)
        if kind == 0:
            ops.append([p, q, -1, -1])             # beta single
        elif kind == 1:
            ops.append([p + n, q + n, -1, -1])     # alpha single
        else:
            ops.append([p + n, r, q + n, s])       # alpha-beta double
    ops = np.array(ops)
    P = max(2, count * 2 // 3)
    order = np.concatenate([np.arange(P), rng.integers(0, P, count - P)])
    rng.shuffle(order)
    scale = rng.choice([1.0, -1.0, 0.5], count)
    return ops, order, scale


@pytest.mark.parametrize("n,count,span,width,seed", [(6, 40, 3, 4, 1), (7, 60, 4, 5, 2), (6, 30, 2, 3, 3)])
def test_window_random_operator_lists(gpu, n, count, span, width, seed):
    """Operator lists without TUPS structure (random singles and mixed-spin
    doubles on nearby orbitals, shared parameters, mixed scales): window
    passes against the oracle, with stored states and in place."""
    h, g = synthetic_integrals(n, 700 + seed)
    ops, order, scale = _local_sz_operators(n, count, seed, span)
    start = hf_index(n)
    w = dict(n=n, nq=2 * n, h=h, g=g, ops=ops, order=order, scale=scale, start_idx=start)
    o = OracleEngine(2 * n, [start], [1.0], ops, order, scale)
    o.set_integrals(h, g, explicit=True)
    P = int(order.max()) + 1
    thetas = seeded_angles(P, 90 + seed, count=5)
    ref = [o.energy_gradient(t) for t in thetas]
    with from_workload(w, storage=2) as a:
        for ck in (1, 0):
            for waves in (4, 8):
                _window(a, width, waves, ck)
                E, G = a.energy_gradients(thetas)
                assert a.get_option(AnsatzSynth.OPT_WINDOW_PASSES) > 0
                assert a.get_option(AnsatzSynth.OPT_WINDOW_REGISTERS) == 0  # not TUPS units
                for v, (E0, g0) in enumerate(ref):
                    assert abs(E[v] - E0) <= E_TOL and np.max(np.abs(G[v] - g0)) <= G_TOL, (ck, waves, v)


@pytest.mark.parametrize("variant", ["alpha_first", "no_double", "tups"])
def test_window_register_forms(gpu, variant):
    """Operator runs on one adjacent pair in both spins that are not the
    five-operator TUPS pattern (the two singles of each group swapped, or the
    paired double left out) run the generic register form (option 28 = 1)
    and agree with the orbit-segment sweeps and the oracle; TUPS itself runs
    the straight-line form (2).  Swapping the commuting alpha and beta
    singles leaves the state unchanged."""
    n, layers = 8, 3
    w, thetas = _tups_case(n, layers, 4411, 70)
    ops, order = w["ops"].copy(), w["order"].copy()
    keep = np.ones(len(ops), bool)
    for u in range(len(ops) // 5):
        b = 5 * u
        if variant == "alpha_first":
            for g in (b, b + 3):
                ops[[g, g + 1]] = ops[[g + 1, g]]
        elif variant == "no_double":
            keep[b + 2] = False
    w2 = dict(w, ops=ops[keep], order=order[keep], scale=w["scale"][keep])
    with from_workload(w2, storage=2) as a:
        _window(a, 0)
        E0, G0 = a.energy_gradients(thetas)
        _window(a, 4)
        E1, G1 = a.energy_gradients(thetas)
        assert a.get_option(AnsatzSynth.OPT_WINDOW_REGISTERS) == (2 if variant == "tups" else 1)
        assert np.max(np.abs(E1 - E0)) <= SAME and np.max(np.abs(G1 - G0)) <= SAME
    o = OracleEngine(2 * n, [w2["start_idx"]], [1.0], w2["ops"], w2["order"], w2["scale"])
    o.set_integrals(w2["h"], w2["g"], explicit=True)
    Eo, go = o.energy_gradient(thetas[3])
    assert abs(E1[3] - Eo) <= E_TOL and np.max(np.abs(G1[3] - go)) <= G_TOL
    if variant == "alpha_first":
        with from_workload(w, storage=2) as a:
            _window(a, 4)
            E2, _ = a.energy_gradients(thetas)
        assert np.max(np.abs(E2 - E1)) <= SAME


def test_window_config3_against_reference(gpu):
    """12 orbitals, 853,776 determinants, TUPS L=6, a 64-vector batch with
    window passes over four and five orbitals (stored states, reachable
    tiles, grouped sigma), replayed from its hipGraph: vector 0 = theta_0
    against the reference CPU build's golden vector, every vector against the
    orbit-segment batch.  Six orbitals do not fit: no window plan."""
    gz = np.load(os.path.join(HERE, "golden", "reference_vectors.npz"), allow_pickle=False)
    w = make_workload(3)
    assert np.array_equal(gz["cfg3_theta"][0], w["theta"])
    rng = np.random.default_rng(w["seed"] + 5)
    thetas = w["theta"][None, :] + 1e-2 * rng.uniform(-np.pi, np.pi, size=(64, w["P"]))
    thetas[0] = w["theta"]
    with from_workload(w) as a:
        _window(a, 0)
        E0, G0 = a.energy_gradients(thetas)
        _window(a, 6)
        assert a.get_option(AnsatzSynth.OPT_WINDOW_PASSES) == 0
        for width, waves, passes in ((5, 4, 11), (4, 8, 16)):
            _window(a, width, waves, 1)
            a.upload_batch(thetas)
            assert a.get_option(AnsatzSynth.OPT_WINDOW_PASSES) == passes
            assert a.get_option(AnsatzSynth.OPT_WINDOW_REGISTERS) == (2 if width == 4 else 0)
            assert a.get_option(AnsatzSynth.OPT_CHECKPOINT_ACTIVE) == 1
            for _ in range(3):  # capture, then replays
                a.run_batch(True)
            E, G = a.download_batch(True)
            assert abs(E[0] - gz["cfg3_E"][0]) <= E_TOL
            assert np.max(np.abs(G[0] - gz["cfg3_grad"][0])) <= G_TOL
            assert np.max(np.abs(E - E0)) <= SAME and np.max(np.abs(G - G0)) <= 1e-11


@pytest.mark.parametrize("n,layers,variant", [(8, 4, "tups"), (10, 3, "tups"), (8, 3, "alpha_first")])
def test_window_vector_major(gpu, n, layers, variant):
    """Window passes in the vector-major sweeps (option 29 = 1, one tile per
    lane, in place): single-vector E+grad and a 5-vector batch below the
    lane threshold against the orbit-segment sweeps (option 29 = 0) and the
    oracle.  The straight-line TUPS form and the generic register form
    (the singles of each group swapped); class task lists with partial
    waves (the tile counts are not multiples of 64)."""
    w, thetas = _tups_case(n, layers, 4500 + n, 5)
    if variant == "alpha_first":
        ops = w["ops"].copy()
        for u in range(len(ops) // 5):
            for g in (5 * u, 5 * u + 3):
                ops[[g, g + 1]] = ops[[g + 1, g]]
        w = dict(w, ops=ops)
    with from_workload(w, storage=2) as a:
        a.set_option(AnsatzSynth.OPT_LANE_MIN, 64)
        a.set_option(AnsatzSynth.OPT_WINDOW, 4)
        a.set_option(AnsatzSynth.OPT_WINDOW_VECTOR, 0)
        ref = [a.energy_gradient(t) for t in thetas]
        Eb0, Gb0 = a.energy_gradients(thetas)
        segs0 = a.get_option(AnsatzSynth.OPT_VECTOR_SEGMENTS)
        a.set_option(AnsatzSynth.OPT_WINDOW_VECTOR, 1)
        assert a.get_option(AnsatzSynth.OPT_WINDOW_VECTOR) == 1
        passes = a.get_option(AnsatzSynth.OPT_WINDOW_PASSES)
        assert a.get_option(AnsatzSynth.OPT_WINDOW_REGISTERS) == (2 if variant == "tups" else 1)
        assert a.get_option(AnsatzSynth.OPT_VECTOR_SEGMENTS) == passes
        print("n=%d L=%d %s: %d window passes vs %d vector segments" % (n, layers, variant, passes, segs0))
        for rep in range(2):  # the second call replays the captured graph
            for v, t in enumerate(thetas):
                E, g = a.energy_gradient(t)
                assert abs(E - ref[v][0]) <= SAME and np.max(np.abs(g - ref[v][1])) <= SAME, (rep, v)
        Eb, Gb = a.energy_gradients(thetas)
        assert np.max(np.abs(Eb - Eb0)) <= SAME and np.max(np.abs(Gb - Gb0)) <= SAME
        # the energy-only batch path
        assert np.max(np.abs(a.energies(thetas) - Eb0)) <= SAME
        # alpha rows staged in LDS (option 32, in use here) against one tile per lane from memory
        assert a.get_option(AnsatzSynth.OPT_WINDOW_STAGED) == 1
        a.set_option(AnsatzSynth.OPT_WINDOW_STAGED, 0)
        assert a.get_option(AnsatzSynth.OPT_WINDOW_STAGED) == 0
        for v, t in enumerate(thetas[:2]):
            E, g = a.energy_gradient(t)
            assert abs(E - ref[v][0]) <= SAME and np.max(np.abs(g - ref[v][1])) <= SAME, v
        a.set_option(AnsatzSynth.OPT_WINDOW_STAGED, 1)
        if n <= 8:
            # the other vector-major callers: Hessian, projected gradient, subspace products
            rng = np.random.default_rng(n + 17)
            proj = rng.standard_normal(1 << (2 * n))
            proj /= np.linalg.norm(proj)
            got = (a.hessian(thetas[1]), a.overlap_gradient(thetas[1], proj), a.subspace_diagonalise(thetas[:3]))
            a.set_option(AnsatzSynth.OPT_WINDOW_VECTOR, 0)
            want = (a.hessian(thetas[1]), a.overlap_gradient(thetas[1], proj), a.subspace_diagonalise(thetas[:3]))
            assert np.max(np.abs(got[0] - want[0])) <= 1e-11
            assert abs(got[1][0] - want[1][0]) <= SAME and np.max(np.abs(got[1][1] - want[1][1])) <= SAME
            assert np.max(np.abs(got[2][1] - want[2][1])) <= 1e-11 and np.max(np.abs(got[2][2] - want[2][2])) <= 1e-11
    if n <= 8:
        o = OracleEngine(2 * n, [w["start_idx"]], [1.0], w["ops"], w["order"], w["scale"])
        o.set_integrals(w["h"], w["g"], explicit=True)
        Eo, go = o.energy_gradient(thetas[2])
        assert abs(ref[2][0] - Eo) <= E_TOL and np.max(np.abs(ref[2][1] - go)) <= G_TOL


def test_window_vector_major_config3(gpu):
    """12 orbitals, 853,776 determinants: single-vector E+grad with the
    default vector-major window passes (on from 2^18 determinants) against
    the reference CPU build's golden vector and the orbit-segment sweeps."""
    gz = np.load(os.path.join(HERE, "golden", "reference_vectors.npz"), allow_pickle=False)
    w = make_workload(3)
    with from_workload(w) as a:
        E, g = a.energy_gradient(w["theta"])
        assert a.get_option(AnsatzSynth.OPT_WINDOW_VECTOR) == -1
        assert a.get_option(AnsatzSynth.OPT_VECTOR_SEGMENTS) == a.get_option(AnsatzSynth.OPT_WINDOW_PASSES) == 16
        assert abs(E - gz["cfg3_E"][0]) <= E_TOL and np.max(np.abs(g - gz["cfg3_grad"][0])) <= G_TOL
        assert a.get_option(AnsatzSynth.OPT_WINDOW_STAGED) == 1
        a.set_option(AnsatzSynth.OPT_WINDOW_STAGED, 0)
        E1, g1 = a.energy_gradient(w["theta"])
        assert abs(E - E1) <= SAME and np.max(np.abs(g - g1)) <= 1e-11
        a.set_option(AnsatzSynth.OPT_WINDOW_VECTOR, 0)
        E0, g0 = a.energy_gradient(w["theta"])
        assert abs(E - E0) <= SAME and np.max(np.abs(g - g0)) <= 1e-11


def test_window_fused_energy(gpu):
    """The energies of a gradient batch from the first adjoint window pass
    (option 30 = 1, the default) against the separate dot (0) and the
    energy-only batch, 70 vectors (two chunks) at 8 orbitals with stored
    states and in place; the gradients are unchanged."""
    w, thetas = _tups_case(8, 4, 4700, 70)
    with from_workload(w, storage=2) as a:
        for ck in (1, 0):
            _window(a, 4, 4, ck)
            assert a.get_option(AnsatzSynth.OPT_WINDOW_ENERGY) == 1
            E1, G1 = a.energy_gradients(thetas)
            Eonly = a.energies(thetas)
            a.set_option(AnsatzSynth.OPT_WINDOW_ENERGY, 0)
            E0, G0 = a.energy_gradients(thetas)
            a.set_option(AnsatzSynth.OPT_WINDOW_ENERGY, 1)
            assert a.get_option(AnsatzSynth.OPT_WINDOW_REGISTERS) == 2
            assert np.max(np.abs(E1 - E0)) <= SAME and np.max(np.abs(E1 - Eonly)) <= SAME, ck
            assert np.array_equal(G1, G0), ck
            Es, gs = a.energy_gradient(thetas[69])
            assert abs(E1[69] - Es) <= SAME


@pytest.mark.parametrize("n,layers", [(8, 4), (10, 3)])
def test_window_hessian_tangents(gpu, n, layers):
    """The Hessian's tangent sweep as window passes (option 31 = 1, the
    default) against the two-group segments (0), for the energy and the
    projected overlap, and the energy Hessian against the oracle at 8
    orbitals."""
    w, thetas = _tups_case(n, layers, 4800 + n, 2)
    with from_workload(w, storage=2) as a:
        assert a.get_option(AnsatzSynth.OPT_WINDOW_REGISTERS) == 2
        assert a.get_option(AnsatzSynth.OPT_WINDOW_HESSIAN) == 1
        rng = np.random.default_rng(n)
        proj = rng.standard_normal(1 << (2 * n)) if n <= 8 else None
        if proj is not None:
            proj /= np.linalg.norm(proj)
        H1 = a.hessian(thetas[0])
        P1 = a.hessian_proj(thetas[1], proj) if proj is not None else None
        a.set_option(AnsatzSynth.OPT_WINDOW_HESSIAN, 0)
        H0 = a.hessian(thetas[0])
        P0 = a.hessian_proj(thetas[1], proj) if proj is not None else None
        scale = max(1.0, np.max(np.abs(H0)))
        assert np.max(np.abs(H1 - H0)) <= 1e-12 * scale
        assert np.max(np.abs(H1 - H1.T)) <= 1e-10 * scale
        if P1 is not None:
            assert np.max(np.abs(P1[0] - P0[0])) <= 1e-12 and np.max(np.abs(P1[1] - P0[1])) <= 1e-12
    if n <= 8:
        o = OracleEngine(2 * n, [w["start_idx"]], [1.0], w["ops"], w["order"], w["scale"])
        o.set_integrals(w["h"], w["g"], explicit=True)
        Ho = o.hessian(thetas[0])[0]
        assert np.max(np.abs(H1 - Ho)) <= 1e-8
