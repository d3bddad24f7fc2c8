"""Canonical TUPS angles for the learned warm start's training targets,
checked on the CPU against the oracle: the canonical form keeps the energy
and the final state (up to its sign), maps angle vectors that give the same
state to the same vector, and is idempotent."""
import numpy as np
import pytest
from scipy.optimize import least_squares

from ansatzevolve_amd.learned import (_canonical_first_unit, _unit_state, canonical_angles, output_slots,
                                      tups_units)
from ansatzevolve_amd.workloads import hf_index, seeded_angles, synthetic_integrals, tups_operators
from oracle.oracle import OracleEngine


def _oracle(n, layers, seed):
    h, g = synthetic_integrals(n, seed)
    ops, order, scale = tups_operators(n, layers)
    o = OracleEngine(2 * n, [hf_index(n)], [1.0], ops, order, scale)
    o.set_integrals(h, g, explicit=True)
    return o


@pytest.mark.parametrize("n,layers", [(4, 2), (6, 3)])
def test_canonical_angles_keep_energy_and_state(n, layers):
    o = _oracle(n, layers, 3100 + n)
    P = 3 * len(tups_units(n, layers))
    for th in seeded_angles(P, 77 + n, count=6):
        c = canonical_angles(th, n, layers)
        assert np.all(c > -np.pi) and np.all(c <= np.pi)
        E0, g0 = o.energy_gradient(th)
        E1, _ = o.energy_gradient(c)
        assert abs(E0 - E1) <= 1e-12
        s0, s1 = o.evolve(th), o.evolve(c)
        assert min(np.max(np.abs(s0 - s1)), np.max(np.abs(s0 + s1))) <= 1e-12
        # first half-layer units in the tied form, every theta_a in (-pi/2, pi/2]
        for layer, p, i in tups_units(n, layers):
            assert -np.pi / 2 < c[i] <= np.pi / 2 + 1e-15
            if layer == 0 and p % 2 == 0:
                assert abs(c[i] + c[i + 2]) <= 1e-12
        assert np.max(np.abs(canonical_angles(c, n, layers) - c)) <= 1e-12


def test_canonical_angles_unique_per_state():
    """Angle vectors reached by the unit symmetries (theta_a + pi,
    theta_c + pi) and by another point of a first unit's gauge orbit (found
    numerically) give the same state and the same canonical vector."""
    n, layers = 6, 2
    o = _oracle(n, layers, 3200)
    P = 3 * len(tups_units(n, layers))
    rng = np.random.default_rng(5)
    moved = 0
    for th in seeded_angles(P, 91, count=4):
        c = canonical_angles(th, n, layers)
        alt = th.copy()
        for layer, p, i in tups_units(n, layers):
            if rng.uniform() < 0.5:
                alt[i] += np.pi
                alt[i + 2] -= np.pi
        # another gauge point of the first unit (pair 0): fix theta_a, solve
        # for theta_b, theta_c reaching the same pair state
        ta, tb, tc = th[0], th[1], th[2]
        r = _unit_state(ta, tb, tc)
        ta2 = ta + 0.37
        sol = least_squares(lambda v: _unit_state(ta2, v[0], v[1]) - r, x0=[tb, tc], xtol=1e-15, ftol=1e-15,
                            gtol=1e-15)
        if np.max(np.abs(_unit_state(ta2, *sol.x) - r)) < 1e-10:
            alt[0], alt[1], alt[2] = ta2, sol.x[0], sol.x[1]
            moved += 1
        E0, _ = o.energy_gradient(th)
        E1, _ = o.energy_gradient(alt)
        assert abs(E0 - E1) <= 1e-12
        assert np.max(np.abs(canonical_angles(alt, n, layers) - c)) <= 1e-9
    assert moved >= 2  # the gauge orbit was actually sampled


def test_first_unit_canonical_form_edge_cases():
    """No rotation, the full pair double (LL -> HH) and its negative."""
    for t in [(0.0, 0.0, 0.0), (0.0, np.pi / 2, 0.0), (0.0, -np.pi / 2, 0.0), (np.pi, 0.0, 0.0)]:
        a, b, c = _canonical_first_unit(*t)
        r0, r1 = _unit_state(*t), _unit_state(a, b, c)
        assert min(np.max(np.abs(r0 - r1)), np.max(np.abs(r0 + r1))) <= 1e-14
        assert abs(a + c) <= 1e-15


def test_output_slots_align_units_across_orbital_counts():
    """A unit on the same pair in the same layer has the same network
    outputs at every orbital count; slots are distinct and fit the grid."""
    s12 = output_slots(12, 6, 12)
    assert sorted(s12) == list(range(198))
    s8 = output_slots(8, 4, 12)
    assert len(set(s8)) == 84 and s8.max() < 198
    u8 = {(layer, p): i for layer, p, i in tups_units(8, 4)}
    u12 = {(layer, p): i for layer, p, i in tups_units(12, 6)}
    for key, i in u8.items():
        for k in range(3):
            assert s8[i + k] == s12[u12[key] + k]


def test_evaluation_parts_merge_and_time_to_target(tmp_path):
    """Host logic of the staged pipeline: the time to 1e-6 Ha of a run's own
    final energy and of the lower one, misses, the summary per orbital
    count, and merging per-count evaluations into one."""
    import json
    from ansatzevolve_amd.warmstart import rows_of, stage_merge, summarise

    def trace(E_end, n_it):  # rows: seconds, energy, |g|_inf, evaluations
        return np.array([[0.1 * i, E_end + 2.0 ** -i, 1.0, 3 * i + 1] for i in range(n_it)])

    held = [dict(n=8, index=0, spacing=1.2), dict(n=8, index=1, spacing=2.0), dict(n=10, index=2, spacing=1.5)]
    plain = [(None, -1.0, trace(-1.0, 30), "stalled"), (None, -2.0, trace(-2.0, 30), "stalled"),
             (None, -3.0, trace(-3.0, 40), "stalled")]
    learned = [(None, -1.0, trace(-1.0, 10), "stalled"), (None, -1.5, trace(-1.5, 30), "stalled"),
               (None, -3.0, trace(-3.0, 20), "stalled")]
    t_eval = {8: 0.001, 10: 0.002}
    rows = rows_of(held, plain, learned, np.array([0.5, 0.5, 0.5]), np.array([0.4, 0.7, 0.3]),
                   [True, False, True], t_eval, 0.01)
    # first iterate within 1e-6 of the own final energy: 2^-i <= 1e-6 from i = 20
    first = min(i for i in range(30) if 2.0 ** -i <= 1e-6)
    assert rows[0]["ev_own_plain"] == 3 * first + 1
    assert np.isinf(rows[0]["ev_own_mlp"])  # 10 iterates never get within 1e-6
    assert np.isinf(rows[1]["t_best_mlp"]) and not np.isinf(rows[1]["t_best_plain"])  # higher minimum: a miss
    assert rows[2]["t_own_mlp"] == rows[2]["ev_own_mlp"] * 0.002 + 0.01
    s = summarise(rows)
    assert s["all"]["count"] == 3 and s["n8"]["count"] == 2 and s["n10"]["count"] == 1
    assert s["all"]["proposals_used"] == 2 and s["all"]["plain_lower_minimum"] == 1
    parts = []
    for n, sel in ((8, rows[:2]), (10, rows[2:])):
        d = dict(counts={}, eval_batch_s=1.0, egrad_seconds={str(n): t_eval[n]}, judge_s_per_instance=0.01,
                 summary=summarise(sel), rows=sel)
        p = tmp_path / ("eval_n%d.json" % n)
        p.write_text(json.dumps(d))
        parts.append(str(p))
    m = stage_merge(str(tmp_path), parts)
    assert m["held_out"] == 3 and m["eval_batch_s"] == 2.0
    assert json.loads((tmp_path / "eval.json").read_text())["summary"] == json.loads(json.dumps(s))
