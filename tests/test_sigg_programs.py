"""Host-side checks of the grouped-target sigma programs (sigg.py): the
generated tables and unrolled programs, interpreted on the CPU, against the
per-target sum of the beta operator they implement, and the committed
header against the generator."""
import os

import numpy as np

from ansatzevolve_amd import sigg

K = 6
NB = 924


def _brute(G, x):
    """y[b] = sum_{r in b, q not in b} G[r][q] (-1)^bp(b, r, q) x[b - r + q]
    + (sum_{e in b} G[e][e]) x[b]: the beta entries of the sparse sigma
    (k_beta_table: singles with the between parity, then the diagonal)."""
    strs = sigg.subsets(12, K)
    rk = {s: i for i, s in enumerate(strs)}
    y = np.zeros_like(x)
    for i, s in enumerate(strs):
        acc = 0.0
        for r in sigg.bits(s):
            for q in range(12):
                if (s >> q) & 1:
                    continue
                sg = -1.0 if sigg.between_parity(s, r, q) else 1.0
                acc += sg * G[r][q] * x[rk[s ^ (1 << r) ^ (1 << q)]]
            acc += G[r][r] * x[i]
        y[i] = acc
    return y


def test_rank_splits_into_lower_and_upper_halves():
    for i, s in enumerate(sigg.subsets(12, K)):
        bl, h = s & 63, s >> 6
        m = bin(bl).count("1")
        assert i == sigg.rank_l(bl, m) + sigg.rh(h, m)


def test_programs_match_per_target_sum():
    rng = np.random.default_rng(5)
    for trial in range(3):
        G = rng.standard_normal((12, 12))
        x = rng.standard_normal(NB)
        ref = _brute(G, x)
        got = sigg.apply_model(G, x, K)
        assert np.max(np.abs(got - ref)) <= 1e-12 * max(1.0, np.max(np.abs(ref)))


def test_a_planted_sign_error_is_caught():
    """A wrong static sign in one FMA of one program must show up."""
    rng = np.random.default_rng(6)
    G = rng.standard_normal((12, 12))
    x = rng.standard_normal(NB)
    ref = _brute(G, x)
    orig = sigg.program

    def bad(Kk, m):
        p = orig(Kk, m)
        if m == 3:
            k, off, fl = p[5]
            t, c, neg = fl[0]
            p[5] = (k, off, [(t, c, not neg)] + fl[1:])
        return p

    sigg.program = bad
    try:
        got = sigg.apply_model(G, x, K)
    finally:
        sigg.program = orig
    assert np.max(np.abs(got - ref)) > 1e-6


def test_plan_covers_every_row_once_and_tables_in_bounds():
    plan = sigg.wave_plan(K)
    assert len(plan) <= 16
    rows = [h for wave in plan for (_, _, rs) in wave for h in rs if h >= 0]
    assert sorted(rows) == list(range(64))
    for wave in plan:
        for (m, off, rs) in wave:
            assert off + sigg.comb(6, m) <= 20
            for h in rs:
                if h < 0:
                    continue
                lay = sigg.row_layout(K, h)
                assert lay["m"] == m
                for (k, o, fl) in sigg.program(K, m):
                    assert 0 <= lay["bases"][k] + o <= (NB - 1) * sigg.SLOT
                for e in sigg.row_entries(K, h):
                    if e & 0x80000000:
                        assert ((e >> 12) & 0x7FFFF) < NB
                    else:
                        assert e % 8 == 0 and e // 8 <= 2 * 144


def test_generated_header_is_current(tmp_path):
    fresh = sigg.emit(path=str(tmp_path / "sigg_prog.inc"))
    committed = os.path.join(os.path.dirname(sigg.__file__), "csrc", "sigg_prog.inc")
    with open(fresh) as f, open(committed) as g:
        assert f.read() == g.read()
