"""The warm-start network's HIP kernels against a float64 numpy restatement:
forward pass, loss, parameter gradients (and central finite differences of
the numpy loss), and one Adam step."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def np_forward(p, X):
    W0, b0, W1, b1, W2, b2 = [q.astype(np.float64) for q in p]
    a1 = np.tanh(X @ W0 + b0)
    a2 = np.tanh(a1 @ W1 + b1)
    return a1, a2, a2 @ W2 + b2


def np_loss_grad(p, X, T, M):
    W0, b0, W1, b1, W2, b2 = [q.astype(np.float64) for q in p]
    X = X.astype(np.float64)
    a1, a2, y = np_forward(p, X)
    ms = M.sum()
    loss = float((M * (y - T) ** 2).sum() / ms)
    dy = 2 * M * (y - T) / ms
    gW2 = a2.T @ dy
    gb2 = dy.sum(0)
    d2 = (dy @ W2.T) * (1 - a2 ** 2)
    gW1 = a1.T @ d2
    gb1 = d2.sum(0)
    d1 = (d2 @ W1.T) * (1 - a1 ** 2)
    gW0 = X.T @ d1
    gb0 = d1.sum(0)
    return loss, np.concatenate([q.reshape(-1) for q in (gW0, gb0, gW1, gb1, gW2, gb2)])


@pytest.fixture
def net(gpu):
    from ansatzevolve_amd.learned import WarmStartMLP
    m = WarmStartMLP(37, 19, hidden=24, seed=5)
    yield m
    m.close()


def data(seed, B=11, nin=37, nout=19):
    rng = np.random.default_rng(seed)
    X = rng.normal(size=(B, nin)).astype(np.float32)
    T = rng.normal(size=(B, nout)).astype(np.float32)
    M = (rng.random((B, nout)) < 0.8).astype(np.float32)
    return X, T, M


def test_forward_matches_numpy(net):
    X, T, M = data(1)
    Y = net.forward(X)
    _, _, Yn = np_forward(net.unpack(), X.astype(np.float64))
    np.testing.assert_allclose(Y, Yn, rtol=1e-4, atol=1e-5)


def test_gradient_matches_numpy_and_finite_differences(net):
    X, T, M = data(2)
    loss, g = net.gradient(X, T, M)
    p = net.unpack()
    ln, gn = np_loss_grad(p, X, T, M)
    assert abs(loss - ln) <= 1e-4 * max(1.0, abs(ln))
    np.testing.assert_allclose(g, gn, rtol=1e-3, atol=1e-5)
    # finite differences of the float64 loss pin the analytic numpy gradient
    flat = net.get_params().astype(np.float64)
    rng = np.random.default_rng(0)
    for idx in rng.choice(len(flat), 12, replace=False):
        e = np.zeros_like(flat)
        e[idx] = 1e-6
        lp = np_loss_grad(net.unpack(flat + e), X, T, M)[0]
        lm = np_loss_grad(net.unpack(flat - e), X, T, M)[0]
        assert abs((lp - lm) / 2e-6 - gn[idx]) < 1e-5 + 1e-4 * abs(gn[idx])


def test_adam_step_matches_numpy(net):
    X, T, M = data(3)
    p0 = net.get_params().astype(np.float64)
    _, g = net.gradient(X, T, M)
    net.train_step(X, T, M, 1e-2)
    p1 = net.get_params()
    # first Adam step with zero moments: p -= lr * g / (|g| + eps)
    exp = p0 - 1e-2 * g / (np.abs(g) + 1e-8)
    np.testing.assert_allclose(p1, exp, rtol=1e-5, atol=1e-6)


def test_training_reduces_loss(net):
    X, T, M = data(4, B=32)
    first = net.train_step(X, T, M, 3e-3)
    for _ in range(200):
        last = net.train_step(X, T, M, 3e-3)
    assert last < 0.5 * first


def test_warm_start_pipeline_gain(gpu):
    """The whole learned warm start on a small family (4 orbitals, 64
    seeded instances, 48 training / 16 held out): plain GPU Newton from the
    standard guess generates the angles, the network is trained on the GPU
    (supervised, then on exact energies), and the proposals are judged by
    exact FP64 GPU energies with the standard guess as fallback.  What the
    network gains: the held-out starting energy gap to the best minimum
    found shrinks, and time/iterations to 1e-6 Ha of it are reported."""
    from ansatzevolve_amd.learned import run_pipeline
    s = run_pipeline([4], 64, epochs=100, energy_steps=200, log=lambda m: None)
    rows = s["rows"]
    assert len(rows) == 16
    for r in rows:
        # fallback rule: a proposal is used only if it does not start higher
        assert r["used_mlp"] == (r["E_mlp_start"] <= r["E_std_start"])
        assert r["gap_start"] <= r["gap_std"] + 1e-15
        assert r["gap_std"] >= -1e-6 and r["gap_start"] >= -1e-6
    assert s["energy_finetune_mean_E"][1] < s["energy_finetune_mean_E"][0]
    a = s["all"]
    print("\nwarm start (4 orbitals, 16 held out): %s" % {k: a[k] for k in a})
    assert a["proposals_used"] >= len(rows) // 2
    assert a["median_start_gap_chosen"] < a["median_start_gap_std"]


def test_warm_start_continuation_targets(gpu):
    """The one-parameter family (site spacing only) with training targets
    that follow one branch of minima by continuation (each training
    instance optimised from its spacing neighbour's optimum, the branch with
    the lower total energy of the two directions): the targets are smooth in
    the spacing, the network fits them closely, and its proposals start much
    nearer a minimum than the standard guess."""
    from ansatzevolve_amd.learned import run_pipeline
    s = run_pipeline([4], 48, epochs=150, energy_steps=0, family="line", targets_kind="continuation",
                     log=lambda m: None)
    rows = s["rows"]
    assert len(rows) == 12 and s["targets"] == "continuation"
    for r in rows:
        assert r["used_mlp"] == (r["E_mlp_start"] <= r["E_std_start"])
    lo, hi = s["supervised_loss"]
    assert hi < 0.2 * lo
    a = s["all"]
    print("\nwarm start, continuation targets (4 orbitals, 12 held out): %s" % {k: a[k] for k in a})
    assert a["median_start_gap_chosen"] < a["median_start_gap_std"]



def test_warmstart_generator_n8_fallback_rule(gpu):
    """The seeded-family pipeline at 8 orbitals: batched L-BFGS targets,
    canonical form, the width-256 network, proposals.  The fallback rule
    holds for every held-out instance (the proposal is used exactly when its
    exact energy is not above theta = 0's, and the chosen start's energy is
    the lower of the two), and the batched start energies equal
    single-context evaluations."""
    from ansatzevolve_amd import from_workload
    from ansatzevolve_amd.learned import canonical_angles
    from ansatzevolve_amd.warmstart import SlotModel, generate, instances, propose

    train, held = instances({8: (6, 6)})
    assert all(w["n"] == 8 and w["P"] == 84 for w in train + held)
    assert len({w["index"] for w in train + held}) == 12  # one seed per instance
    gen, _ = generate(train, max_evals=150, log=lambda m: None)
    for (x, E, ev, why), w in zip(gen, train):
        assert ev <= 150 and np.isfinite(E)
        with from_workload(w) as a:
            assert E <= a.energy(np.zeros(w["P"])) + 1e-12  # descent from the standard guess
            assert abs(a.energy(canonical_angles(x, 8, 4)) - E) <= 1e-10  # canonical form keeps the energy
    model = SlotModel()
    model.fit(train, [canonical_angles(g[0], 8, 4) for g in gen], epochs=20)
    props = model.predict(held)
    starts, E_std, E_prop, used, _ = propose(model, held)
    model.close()
    for w, p, s, es, ep, u in zip(held, props, starts, E_std, E_prop, used):
        with from_workload(w) as a:
            assert abs(a.energy(np.zeros(w["P"])) - es) <= 1e-12
            assert abs(a.energy(p) - ep) <= 1e-12
            assert u == (ep <= es)
            assert np.array_equal(s, p if u else np.zeros(w["P"]))
            assert abs(a.energy(s) - min(es, ep)) <= 1e-12
