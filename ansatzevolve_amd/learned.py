"""Learned warm start for the TUPS optimiser.

A 3-layer FP32 network (width 256, tanh) maps integral features to initial
angles.  Training and inference run on the GPU through the library's
mlp* entry points (hand-written HIP kernels); this module prepares
features and drives them.

Features of an n-orbital Hamiltonian, padded to nmax orbitals:
  * h, n x n;
  * the leading four factors of the two-electron supermatrix
    G[(pq),(rs)] = g_pqrs (symmetric positive semi-definite for the seeded
    integrals): eigenvectors of the four largest eigenvalues times the
    square root of the eigenvalue, each reshaped to n x n, sign fixed so that
    its largest-magnitude entry is positive.
Angles are predicted per parameter and padded to the largest parameter
count; a mask marks the real ones.

Exact FP64 energies and gradients of many instances (uneven sector sizes)
come from one multi-instance graph launch per step (MultiBatch): the energy
fine-tuning stage and the fallback judgement of the proposals.
"""
import ctypes

import numpy as np

from . import load_library

_fp = ctypes.POINTER(ctypes.c_float)


def _f(a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(_fp)


def features(h, g, nmax, nfactors=4):
    n = h.shape[0]
    G = np.asarray(g).reshape(n * n, n * n)
    G = 0.5 * (G + G.T)
    w, V = np.linalg.eigh(G)
    order = np.argsort(w)[::-1][:nfactors]
    out = [np.pad(h, ((0, nmax - n), (0, nmax - n)))]
    for i in order:
        f = V[:, i] * np.sqrt(max(w[i], 0.0))
        if f[np.argmax(np.abs(f))] < 0:
            f = -f
        out.append(np.pad(f.reshape(n, n), ((0, nmax - n), (0, nmax - n))))
    return np.concatenate([o.reshape(-1) for o in out])


class WarmStartMLP:
    """The network on the GPU (mlpCreate / mlpForward / mlpGradient / mlpTrainStep)."""

    def __init__(self, nin, nout, hidden=256, seed=0):
        self.lib = load_library()
        self.h = self.lib.mlpCreate(nin, hidden, nout, seed)
        if not self.h:
            raise RuntimeError("mlpCreate failed")
        self.nin, self.nout, self.hidden = nin, nout, hidden

    def close(self):
        if self.h:
            self.lib.mlpDestroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def nparams(self):
        return self.lib.mlpParameterCount(self.h)

    def get_params(self):
        p = np.zeros(self.nparams(), dtype=np.float32)
        assert self.lib.mlpGetParameters(self.h, p.ctypes.data_as(_fp)) == 0
        return p

    def set_params(self, p):
        p, pp = _f(p)
        assert self.lib.mlpSetParameters(self.h, pp) == 0

    def unpack(self, p=None):
        """(W0, b0, W1, b1, W2, b2) as numpy arrays."""
        p = self.get_params() if p is None else p
        i, hd, o = self.nin, self.hidden, self.nout
        sizes = [(i, hd), (hd,), (hd, hd), (hd,), (hd, o), (o,)]
        out, off = [], 0
        for s in sizes:
            cnt = int(np.prod(s))
            out.append(p[off:off + cnt].reshape(s))
            off += cnt
        return out

    def forward(self, X):
        X, xp = _f(X)
        Y = np.zeros((X.shape[0], self.nout), dtype=np.float32)
        assert self.lib.mlpForward(self.h, X.shape[0], xp, Y.ctypes.data_as(_fp)) == 0
        return Y

    def gradient(self, X, T, M):
        X, xp = _f(X)
        T, tp = _f(T)
        M, mp = _f(M)
        loss = ctypes.c_float()
        g = np.zeros(self.nparams(), dtype=np.float32)
        assert self.lib.mlpGradient(self.h, X.shape[0], xp, tp, mp, ctypes.byref(loss), g.ctypes.data_as(_fp)) == 0
        return loss.value, g

    def train_step(self, X, T, M, lr):
        X, xp = _f(X)
        T, tp = _f(T)
        M, mp = _f(M)
        loss = ctypes.c_float()
        assert self.lib.mlpTrainStep(self.h, X.shape[0], xp, tp, mp, ctypes.c_float(lr), ctypes.byref(loss)) == 0
        return loss.value


# ------------------------------------------------------------------ pipeline
# Instances of the learned warm-start workload: n spatial orbitals at half
# filling, TUPS with L = n / 2 layers, the alternating determinant as the
# start.  Three families:
#   "random": the seeded integral generator with its own seed per instance
#     (site spacing s ~ U(1.0, 2.5)): every instance has its own random
#     two-electron factors;
#   "smooth": the generator at one fixed seed per orbital count, so an
#     instance is fixed by two continuous parameters, the site spacing
#     s ~ U(1.0, 2.5) and a scale of the two-electron integrals
#     lambda ~ U(0.6, 1.4) (a chain stretched or compressed, with stronger or
#     weaker interaction);
#   "line": the smooth family at lambda = 1, so the site spacing alone fixes
#     an instance.
WARM_SEED = 20260500
WARM_SMOOTH_SEED = 20260700


def warm_instance(index, n, family="smooth"):
    from .workloads import hf_index, synthetic_integrals, tups_operators
    rng = np.random.default_rng(WARM_SEED + 1000 * n + index)
    spacing = float(rng.uniform(1.0, 2.5))
    if family == "random":
        lam = 1.0
        h, g = synthetic_integrals(n, WARM_SEED + 100000 + 1000 * n + index, spacing=spacing)
    elif family == "smooth":
        lam = float(rng.uniform(0.6, 1.4))
        h, g = synthetic_integrals(n, WARM_SMOOTH_SEED + n, spacing=spacing)
        g = lam * g
    elif family == "line":
        # one parameter: the smooth family's base integrals at site spacings
        # spread quasi-randomly over [1.0, 2.5] (held-out instances fall
        # between training ones)
        lam = 1.0
        spacing = 1.0 + 1.5 * float(np.mod((index + 1) * 0.6180339887498949, 1.0))
        h, g = synthetic_integrals(n, WARM_SMOOTH_SEED + n, spacing=spacing)
    else:
        raise ValueError("unknown instance family %r" % family)
    ops, order, scale = tups_operators(n, n // 2)
    return dict(n=n, nq=2 * n, h=h, g=g, ops=ops, order=order, scale=scale, start_idx=hf_index(n),
                P=int(order.max()) + 1, spacing=spacing, lam=lam, index=index)


def wrap_angles(t):
    """Angles into (-pi, pi]: every TUPS rotation is 2 pi periodic."""
    return np.pi - np.mod(np.pi - np.asarray(t, dtype=np.float64), 2 * np.pi)


# ------------------------------------------------------------ TUPS angles
# Parameter order of the TUPS workload (workloads.tups_operators): per layer,
# units on pairs (p, p+1), first p = 0, 2, 4, ... then p = 1, 3, ...
# (0-based lower orbital), three parameters per unit: theta_a (the two
# singles a+_p a_p+1 of both spins), theta_b (the pair double), theta_c (the
# singles again).
def tups_units(n, layers):
    """[(layer, lower orbital p, index of theta_a)] in parameter order."""
    out, par = [], 0
    for layer in range(layers):
        for start in (0, 1):
            for p in range(start, n - 1, 2):
                out.append((layer, p, par))
                par += 3
    return out


def output_slots(n, layers, nmax):
    """Network output index of every parameter: (layer, pair, kind) on the
    grid of the largest orbital count, so that a unit on the same pair in
    the same layer shares its outputs across orbital counts."""
    slots = np.zeros(3 * len(tups_units(n, layers)), dtype=np.int64)
    for layer, p, i in tups_units(n, layers):
        for k in range(3):
            slots[i + k] = (layer * (nmax - 1) + p) * 3 + k
    return slots


def _pair_single(t):
    """The two singles of a unit at angle t on the two-electron states of
    one spatial pair, basis (LL, LH, HL, HH) = (beta, alpha) in the lower (L)
    or upper (H) orbital.  Each electron's orbital turns by t:
    |L> -> cos t |L> + sin t |H>, |H> -> cos t |H> - sin t |L> (the rotation
    rule psi_b <- C psi_b + s S psi_b', psi_b' <- -s S psi_b + C psi_b' with
    b = H, b' = L and sign s = +1: the two orbitals are adjacent, so no
    occupied orbital lies between them)."""
    c, s = np.cos(t), np.sin(t)
    r = np.array([[c, -s], [s, c]])
    return np.kron(r, r)


def _pair_double(t):
    """The pair double (both electrons L <-> H): LL -> cos t LL + sin t HH,
    HH -> cos t HH - sin t LL (sign +1: its sign mask is empty)."""
    c, s = np.cos(t), np.sin(t)
    d = np.eye(4)
    d[0, 0], d[3, 0], d[0, 3], d[3, 3] = c, s, -s, c
    return d


def _unit_state(ta, tb, tc):
    x = np.array([1.0, 0.0, 0.0, 0.0])
    return _pair_single(tc) @ _pair_double(tb) @ _pair_single(ta) @ x


def _canonical_first_unit(ta, tb, tc):
    """A unit acting on its pair's start configuration (both electrons in
    the lower orbital) reaches the unit sphere of the 3-dimensional singlet
    space {LL, (LH + HL) / sqrt 2, HH}: three angles for two degrees of
    freedom, and the state's sign is free (it only flips the total state).
    Canonical form: theta_c = -theta_a = -t1 with t1 in (-pi/4, pi/4], the
    sign that makes the LL amplitude positive.  The singles rotate about the
    axis A = (LL + HH) / sqrt 2 (LL + HH is invariant), so S(-t1) D(t2) S(t1)
    is a rotation about the axis n = S(-t1) y, with n(t1) = (sin 2t1 / sqrt 2,
    cos 2t1, -sin 2t1 / sqrt 2); it carries LL to r exactly when n is
    orthogonal to both A and LL - r."""
    r = _unit_state(ta, tb, tc)
    if r[0] < 0 or (abs(r[0]) < 1e-14 and r[1] + r[2] < 0):
        r = -r
    r3 = np.array([r[0], (r[1] + r[2]) / np.sqrt(2.0), r[3]])
    A = np.array([1.0, 0.0, 1.0]) / np.sqrt(2.0)
    nv = np.cross(A, np.array([1.0, 0.0, 0.0]) - r3)
    if np.linalg.norm(nv) < 1e-12:
        nv = np.array([0.0, 1.0, 0.0])
    nv /= np.linalg.norm(nv)
    if nv[1] < 0 or (nv[1] == 0 and nv[0] < 0):
        nv = -nv
    t1 = 0.5 * np.arctan2(np.sqrt(2.0) * nv[0], nv[1])
    u = _pair_single(t1) @ np.array([1.0, 0.0, 0.0, 0.0])
    v = _pair_single(t1) @ r
    t2 = float(wrap_angles(np.arctan2(v[3], v[0]) - np.arctan2(u[3], u[0])))
    if np.max(np.abs(_unit_state(t1, t2, -t1) - r)) > 1e-9:
        return ta, tb, tc  # not reached (kept as a guard)
    return t1, t2, -t1


def canonical_angles(theta, n, layers):
    """Angles of the same energy in one canonical form per symmetry class:
    * every unit: (theta_a + pi, theta_c + pi) gives the same operator (the
      singles at theta + pi differ by (-1)^(electrons of that spin in the
      pair), which commutes with the whole unit), so theta_a is brought into
      (-pi/2, pi/2];
    * units of the first half-layer, acting on the alternating start
      determinant's pairs (lower orbital occupied in both spins): the form
      of _canonical_first_unit;
    * all angles in (-pi, pi]."""
    t = np.array(theta, dtype=np.float64).copy()
    for layer, p, i in tups_units(n, layers):
        ta, tb, tc = t[i], t[i + 1], t[i + 2]
        if layer == 0 and p % 2 == 0:
            ta, tb, tc = _canonical_first_unit(ta, tb, tc)
        k = np.round(ta / np.pi)
        ta, tc = ta - k * np.pi, tc - k * np.pi
        if ta <= -np.pi / 2:
            ta, tc = ta + np.pi, tc + np.pi
        t[i], t[i + 1], t[i + 2] = ta, tb, tc
    return wrap_angles(t)


def time_to(trace, target):
    """(seconds, Newton iterations) until the trace's energy <= target, or
    (inf, inf) if it never gets there.  Trace rows: seconds, energy, |g|,
    Hessian evaluations, energy evaluations."""
    for i, row in enumerate(trace):
        if row[1] <= target:
            return float(row[0]), i
    return float("inf"), float("inf")


class WarmStart:
    """Feature normalisation + the GPU network + angle padding for a set of
    orbital counts; predicts per-parameter angles for an instance."""

    def __init__(self, nmax, pmax, hidden=256, seed=0):
        self.nmax, self.pmax = nmax, pmax
        self.nin = 5 * nmax * nmax
        self.net = WarmStartMLP(self.nin, pmax, hidden=hidden, seed=seed)
        self.mu = np.zeros(self.nin)
        self.sd = np.ones(self.nin)

    def close(self):
        self.net.close()

    def inputs(self, insts):
        X = np.array([features(w["h"], w["g"], self.nmax) for w in insts])
        return ((X - self.mu) / self.sd).astype(np.float32)

    def fit(self, insts, targets, epochs=300, lr=1e-3, batch=64, seed=0):
        """Masked MSE on wrapped optimised angles, Adam on the GPU."""
        X = np.array([features(w["h"], w["g"], self.nmax) for w in insts])
        self.mu, self.sd = X.mean(0), X.std(0) + 1e-8
        Xn = ((X - self.mu) / self.sd).astype(np.float32)
        T = np.zeros((len(insts), self.pmax), np.float32)
        Mk = np.zeros_like(T)
        for i, t in enumerate(targets):
            T[i, :len(t)] = wrap_angles(t)
            Mk[i, :len(t)] = 1
        rng = np.random.default_rng(seed)
        losses = []
        for ep in range(epochs):
            perm = rng.permutation(len(insts))
            for b in range(0, len(perm), batch):
                j = perm[b:b + batch]
                losses.append(self.net.train_step(Xn[j], T[j], Mk[j], lr))
        return losses

    def predict(self, insts):
        Y = self.net.forward(self.inputs(insts)).astype(np.float64)
        return [Y[i, :w["P"]] for i, w in enumerate(insts)]


def solve(w, theta0, stop_energy=float("nan")):
    """Plain GPU Newton from theta0; returns (theta, E, trace)."""
    from . import from_workload
    with from_workload(w) as a:
        return a.optimise(np.asarray(theta0, dtype=np.float64), stop_energy=stop_energy)


def consistent_targets(insts, sols, neighbours=2, log=None):
    """Training targets that vary smoothly across the family: for every
    training instance the plain optimiser is also started from the optimised
    angles of its nearest training neighbours in (spacing, lambda) (same
    orbital count), and the lowest minimum found is kept.  The optimiser from
    the standard guess alone lands in different local minima on neighbouring
    instances, which no regression can fit."""
    out = [(np.asarray(t), float(E)) for t, E, _ in sols]
    params = np.array([[w["spacing"], w["lam"]] for w in insts])
    for i, w in enumerate(insts):
        same = [j for j in range(len(insts)) if j != i and insts[j]["n"] == w["n"]]
        if not same:
            continue
        d = np.linalg.norm(params[same] - params[i], axis=1)
        for j in [same[k] for k in np.argsort(d)[:neighbours]]:
            th, E, _ = solve(w, sols[j][0])
            if E < out[i][1] - 1e-9:
                out[i] = (np.asarray(th), float(E))
        if log and i % 20 == 0:
            log("[warmstart] consistent targets %d/%d" % (i + 1, len(insts)))
    return out


def continuation_targets(insts, sols, log=None):
    """Training targets along one branch of minima: per orbital count the
    training instances in order of site spacing, the first from its plain
    optimum, every next one optimised from the previous one's angles.  The
    optimiser from the standard guess lands in different minima on
    neighbouring instances; a branch followed by continuation varies
    smoothly with the spacing, so a regression can interpolate it.  The
    branch is followed from both ends of the spacing range; per orbital
    count the one with the lower total energy is kept (taking the lower of
    the two per instance would jump where they cross)."""
    out = [(np.asarray(t), float(E)) for t, E, _ in sols]
    for n in sorted({w["n"] for w in insts}):
        idx = sorted([i for i, w in enumerate(insts) if w["n"] == n], key=lambda i: insts[i]["spacing"])
        best = None
        for order in (idx, idx[::-1]):
            br = {order[0]: out[order[0]]}
            for a, b in zip(order[:-1], order[1:]):
                th, E, _ = solve(insts[b], br[a][0])
                br[b] = (np.asarray(th), float(E))
            total = sum(E for _, E in br.values())
            if best is None or total < best[0]:
                best = (total, br)
        for i, t in best[1].items():
            out[i] = t
        if log:
            log("[warmstart] continuation targets n=%d: %d instances, total E %.6f" % (n, len(idx), best[0]))
    return out


def start_energies(insts, starts):
    """Exact FP64 energies of candidate starts, starts[i] (k x P_i) for
    instance i: one launch over all instances (uneven sizes)."""
    from . import MultiBatch, from_workload
    ctxs = [from_workload(w) for w in insts]
    try:
        with MultiBatch(ctxs) as mb:
            for i, st in enumerate(starts):
                mb.upload(i, np.atleast_2d(st))
            mb.run(False)
            return [mb.download(i, with_gradient=False) for i in range(len(insts))]
    finally:
        for a in ctxs:
            a.close()


def finetune_energy(model, insts, steps=300, lr=3e-4, log=None):
    """Second training stage: minimise the mean exact FP64 energy of the
    network's proposals over the training instances.  dE/dtheta comes from
    the library's adjoint gradient on the GPU; it enters the network's
    backward pass through the masked-MSE kernel with the target
    t = y - (sum m) / (2 B) * dE/dtheta, whose loss gradient is exactly
    (1 / B) dE/dtheta."""
    from . import MultiBatch, from_workload
    engines = [from_workload(w) for w in insts]
    mb = MultiBatch(engines)
    X = model.inputs(insts)
    B = len(insts)
    Mk = np.zeros((B, model.pmax), np.float32)
    for i, w in enumerate(insts):
        Mk[i, :w["P"]] = 1
    msum = float(Mk.sum())
    hist = []
    try:
        for step in range(steps):
            Y = model.net.forward(X).astype(np.float64)
            G = np.zeros((B, model.pmax))
            Es = np.zeros(B)
            # every instance's exact energy and gradient: one launch over all
            for i, w in enumerate(insts):
                mb.upload(i, Y[i, :w["P"]][None, :])
            mb.run(True)
            for i, w in enumerate(insts):
                e, g = mb.download(i)
                Es[i], G[i, :w["P"]] = e[0], g[0]
            T = (Y - msum / (2.0 * B) * G).astype(np.float32)
            model.net.train_step(X, T, Mk, lr)
            hist.append(float(Es.mean()))
            if log and step % 50 == 0:
                log("[warmstart] energy fine-tuning step %d: mean E %.6f" % (step, hist[-1]))
    finally:
        mb.close()
        for a in engines:
            a.close()
    return hist


def run_pipeline(ns, count, train_frac=0.75, epochs=300, energy_steps=300, hidden=256, log=print,
                 family="smooth", neighbours=2, targets_kind="consistent"):
    """Generate (plain optimiser from the standard guess theta = 0; for the
    training set also from neighbours' optima, keeping the lowest minimum),
    train (supervised on the optimised angles, then energy fine-tuning), and
    evaluate time-to-1e-6 Ha on the held-out instances with the fallback
    rule: the network's proposal is used only if its exact energy is not
    above the standard guess's.  The target energy of an instance is the
    lower of the two runs' final energies plus 1e-6 Ha; a run that ends in a
    higher minimum never reaches it.  Returns a JSON-able summary."""
    import time
    insts = []
    for n in ns:
        insts += [warm_instance(i, n, family) for i in range(count)]
    rng = np.random.default_rng(WARM_SEED)
    order = rng.permutation(len(insts))
    insts = [insts[i] for i in order]
    ntr = int(train_frac * len(insts))
    train, held = insts[:ntr], insts[ntr:]
    t0 = time.time()
    plain = []
    for k, w in enumerate(insts):
        plain.append(solve(w, np.zeros(w["P"])))
        if k % 20 == 0:
            log("[warmstart] plain solves %d/%d (%.0f s)" % (k + 1, len(insts), time.time() - t0))
    targets = [(plain[k][0], plain[k][1]) for k in range(ntr)]
    if targets_kind == "continuation":
        targets = continuation_targets(train, plain[:ntr], log=log)
    elif neighbours:
        targets = consistent_targets(train, plain[:ntr], neighbours=neighbours, log=log)
    improved = int(sum(targets[k][1] < plain[k][1] - 1e-9 for k in range(ntr)))
    gen_s = time.time() - t0
    model = WarmStart(max(ns), max(w["P"] for w in insts), hidden=hidden)
    t1 = time.time()
    losses = model.fit(train, [t for t, _ in targets], epochs=epochs)
    log("[warmstart] supervised stage: %d steps, loss %.4f -> %.4f" % (len(losses), losses[0], losses[-1]))
    ehist = finetune_energy(model, train, steps=energy_steps, log=log) if energy_steps else []
    train_s = time.time() - t1
    t2 = time.time()
    props = model.predict(held)
    Eall = start_energies(held, [np.stack([np.zeros(w["P"]), p]) for w, p in zip(held, props)])
    choose_s = (time.time() - t2) / max(1, len(held))  # inference + exact start energies, per instance
    rows = []
    for k, (w, p) in enumerate(zip(held, props)):
        Es = Eall[k]
        use = bool(Es[1] <= Es[0])
        th1, E1, tr1 = solve(w, p if use else np.zeros(w["P"]))
        th0, E0, tr0 = plain[ntr + k]
        best = min(E0, E1)
        t_plain, i_plain = time_to(tr0, best + 1e-6)
        t_mlp, i_mlp = time_to(tr1, best + 1e-6)
        if k % 10 == 0:
            log("[warmstart] held-out solves %d/%d" % (k + 1, len(held)))
        rows.append(dict(n=w["n"], index=w["index"], E_std_start=float(Es[0]), E_mlp_start=float(Es[1]),
                         used_mlp=use, E_plain=float(E0), E_mlp=float(E1), t_plain=t_plain, t_mlp=t_mlp + choose_s,
                         it_plain=i_plain, it_mlp=i_mlp, t_plain_conv=float(tr0[-1][0]) if len(tr0) else 0.0,
                         t_mlp_conv=float(tr1[-1][0]) + choose_s if len(tr1) else 0.0,
                         gap_std=float(Es[0] - best), gap_start=float(min(Es[0], Es[1]) - best)))
    model.close()

    def med(key, sel=None):
        v = np.array([r[key] for r in rows if sel is None or sel(r)], dtype=float)
        return float(np.median(v)) if len(v) else None

    def summ(sel=None):
        rs = [r for r in rows if sel is None or sel(r)]
        return dict(count=len(rs), median_t_plain=med("t_plain", sel), median_t_mlp=med("t_mlp", sel),
                    median_t_plain_conv=med("t_plain_conv", sel), median_t_mlp_conv=med("t_mlp_conv", sel),
                    median_it_plain=med("it_plain", sel), median_it_mlp=med("it_mlp", sel),
                    plain_misses=int(sum(np.isinf(r["t_plain"]) for r in rs)),
                    mlp_misses=int(sum(np.isinf(r["t_mlp"]) for r in rs)),
                    proposals_used=int(sum(r["used_mlp"] for r in rs)),
                    median_start_gap_std=med("gap_std", sel), median_start_gap_chosen=med("gap_start", sel))
    summary = dict(ns=list(ns), family=family, neighbours=neighbours, targets=targets_kind,
                   targets_improved=improved,
                   instances=len(insts), train=ntr, held_out=len(held), generate_s=gen_s,
                   train_s=train_s, supervised_loss=[float(losses[0]), float(losses[-1])],
                   energy_finetune_mean_E=[ehist[0], ehist[-1]] if ehist else None, all=summ(), rows=rows)
    for n in ns:
        summary["n%d" % n] = summ(lambda r, n=n: r["n"] == n)
    return summary


def _main():
    import argparse
    import json
    import sys
    p = argparse.ArgumentParser(description="learned warm start: generate, train, evaluate")
    p.add_argument("--ns", type=int, nargs="+", default=[6, 8])
    p.add_argument("--count", type=int, default=80, help="instances per orbital count")
    p.add_argument("--epochs", type=int, default=300, help="supervised epochs")
    p.add_argument("--energy-steps", type=int, default=0,
                   help="energy fine-tuning steps (it moves proposals off the branch the targets follow)")
    p.add_argument("--family", default="line", choices=["smooth", "random", "line"], help="instance family")
    p.add_argument("--targets", default="continuation", choices=["consistent", "continuation"],
                   help="training targets: lowest of neighbours' optima, or one branch by continuation")
    p.add_argument("--neighbours", type=int, default=2,
                   help="training targets: also start from this many neighbours' optima (0: plain only)")
    p.add_argument("--out", default="")
    a = p.parse_args()
    s = run_pipeline(a.ns, a.count, epochs=a.epochs, energy_steps=a.energy_steps, family=a.family,
                     neighbours=a.neighbours, targets_kind=a.targets,
                     log=lambda m: print(m, file=sys.stderr, flush=True))
    txt = json.dumps(s)
    if a.out:
        with open(a.out, "w") as f:
            f.write(txt)
    brief = {k: v for k, v in s.items() if k != "rows"}
    print(json.dumps(brief))


if __name__ == "__main__":
    _main()
