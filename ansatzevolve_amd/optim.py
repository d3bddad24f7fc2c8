"""Gradient-based optimiser (L-BFGS) over the GPU energy and adjoint gradient.

The reference's plain optimiser is Newton with exact Hessians
(src/tupsquantities.cpp:704-1236, restated in csrc/newton.cpp).  A Hessian
costs 2 ms, 12 ms and 250 ms at 8, 10 and 12 orbitals (TUPS L = n / 2) on
one MI355X against 0.6, 1.1 and 3.8 ms for an energy + gradient, and Newton
needs several hundred iterations from the standard guess, so optimising
hundreds of Hamiltonians (the learned warm start's training data) uses this
quasi-Newton method instead:

* two-loop L-BFGS direction from the last `memory` curvature pairs (pairs
  with s.y <= 1e-12 |s| |y| are skipped);
* backtracking line search: step 1 (the first iteration: 0.1 / |g|_inf,
  at most 1), Armijo constant 1e-4, halving up to 40 times; a failed search
  drops the history and retries along -g;
* stop when |g|_inf < gtol, when the energy has decreased by less than
  `stall` over the last `stall_window` accepted iterations, or after
  `max_evals` energy + gradient evaluations.

One optimisation is a state machine (`Lbfgs`): it proposes a point, takes
its energy and gradient, proposes the next.  `minimise` drives one instance
through its context; `minimise_batched` drives many instances (uneven sector
sizes) together, one multi-instance graph launch per round, every instance
evaluating its own next point.  The two give the same iterates (batched and
single evaluations agree to 1e-12).
"""
import time

import numpy as np


class Lbfgs:
    """One L-BFGS optimisation.  trace rows: (seconds, energy, |g|_inf,
    evaluations) at every accepted iterate, the start included."""

    def __init__(self, x0, memory=20, gtol=1e-7, stall=1e-11, stall_window=20, max_evals=4000):
        self.x = np.array(x0, dtype=np.float64)
        self.m, self.gtol, self.stall, self.max_evals = memory, gtol, stall, max_evals
        self.window = stall_window
        self.S, self.Y = [], []
        self.E = self.g = None
        self.evals = 0
        self.trial = self.x.copy()
        self.d = None
        self.alpha = 1.0
        self.tries = 0
        self.done = False
        self.reason = ""
        self.trace = []
        self.t0 = None

    # the point whose energy and gradient are wanted next
    def point(self):
        return self.trial

    def _direction(self):
        q = -self.g.copy()
        a = []
        for s, y in zip(reversed(self.S), reversed(self.Y)):
            r = 1.0 / np.dot(y, s)
            ai = r * np.dot(s, q)
            a.append((ai, r, s, y))
            q -= ai * y
        if self.S:
            s, y = self.S[-1], self.Y[-1]
            q *= np.dot(s, y) / np.dot(y, y)
        for ai, r, s, y in reversed(a):
            b = r * np.dot(y, q)
            q += (ai - b) * s
        return q

    def _new_search(self, first=False):
        self.d = self._direction()
        if np.dot(self.d, self.g) >= 0:  # not a descent direction: restart
            self.S, self.Y = [], []
            self.d = -self.g.copy()
        gi = float(np.max(np.abs(self.g)))
        self.alpha = min(1.0, 0.1 / gi) if first or not self.S else 1.0
        self.tries = 0
        self.trial = self.x + self.alpha * self.d

    def _accept(self, E, g):
        s = self.trial - self.x
        y = g - self.g
        sy = float(np.dot(s, y))
        if sy > 1e-12 * np.linalg.norm(s) * np.linalg.norm(y):
            self.S.append(s)
            self.Y.append(y)
            if len(self.S) > self.m:
                self.S.pop(0)
                self.Y.pop(0)
        self.x, self.E, self.g = self.trial.copy(), E, g

    def _record(self):
        self.trace.append((time.perf_counter() - self.t0, self.E, float(np.max(np.abs(self.g))), self.evals))

    def _check_done(self):
        if np.max(np.abs(self.g)) < self.gtol:
            self.done, self.reason = True, "gradient"
        elif len(self.trace) > self.window and self.trace[-1 - self.window][1] - self.E < self.stall:
            self.done, self.reason = True, "stalled"
        elif self.evals >= self.max_evals:
            self.done, self.reason = True, "evaluations"

    def tell(self, E, g):
        """Energy and gradient at point()."""
        if self.t0 is None:
            self.t0 = time.perf_counter()
        self.evals += 1
        g = np.asarray(g, dtype=np.float64)
        if self.E is None:  # the start
            self.E, self.g = float(E), g.copy()
            self._record()
            self._check_done()
            if not self.done:
                self._new_search(first=True)
            return
        if np.isfinite(E) and E <= self.E + 1e-4 * self.alpha * float(np.dot(self.g, self.d)):
            self._accept(float(E), g)
            self._record()
            self._check_done()
            if not self.done:
                self._new_search()
            return
        self.tries += 1
        if self.tries > 40:
            if self.S:  # drop the history, search along -g
                self.S, self.Y = [], []
                self._new_search(first=True)
            else:
                self.done, self.reason = True, "line search"
            return
        if self.evals >= self.max_evals:
            self.done, self.reason = True, "evaluations"
            return
        self.alpha *= 0.5
        self.trial = self.x + self.alpha * self.d

    def result(self):
        return self.x.copy(), float(self.E), np.array(self.trace)


def minimise(ctx, x0, **kw):
    """L-BFGS on one AnsatzSynth context from per-parameter angles x0;
    returns (angles, energy, trace, reason)."""
    st = Lbfgs(x0, **kw)
    st.t0 = time.perf_counter()
    while not st.done:
        E, g = ctx.energy_gradient(st.point())
        st.tell(E, g)
    x, E, tr = st.result()
    return x, E, tr, st.reason


def minimise_batched(ctxs, x0s, log=None, **kw):
    """L-BFGS on many contexts at once: every round one multi-instance graph
    launch evaluates each running instance's next point.  The launch is
    rebuilt over the running instances when a quarter of them have
    finished.  Returns a list of (angles, energy, trace, reason); trace
    times are those of the whole batch."""
    from . import MultiBatch
    states = [Lbfgs(x, **kw) for x in x0s]
    t0 = time.perf_counter()
    for s in states:
        s.t0 = t0
    rounds = 0
    last = t0
    active = [i for i, s in enumerate(states) if not s.done]
    while active:
        with MultiBatch([ctxs[i] for i in active]) as mb:
            while True:
                for k, i in enumerate(active):
                    mb.upload(k, states[i].point()[None, :])
                mb.run(True)
                for k, i in enumerate(active):
                    E, G = mb.download(k)
                    if not states[i].done:
                        states[i].tell(float(E[0]), G[0])
                rounds += 1
                running = [i for i in active if not states[i].done]
                now = time.perf_counter()
                if log and now - last > 30:
                    last = now
                    log("[lbfgs] round %d: %d of %d running (%.0f s)" % (rounds, len(running), len(states), now - t0))
                if len(running) < 0.75 * len(active) or not running:
                    break
        active = running
    return [s.result() + (s.reason,) for s in states]
