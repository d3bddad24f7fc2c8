"""Writes reference_vectors.npz: energies, gradients, Hessians and states
computed by the reference's own CPU build (oracle/_ref/libref_ansatz.so,
built from the reference sources by `make -C oracle ref`).

Inputs are regenerated from ansatzevolve_amd.workloads (seeded); each
integral set is stored with a checksum so a changed generator is caught.
Run in the build container:  python tests/golden/make_golden.py [--big]
    python tests/golden/make_golden.py --cfg4   (adds only the 14-orbital
    workload, 11,778,624 determinants, to the existing file; tens of minutes)
    python tests/golden/make_golden.py --proj   (adds the projected Hessian
    and iterative-optimiser cases)
    python tests/golden/make_golden.py --complex   (adds the complex-mode
    cases, from the reference built with its complex mode: make -C oracle
    ref-complex)
"""
import hashlib
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from ansatzevolve_amd.workloads import (hf_index, make_workload, seeded_angles, synthetic_integrals,  # noqa: E402
                                        tups_operators)
from oracle.oracle import OracleEngine, RefEngine, RefFusedEvolve  # noqa: E402

OUT = os.path.join(HERE, "reference_vectors.npz")


def digest(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a, dtype=np.float64).tobytes())
    return h.hexdigest()[:16]


def integral_case(n, layers, seed, nsets, hessian=True):
    h, g = synthetic_integrals(n, seed)
    ops, order, scale = tups_operators(n, layers)
    P = int(order.max()) + 1
    thetas = seeded_angles(P, seed + 7, count=nsets)
    r = RefEngine(2 * n, [hf_index(n)], [1.0], ops, order, scale, h=h, g=g)
    E, G, H = [], [], []
    for th in thetas:
        E.append(r.energy(th))
        G.append(r.gradient(th))
        if hessian:
            H.append(r.hessian(th))
    r.close()
    out = {"theta": thetas, "E": np.array(E), "grad": np.array(G), "digest": digest(h, g)}
    if hessian:
        out["hess"] = np.array(H)
    return out


def cfg4_only():
    """E and gradient of the 14-orbital workload at theta_0 and one more
    seeded angle vector, by the reference's matrix-free path (its
    FusedEvolve built as its benchmark mode does), merged into OUT."""
    old = np.load(OUT, allow_pickle=False)
    data = {key: old[key] for key in old.files}
    meta = json.loads(str(old["meta_json"]))
    w = make_workload(4)
    fe = RefFusedEvolve(w["nq"], w["nup"], w["ndown"], [w["start_idx"]], [1.0], w["ops"], w["order"],
                        w["scale"], w["h"], w["g"])
    th = np.stack([w["theta"], seeded_angles(w["P"], w["seed"] + 2)])
    E, G, T = [], [], []
    for t in th:
        t0 = time.time()
        e, g, tt = fe.egrad(w["scale"] * t[w["order"]], reps=1)
        E.append(e)
        G.append(g)
        T.append(tt)
        print("cfg4 E=%.15f %.1fs" % (e, time.time() - t0), file=sys.stderr, flush=True)
    data["cfg4_theta"] = th
    data["cfg4_E"] = np.array(E)
    data["cfg4_grad"] = np.array(G)
    meta["cfg4"] = {"config": 4, "seed": w["seed"], "digest": digest(w["h"], w["g"]),
                    "ref_egrad_seconds": T, "ref_threads": fe.threads()}
    fe.close()
    data["meta_json"] = np.array(json.dumps(meta))
    np.savez_compressed(OUT, **data)
    print("wrote", OUT, file=sys.stderr)


def proj_cases():
    """The reference build's FusedEvolve::evolveHessianProj (random
    projection vectors over the sector) and TUPSQuantities::iterativeTups
    (from small seeded angles) on integral workloads, merged into OUT."""
    old = np.load(OUT, allow_pickle=False)
    data = {key: old[key] for key in old.files}
    meta = json.loads(str(old["meta_json"]))
    for name, n, layers, seed in (("projn4L2", 4, 2, 2101), ("projn6L2", 6, 2, 2102)):
        h, g = synthetic_integrals(n, seed)
        ops, order, scale = tups_operators(n, layers)
        P = int(order.max()) + 1
        fe = RefFusedEvolve(2 * n, n // 2, n // 2, [hf_index(n)], [1.0], ops, order, scale, h, g)
        o = OracleEngine(2 * n, [hf_index(n)], [1.0], ops, order, scale)
        st = o.states()
        rng = np.random.default_rng(seed + 5)
        thetas = seeded_angles(P, seed + 7, count=2)
        projs = rng.standard_normal((2, len(st)))
        H, G, V = [], [], []
        for th, pr in zip(thetas, projs):
            full = np.zeros(1 << (2 * n))
            full[st] = pr
            hh, gg, vv = fe.hessian_proj(scale * th[order], full)
            H.append(hh)
            G.append(gg)
            V.append(vv)
        data[name + "_theta"], data[name + "_proj"] = thetas, projs
        data[name + "_hess"], data[name + "_grad"], data[name + "_ov"] = np.array(H), np.array(G), np.array(V)
        # the iterative optimiser from small angles
        th0 = 0.3 * seeded_angles(P, seed + 1)
        x, E = fe.iterative(scale * th0[order])
        data[name + "_iter_start"], data[name + "_iter_angles"], data[name + "_iter_E"] = th0, x, np.array([E])
        meta[name] = {"n": n, "layers": layers, "seed": seed, "digest": digest(h, g)}
        fe.close()
        print(name, "iterative E %.15f" % E, file=sys.stderr, flush=True)
    data["meta_json"] = np.array(json.dumps(meta))
    np.savez_compressed(OUT, **data)
    print("wrote", OUT, file=sys.stderr)


def complex_cases():
    """The reference build in its complex mode (oracle/_ref/
    libref_ansatz_complex.so): the test.F90 set-up, the same with number
    operators appended (one per spin-orbital, two pair doubles, one
    out-of-order double that never acts) and a complex start, and number
    operators interleaved between the excitations with shared parameters.
    Energies, gradients and final states, merged into OUT."""
    from oracle.oracle import RefComplex
    old = np.load(OUT, allow_pickle=False)
    data = {key: old[key] for key in old.files}
    with open(os.path.join(HERE, "testf90_kat.json")) as f:
        k = json.load(f)
    ops1, order1 = np.array(k["operators_1based"]), np.array(k["order_1based"])
    ti, tj = np.array(k["ham_i_1based"]) - 1, np.array(k["ham_j_1based"]) - 1
    tv = np.array(k["ham_coeffs"])
    s0 = k["initial_index_1based"] - 1
    extra = np.array([[p, p, 0, 0] for p in range(1, 9)] + [[1, 5, 1, 5], [2, 6, 2, 6], [5, 1, 1, 5]])
    inter_ops, inter_ord = [], []
    for j, (op, od) in enumerate(zip(ops1, order1)):
        inter_ops.append(list(op))
        inter_ord.append(od)
        if j % 3 == 1:
            inter_ops.append([1 + (j % 8), 1 + (j % 8), 0, 0])
            inter_ord.append(od)
    inter_ops.append([3, 7, 3, 7])
    inter_ord.append(order1.max() + 1)
    cases = [("cplx_kat", ops1, order1, [s0], [1.0]),
             ("cplx_num", np.concatenate([ops1, extra]),
              np.concatenate([order1, order1.max() + 1 + np.arange(len(extra))]), [s0, 172], [1.0, 0.3j]),
             ("cplx_mix", np.array(inter_ops), np.array(inter_ord), [s0, 172], [1.0, 0.5 - 0.2j])]
    for name, o1, od1, idx, co in cases:
        co = np.array(co, dtype=complex)
        co = co / np.linalg.norm(co)
        r = RefComplex(8, idx, co, o1 - 1, od1 - 1, None, ti, tj, tv)
        P = int(od1.max())
        th = [np.array(k["angles_hf"] + [0.0] * (P - 9)), np.array(k["angles_min"] + [0.0] * (P - 9))] + \
             list(seeded_angles(P, 4343, count=2))
        data[name + "_ops1"], data[name + "_order1"] = o1, od1
        data[name + "_start_idx0"], data[name + "_start_re"], data[name + "_start_im"] = \
            np.array(idx), co.real.copy(), co.imag.copy()
        data[name + "_theta"] = np.array(th)
        data[name + "_E"] = np.array([r.energy(t) for t in th])
        data[name + "_grad"] = np.array([r.gradient(t) for t in th])
        st = np.array([r.state(t) for t in th])
        data[name + "_state_re"], data[name + "_state_im"] = st.real.copy(), st.imag.copy()
        r.close()
        print(name, data[name + "_E"], file=sys.stderr, flush=True)
    np.savez_compressed(OUT, **data)
    print("wrote", OUT, file=sys.stderr)


def main():
    if "--complex" in sys.argv:
        complex_cases()
        return
    if "--cfg4" in sys.argv:
        cfg4_only()
        return
    if "--proj" in sys.argv:
        proj_cases()
        return
    big = "--big" in sys.argv
    data = {}
    meta = {}
    # test.F90 set-up (triplet Hamiltonian), two KAT angle sets + two random
    with open(os.path.join(HERE, "testf90_kat.json")) as f:
        k = json.load(f)
    ops = np.array(k["operators_1based"]) - 1
    order = np.array(k["order_1based"]) - 1
    trip = (np.array(k["ham_i_1based"]) - 1, np.array(k["ham_j_1based"]) - 1, np.array(k["ham_coeffs"]))
    r = RefEngine(8, [k["initial_index_1based"] - 1], [1.0], ops, order, triplets=trip)
    thetas = [np.array(k["angles_hf"]), np.array(k["angles_min"])] + list(seeded_angles(9, 4242, count=2))
    data["kat_theta"] = np.array(thetas)
    data["kat_E"] = np.array([r.energy(t) for t in thetas])
    data["kat_grad"] = np.array([r.gradient(t) for t in thetas])
    data["kat_hess"] = np.array([r.hessian(t) for t in thetas])
    data["kat_state"] = np.array([r.state(t, 256) for t in thetas])
    r.close()
    # integral Hamiltonians, TUPS, HF start
    cases = [("n4L2", 4, 2, 1001, 3, True), ("n6L2", 6, 2, 1006, 2, True), ("n8L2", 8, 2, 1008, 2, True),
             ("n6L3", 6, 3, 1016, 2, True)]
    for name, n, layers, seed, nsets, hess in cases:
        t0 = time.time()
        c = integral_case(n, layers, seed, nsets, hess)
        for key, v in c.items():
            if key == "digest":
                meta[name] = {"n": n, "layers": layers, "seed": seed, "digest": v}
            else:
                data[name + "_" + key] = v
        print(name, "%.1fs" % (time.time() - t0), file=sys.stderr)
    # the 10-orbital workload (63,504 determinants)
    t0 = time.time()
    w = make_workload(2)
    r = RefEngine(w["nq"], [w["start_idx"]], [1.0], w["ops"], w["order"], w["scale"], h=w["h"], g=w["g"])
    th2 = np.stack([w["theta"], seeded_angles(w["P"], w["seed"] + 2)])
    data["cfg2_theta"] = th2
    data["cfg2_E"] = np.array([r.energy(t) for t in th2])
    data["cfg2_grad"] = np.array([r.gradient(t) for t in th2])
    data["cfg2_hess"] = r.hessian(th2[0])[None]
    r.close()
    meta["cfg2"] = {"config": 2, "seed": w["seed"], "digest": digest(w["h"], w["g"])}
    print("cfg2 %.1fs" % (time.time() - t0), file=sys.stderr)
    if big:
        # the 12-orbital workload (853,776 determinants): the reference's
        # matrix-free path, FusedEvolve built as its benchmark mode does
        t0 = time.time()
        w = make_workload(3)
        fe = RefFusedEvolve(w["nq"], w["nup"], w["ndown"], [w["start_idx"]], [1.0], w["ops"], w["order"],
                            w["scale"], w["h"], w["g"])
        x = w["scale"] * w["theta"][w["order"]]
        E, g, t = fe.egrad(x, reps=1)
        data["cfg3_theta"] = w["theta"][None]
        data["cfg3_E"] = np.array([E])
        data["cfg3_grad"] = g[None]
        meta["cfg3"] = {"config": 3, "seed": w["seed"], "digest": digest(w["h"], w["g"]),
                        "ref_egrad_seconds": t, "ref_threads": fe.threads()}
        fe.close()
        print("cfg3 %.1fs (egrad %.2fs)" % (time.time() - t0, t), file=sys.stderr)
    data["meta_json"] = np.array(json.dumps(meta))
    if os.path.exists(OUT) and not big:
        # keep previously generated large cases
        old = np.load(OUT, allow_pickle=False)
        for key in old.files:
            if key.startswith(("cfg3", "cfg4", "proj", "cplx")) and key not in data:
                data[key] = old[key]
        om = json.loads(str(old["meta_json"]))
        for c in [k for k in om if k.startswith(("cfg3", "cfg4", "proj"))]:
            if c in om and c not in meta:
                meta[c] = om[c]
        data["meta_json"] = np.array(json.dumps(meta))
    np.savez_compressed(OUT, **data)
    print("wrote", OUT, file=sys.stderr)


if __name__ == "__main__":
    main()
