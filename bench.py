"""Benchmark: TUPS energy + adjoint-gradient evaluations per second.

Workload: 12 spatial orbitals / 12 electrons, S_z = 0 sector of 853,776
determinants, seeded synthetic integrals, TUPS with 6 layers (198
parameters, 330 operators), the automatic Hamiltonian form (matrix-free at
12 orbitals; --storage 1 forces the explicit rows, 18.6 GB).  One
step = one batch of 64 angle vectors (theta_b = theta_0 + 1e-2 U(-pi, pi))
evaluated to energy + gradient.  Inputs (Hamiltonian, psi_0, the angle
batch) are resident in HBM before the timed region; a step enqueues device
work only.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--nvec 64]
                    [--config 3] [--storage 0|1|2] [--no-cpu-baseline]
                    [--no-autotune] [--calib-copy]

Prints one JSON line.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
# LDS: 256 CUs x 128 bytes per clock x 2.4 GHz peak engine clock
LDS_PEAK_GBS = 256 * 128 * 2.4
# results must agree with the reference build / the single-vector path
E_TOL, G_TOL, SAME_TOL = 1e-10, 1e-9, 1e-10


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--nvec", type=int, default=64)
    p.add_argument("--config", type=int, default=3)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--no-warm-start", action="store_true",
                   help="skip the short learned-vs-plain time-to-1e-6 Ha run reported beside the throughput")
    p.add_argument("--cpu-threads", type=int, default=0)
    p.add_argument("--storage", type=int, default=0,
                   help="Hamiltonian form: 0 automatic, 1 explicit rows, 2 matrix-free sigma")
    p.add_argument("--no-autotune", action="store_true",
                   help="keep the library's default sweep kind, fusion depth and batch layout")
    p.add_argument("--lane-groups", type=int, default=0,
                   help="force the lane-layout orbit groups per pass (1 or 2) after autotuning")
    p.add_argument("--autotune-cache", default="",
                   help="autotuner record file: a recorded choice for this shape is re-applied without measuring, "
                        "a measured one is appended")
    p.add_argument("--no-extras", action="store_true",
                   help="skip the Hessian timings and the multi-instance launch comparison reported after the "
                        "timed region")
    p.add_argument("--eager", action="store_true",
                   help="launch every kernel directly instead of replaying hipGraphs (for profiler counter passes)")
    p.add_argument("--no-exchange", action="store_true",
                   help="both same-spin DGEMMs even where the batch is symmetric under the exchange of the alpha "
                        "and beta strings (engine option 33 = 0)")
    p.add_argument("--calib-copy", action="store_true",
                   help="run one plain 2 GiB device copy first (HBM counter calibration under rocprofv3 --pmc)")
    return p.parse_args()


def cpu_baseline(config, threads, reps=3):
    """The reference's own CPU build (oracle/_ref, compiled from its sources):
    FusedEvolve::evolve + evolveDerivative(&E) on one angle vector, its
    matrix-free Hamiltonian path, in a child process with `threads` threads.
    Set-up (integral files read, operator list built) and one warm-up E+grad
    (the reference builds its fused-block cache lazily on the first evolve,
    src/fusedevolve.cpp:2111-2112) are timed apart; the rate is the median of
    `reps` further E+grads, as the reference's benchmark mode repeats its
    evaluation (src/main.cpp:404-435)."""
    code = r"""
import json, sys, time, os
sys.path.insert(0, %r)
import numpy as np
from ansatzevolve_amd.workloads import make_workload
from oracle.oracle import RefFusedEvolve
w = make_workload(%d)
t0 = time.time()
fe = RefFusedEvolve(w["nq"], w["nup"], w["ndown"], [w["start_idx"]], [1.0], w["ops"], w["order"],
                    w["scale"], w["h"], w["g"])
setup = time.time() - t0
x = w["scale"] * w["theta"][w["order"]]
E, g, t_first = fe.egrad(x, reps=1)
times = []
for r in range(%d):
    E, g, t = fe.egrad(x, reps=1)
    times.append(t)
    print("[cpu] E+grad %%d: %%.2f s" %% (r, t), file=sys.stderr, flush=True)
print("RESULT " + json.dumps({"times": times, "first": t_first, "setup": setup, "threads": fe.threads(), "E": E}))
""" % (ROOT, config, reps)
    env = dict(os.environ)
    env["OMP_NUM_THREADS"] = str(threads)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=1200)
    for line in r.stdout.splitlines():
        if line.startswith("RESULT "):
            return json.loads(line[7:])
    raise RuntimeError("reference CPU run failed: " + r.stderr[-2000:])


def extra_measurements(a, w):
    """After the timed region: per-parameter Hessian wall time (one call
    after a warm-up) at the benchmark configuration and at 10 orbitals, and
    32 independent 8-orbital instances evaluated (E + gradient, one angle
    vector each) as one multi-instance graph launch against one
    single-instance call after another."""
    from ansatzevolve_amd import MultiBatch, from_workload
    from ansatzevolve_amd.learned import warm_instance
    from ansatzevolve_amd.workloads import make_workload
    out = {}
    hs = {}
    for cfg, ctx, wl in ((3, a, w), (2, None, None)):
        own = ctx is None
        if own:
            wl = make_workload(cfg)
            ctx = from_workload(wl)
        ctx.hessian(wl["theta"])
        t = time.perf_counter()
        ctx.hessian(wl["theta"] + 1e-3)
        hs["config%d" % cfg] = {"ms": 1e3 * (time.perf_counter() - t), "parameters": wl["P"],
                                "determinants": ctx.dimensions()[0]}
        if own:
            ctx.close()
    out["hessian"] = hs
    insts = [warm_instance(i, 8, "random") for i in range(32)]
    ctxs = [from_workload(x) for x in insts]
    try:
        th = [np.full(x["P"], 0.05) for x in insts]
        for c, t0 in zip(ctxs, th):
            c.energy_gradient(t0)
        K = 20
        t = time.perf_counter()
        for r in range(K):
            for c, t0 in zip(ctxs, th):
                c.energy_gradient(t0 + 1e-4 * r)
        seq = time.perf_counter() - t
        worst = 0.0
        with MultiBatch(ctxs) as mb:
            for r in range(K + 2):
                if r == 2:
                    t = time.perf_counter()
                for i, t0 in enumerate(th):
                    mb.upload(i, (t0 + 1e-4 * r)[None, :])
                mb.run(True)
                res = [mb.download(i) for i in range(len(ctxs))]
            multi = time.perf_counter() - t
            for i, (c, t0) in enumerate(zip(ctxs, th)):  # last round against single calls
                E, g = c.energy_gradient(t0 + 1e-4 * (K + 1))
                worst = max(worst, abs(float(res[i][0][0]) - E), float(np.max(np.abs(res[i][1][0] - g))))
        out["multi_instance"] = {
            "instances": "32 seeded 8-orbital Hamiltonians (4,900 determinants, 84 parameters), one angle vector "
                         "each, energy + gradient",
            "one_launch_egrad_per_s": 32 * K / multi, "one_after_another_egrad_per_s": 32 * K / seq,
            "max_diff_vs_single": worst}
    finally:
        for c in ctxs:
            c.close()
    return out


def calib_copy(nbytes):
    """hipMalloc two buffers, fill one, copy it to the other (device to
    device), free both; returns 0 on success."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    size = ctypes.c_size_t(nbytes)
    if hip.hipMalloc(ctypes.byref(a), size) or hip.hipMalloc(ctypes.byref(b), size):
        return 1
    rc = hip.hipMemset(a, 0, size) or hip.hipMemcpy(b, a, size, 3) or hip.hipDeviceSynchronize()  # 3: device to device
    hip.hipFree(a)
    hip.hipFree(b)
    return int(rc)


def measured_traffic(kernel, config, nvec):
    """HBM bytes per launch of `kernel` from the counter passes recorded in
    profiles/ (FETCH_SIZE and WRITE_SIZE, scaled by a plain copy of known
    size), for the workload they were taken on; None when not recorded."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "hbm_traffic_*.md"))):
        text = open(path).read()
        if "```json" not in text:
            continue
        rec = json.loads(text.split("```json", 1)[1].split("```", 1)[0])
        if rec.get("config") != config or rec.get("nvec") != nvec:
            continue
        for name, t in rec.get("kernels", {}).items():
            if kernel.startswith(name):
                return {"bytes_per_launch": t["fetch_bytes"] + t["write_bytes"], "fetch_bytes": t["fetch_bytes"],
                        "write_bytes": t["write_bytes"], "source": os.path.relpath(path, ROOT)}
    return None


def window_vm(a, opts, N):
    """True when the vector-major sweeps run as window passes (option 29:
    -1 from 2^18 determinants on, 1 always; register tiles only)."""
    wv = a.get_option(a.OPT_WINDOW_VECTOR)
    return opts["window_registers"] > 0 and (wv == 1 or (wv == -1 and N >= 1 << 18))


def host_cpu_share():
    """Host threads for the CPU baseline: OMP_NUM_THREADS when set (the
    GPU hosts here allot each GPU a fixed CPU share and export it there),
    else every logical CPU of the host."""
    v = os.environ.get("OMP_NUM_THREADS", "")
    if v.isdigit() and int(v) > 0:
        return int(v)
    return os.cpu_count() or 1


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path; run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    from ansatzevolve_amd import from_workload, load_library
    from ansatzevolve_amd.workloads import make_workload

    load_library()
    if args.calib_copy:
        # a plain 2 GiB device-to-device copy through the HIP runtime (its
        # blit kernel), the known size the counter passes are scaled by;
        # no torch in this process: with torch loaded, rocprofv3's counter
        # collection did not see the library's kernels and a rocBLAS call
        # failed
        nbytes = 1 << 31
        rc = calib_copy(nbytes)
        print("[bench] calibration copy: %d bytes read, %d written (status %d)" % (nbytes, nbytes, rc),
              file=sys.stderr, flush=True)
    w = make_workload(args.config)
    N, M, P = w["N"], w["M"], w["P"]
    rng = np.random.default_rng(w["seed"] + 3)
    thetas = w["theta"][None, :] + 1e-2 * rng.uniform(-math.pi, math.pi, size=(args.nvec, P))
    thetas[0] = w["theta"]
    singles = w["theta"][None, :] + 1e-2 * rng.uniform(-math.pi, math.pi, size=(7, P))

    print("[bench] set-up: %d determinants, %d operators, %d parameters" % (N, M, P), file=sys.stderr, flush=True)
    t0 = time.time()
    a = from_workload(w, storage=args.storage)
    d, M2, P2 = a.dimensions()  # constructs: device sector maps, operators, Hamiltonian rows
    setup_s = time.time() - t0
    if args.eager:
        a.set_option(a.OPT_GRAPH, 0)
    if args.no_exchange:
        a.set_option(a.OPT_SIGMA_EXCHANGE, 0)
    assert (d, M2, P2) == (N, M, P)

    # parity of the benchmarked path against the reference CPU build's golden
    # vector for theta_0 (tests/golden/reference_vectors.npz, when present)
    parity = None
    gold = os.path.join(ROOT, "tests", "golden", "reference_vectors.npz")
    if os.path.exists(gold):
        gz = np.load(gold, allow_pickle=False)
        key = "cfg%d_" % args.config
        if key + "E" in gz.files and np.array_equal(gz[key + "theta"][0], w["theta"]):
            E1, g1 = a.energy_gradient(w["theta"])
            parity = {"dE": abs(E1 - float(gz[key + "E"][0])),
                      "max_dgrad": float(np.max(np.abs(g1 - gz[key + "grad"][0]))), "against": "reference CPU build"}

    # single-vector E+grad through the reference-style entry point
    t1 = time.time()
    nsingle = 3
    for i in range(nsingle):
        a.energy_gradient(singles[i])
    single_rate = nsingle / (time.time() - t1)
    # kernel-class breakdown of the single-vector path (eager launches)
    a.set_kernel_timing(True)
    for i in range(nsingle):
        a.energy_gradient(singles[i + 3])
    single_breakdown = {}
    single_launches = 0
    for c, name in ((0, "hamiltonian"), (1, "forward_sweep"), (2, "adjoint_sweep"), (3, "same_spin_dgemm")):
        ms, cnt = a.kernel_timing(c)
        single_breakdown[name] = ms / nsingle
        single_launches += cnt
    a.set_kernel_timing(False)
    single_ms = 1000.0 / single_rate
    # one vector of 853,776 amplitudes is 6.8 MB: it stays in the on-chip
    # caches, and the sequence is a chain of short dependent launches (17
    # window passes per sweep, each of about 900 one-wave class tasks, fewer
    # than the SIMDs); its bound is latency: the heaviest task's chain of
    # dependent loads, phases and stores in every launch
    single_bound = {
        "bound": "latency",
        "ms_per_egrad": single_ms,
        "sweep_segments_per_sweep": a.get_option(a.OPT_VECTOR_SEGMENTS),
        "timed_intervals_per_egrad": single_launches / nsingle,
        "ideal_hbm_bytes_per_egrad": 40.0 * N,
        "hbm_frac_of_ideal_bytes": 40.0 * N / (single_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
    }

    tune = None
    if not args.no_autotune:
        print("[bench] autotuning the batch (sweep kind x fusion depth x layout)", file=sys.stderr, flush=True)
        rep = a.autotune(thetas, cache=args.autotune_cache or None, reps=2)
        tune = {"choice": rep["choice"], "source": rep.get("source"), "ms": rep.get("ms"),
                "max_dev": rep.get("max_dev"),
                "sigma_gemms": rep.get("sigma_gemms"),
                "candidates": [[c["choice"]["sweep"], c["choice"]["depth"], c["choice"]["lane_min"],
                                c["choice"]["groups"], c["choice"]["groups_fwd"], c["choice"]["checkpoint"],
                                c["choice"]["sigma_halves"], c["choice"]["sigma_grouped"], c["choice"]["window"],
                                round(c["ms"], 3)]
                               for c in rep.get("candidates", [])],
                "candidate_fields": ["sweep", "depth", "lane_min", "groups", "groups_fwd", "checkpoint",
                                     "sigma_halves", "sigma_grouped", "window", "ms"]}
    if args.lane_groups:
        a.set_option(a.OPT_LANE_GROUPS, args.lane_groups)
        a.set_option(a.OPT_LANE_GROUPS_FWD, args.lane_groups)
    opts = {"sweep": a.get_option(a.OPT_SWEEP), "fuse_depth": a.get_option(a.OPT_FUSE_DEPTH),
            "lane_min": a.get_option(a.OPT_LANE_MIN), "graph": a.get_option(a.OPT_GRAPH),
            "lane_groups": a.get_option(a.OPT_LANE_GROUPS), "lane_groups_fwd": a.get_option(a.OPT_LANE_GROUPS_FWD),
            "lane_segments": a.get_option(a.OPT_LANE_SEGMENTS), "checkpoint": a.get_option(a.OPT_CHECKPOINT),
            "sigma_halves": a.get_option(a.OPT_SIGMA_HALVES), "sigma_grouped": a.get_option(a.OPT_SIGMA_GROUPED),
            "gemm_solution_beta": a.get_option(a.OPT_GEMM_SOLUTION_BETA),
            "gemm_solution_alpha": a.get_option(a.OPT_GEMM_SOLUTION_ALPHA),
            "window": a.get_option(a.OPT_WINDOW)}

    # timed region: one hipGraph launch per step, inputs resident in HBM
    a.upload_batch(thetas)
    opts["checkpoint_active"] = a.get_option(a.OPT_CHECKPOINT_ACTIVE)
    opts["reachable"] = a.get_option(a.OPT_REACHABLE)
    opts["window_passes"] = a.get_option(a.OPT_WINDOW_PASSES)
    opts["window_registers"] = a.get_option(a.OPT_WINDOW_REGISTERS)
    opts["window_vector"] = a.get_option(a.OPT_WINDOW_VECTOR)
    opts["sigma_exchange"] = a.get_option(a.OPT_SIGMA_EXCHANGE)
    amp = {"forward": a.get_option(a.OPT_SWEEP_AMP_FORWARD),
           "adjoint_lambda": a.get_option(a.OPT_SWEEP_AMP_ADJOINT_LAMBDA),
           "adjoint_state": a.get_option(a.OPT_SWEEP_AMP_ADJOINT_STATE)}
    for _ in range(args.warmup):
        a.run_batch(True)
    a.download_batch(True)  # drains the stream
    print("[bench] timing %d steps of %d vectors" % (args.steps, args.nvec), file=sys.stderr, flush=True)
    t_start = time.perf_counter()
    for _ in range(args.steps):
        a.run_batch(True)
    E, G = a.download_batch(True)  # waits for the stream
    t_end = time.perf_counter()
    elapsed = t_end - t_start

    # kernel-class breakdown: the same sequence launched eagerly with HIP
    # events around each class on the engine's stream
    nbd = max(1, min(args.steps, 3))
    a.set_kernel_timing(True)
    for _ in range(nbd):
        a.run_batch(True)
    a.download_batch(True)
    ham_ms, ham_n = a.kernel_timing(0)
    fwd_ms, fwd_n = a.kernel_timing(1)
    rev_ms, rev_n = a.kernel_timing(2)
    gemm_ms, gemm_n = a.kernel_timing(3)
    a.set_kernel_timing(False)
    ham_kind = a.get_option(a.OPT_HAM_KIND)
    nseg = a.get_option(a.OPT_VECTOR_SEGMENTS)  # launches per vector-major sweep
    lanes = opts["lane_min"] > 0 and args.nvec >= opts["lane_min"]

    # batched results must equal single-vector results
    Es, gs = a.energy_gradient(thetas[-1])
    batch_consistency = max(abs(Es - E[-1]), float(np.max(np.abs(gs - G[-1]))))

    evals = args.steps * args.nvec
    value = evals / elapsed
    nnz = N * w_row_width(w)
    n_, ku, kd = w["n"], w["nup"], w["ndown"]
    rooflines = {}

    def per_launch(ms, cnt, per_interval):
        # lane-layout sweeps time each launch; other paths one interval per sweep
        return ms / max(1, cnt * per_interval)

    def hbm(kernel, bytes_launch, avg_ms, launches, per_step_ms, per_pass_bytes=None):
        ach = bytes_launch / (avg_ms * 1e-3) / 1e9
        r = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
             "traffic": None, "kernel": kernel, "avg_launch_ms": avg_ms, "launches": launches,
             "algorithmic_bytes_per_launch": bytes_launch, "ms_per_step": per_step_ms}
        if per_pass_bytes is not None:
            # the bytes each launch actually has to move (its own pass over
            # the amplitudes it touches): the streaming efficiency, apart
            # from the fusion gap that `frac` includes
            pa = per_pass_bytes / (avg_ms * 1e-3) / 1e9
            r["per_pass"] = {"bytes_per_launch": per_pass_bytes, "achieved": pa, "frac": pa / HBM_PEAK_GBS}
        return r

    def fp64(kernel, flops_launch, avg_ms, launches, per_step_ms, note):
        ach = flops_launch / (avg_ms * 1e-3) / 1e12
        return {"bound": "mfma", "achieved": ach, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": ach / FP64_PEAK_TFLOPS, "traffic": None, "kernel": kernel, "avg_launch_ms": avg_ms,
                "launches": launches, "algorithmic_flops_per_launch": flops_launch, "ms_per_step": per_step_ms,
                "note": note}

    # sweeps.  `frac` counts the ideal-fusion bytes: every vector read and
    # written once per sweep, 16 B per amplitude per vector, spread over the
    # sweep's launches.  `per_pass` counts what each launch has to move: 8 B
    # per amplitude per access over the amplitudes the engine reports per
    # step (summed over the segments): the forward sweep reads and writes the
    # reachable part of each state, the adjoint sweep reads and writes lambda
    # on every determinant some operator of the segment acts on, and reads
    # the stored state there where it can be nonzero (with stored states;
    # in place it reads and writes psi like lambda).  Vector-major: every
    # segment reads and writes each vector once (forward: psi; adjoint: psi
    # and lambda).
    two_vm = a.get_option(a.OPT_SWEEP) == 3 and a.get_option(a.OPT_SWEEP_TWO_GROUP_AVAILABLE) == 1
    k_fwd = ("k_orb2v<false>" if opts["lane_groups_fwd"] == 2 else "k_orbv<false>") + " (lane layout)" if lanes \
        else ("k_orb2s<false>" if two_vm else "k_orb<false>")
    k_rev = ("k_orb2v<true>" if opts["lane_groups"] == 2 else "k_orbv<true>") + " (lane layout)" if lanes \
        else ("k_orb2s<true>" if two_vm else "k_orb<true>")
    if lanes and opts["window_passes"] > 0:
        # window passes: register tiles (k_winr; TUPS units straight-line) or LDS tiles (k_win)
        form = {2: "k_winr<%s, TUPS units>", 1: "k_winr<%s>"}.get(opts["window_registers"], "k_win<%s>")
        k_fwd = (form % "false") + " (lane layout, %d window passes)" % opts["window_passes"]
        k_rev = (form % "true") + " (lane layout, %d window passes)" % opts["window_passes"]
    elif not lanes and window_vm(a, opts, N):
        # vector-major window passes: one register tile per lane (k_winv)
        tu = ", TUPS units" if opts["window_registers"] == 2 else ""
        k_fwd = "k_winv<false%s> (%d window passes)" % (tu, opts["window_passes"])
        k_rev = "k_winv<true%s> (%d window passes)" % (tu, opts["window_passes"])
    if lanes:
        fwd_launches, rev_launches = max(1, fwd_n // nbd), max(1, rev_n // nbd)
        fwd_pass = 16.0 * amp["forward"] * args.nvec / fwd_launches
        rev_state = 8.0 * amp["adjoint_state"] if opts.get("checkpoint_active") else 16.0 * amp["adjoint_lambda"]
        rev_pass = (16.0 * amp["adjoint_lambda"] + rev_state) * args.nvec / rev_launches
    else:
        fwd_launches = rev_launches = max(1, nseg)
        fwd_pass = 16.0 * N * args.nvec
        rev_pass = (24.0 if opts.get("checkpoint_active") else 32.0) * N * args.nvec
    ideal_sweep = 16.0 * N * args.nvec  # one read + one write of every vector per sweep
    rooflines["forward_sweep"] = hbm(k_fwd, ideal_sweep / fwd_launches,
                                     per_launch(fwd_ms, fwd_n, 1 if lanes else nseg),
                                     fwd_n if lanes else fwd_n * nseg, fwd_ms / nbd, fwd_pass)
    rooflines["adjoint_sweep"] = hbm(k_rev, ideal_sweep / rev_launches,
                                     per_launch(rev_ms, rev_n, 1 if lanes else nseg),
                                     rev_n if lanes else rev_n * nseg, rev_ms / nbd, rev_pass)
    if ham_kind == 3 and gemm_n > 0 and not lanes:
        # vector-major split sigma (k_sig1r): per target and entry one
        # coefficient and one source amplitude read from LDS, 16 B per FMA
        per_el = ku * (n_ - ku) * kd * (n_ - kd) + ku * (n_ - ku) * kd + kd * (n_ - kd) + 1
        na = math.comb(n_, ku)
        nb = math.comb(n_, kd)
        avg = ham_ms / max(ham_n, 1)
        lds_bytes = 16.0 * N * per_el * args.nvec / max(1.0, ham_n / nbd)
        ach = lds_bytes / (avg * 1e-3) / 1e9
        rooflines["hamiltonian"] = {
            "bound": "lds", "achieved": ach, "peak": LDS_PEAK_GBS, "unit": "GB/s", "frac": ach / LDS_PEAK_GBS,
            "traffic": None, "kernel": "k_sig1r (vector-major sigma, %d vectors)" % args.nvec,
            "avg_launch_ms": avg, "launches": ham_n, "algorithmic_lds_bytes_per_launch": lds_bytes,
            "ms_per_step": ham_ms / nbd, "fp64_tflops": 2.0 * N * per_el * args.nvec / nbd / (ham_ms / nbd * 1e-3) / 1e12,
            "note": "LDS peak = 256 CUs x 128 B/clock x 2.4 GHz; 16 B of LDS reads per FMA (random gathers)"}
        gflops = 2.0 * (na * na * nb + nb * nb * na) * args.nvec
        rooflines["same_spin_dgemm"] = fp64("rocBLAS dgemm_strided_batched x2 (dense same-spin rows)", gflops,
                                            gemm_ms / max(gemm_n, 1), gemm_n, gemm_ms / nbd,
                                            "dense FLOPs of both GEMMs on FP64 MFMA")
    elif ham_kind == 3 and gemm_n > 0:
        halves = a.get_option(a.OPT_SIGMA_HALVES) and a.get_option(a.OPT_SIGMA_HALVES_AVAILABLE)
        grouped = a.get_option(a.OPT_SIGMA_GROUPED) and a.get_option(a.OPT_SIGMA_GROUPED_AVAILABLE)
        sig_name = "k_sigg_sparse" if grouped else "k_sigh_sparse" if halves else "k_sigw_sparse"
        # the sparse lane sigma is bound by LDS reads: one workgroup per
        # (alpha row, 16-vector sub-chunk) stages 1 + ku (n - ku) source rows
        # (identity and the alpha singles), each 128 B per beta string, and
        # reads them back as 512-byte wave reads (64 lanes x 8 B); the
        # grouped kernel's generated programs make a fixed number of reads
        # per staged row, the one-row kernel one read per entry and target
        per_el = ku * (n_ - ku) * kd * (n_ - kd) + ku * (n_ - ku) * kd + kd * (n_ - kd) + 1
        na = math.comb(n_, ku)
        nb = math.comb(n_, kd)
        nsub = 4 * math.ceil(args.nvec / 64)
        nsrc = 1 + ku * (n_ - ku)
        if grouped:
            from ansatzevolve_amd import sigg
            plan = sigg.wave_plan(kd)
            plen = {m: len(sigg.program(kd, m)) for wv in plan for (m, _, _) in wv}
            reads = sum(plen[m] for wv in plan for (m, _, _) in wv)
            if opts.get("sigma_exchange") == 1:
                # one triangle of targets: in alpha row ar a wave skips a task
                # whose rows' beta ranks all lie below ar (mean over the rows)
                def top(h):
                    m = kd - bin(h).count("1")
                    return sigg.rh(h, m) + math.comb(6, m) - 1
                tops = {(m, tuple(rows)): max(top(h) for h in rows if h >= 0)
                        for wv in plan for (m, _, rows) in wv}
                reads = sum(plen[m] for ar in range(na) for wv in plan for (m, _, rows) in wv
                            if tops[(m, tuple(rows))] >= ar) / na
        else:
            # one read per (entry, four targets): the beta singles and the diagonal of every target
            reads = (nb + 3) // 4 * (kd * (n_ - kd) + 1)
        lds_bytes = na * nsub * nsrc * (reads * 512.0 + nb * 128.0)
        lds_peak = LDS_PEAK_GBS
        avg = ham_ms / max(ham_n, 1)
        ach = lds_bytes / (avg * 1e-3) / 1e9
        flops = 2.0 * N * per_el * args.nvec
        rooflines["hamiltonian"] = {
            "bound": "lds", "achieved": ach, "peak": lds_peak, "unit": "GB/s", "frac": ach / lds_peak,
            "traffic": None, "kernel": "%s (lane-layout sigma, %d vectors)" % (sig_name, args.nvec),
            "avg_launch_ms": avg, "launches": ham_n, "algorithmic_lds_bytes_per_launch": lds_bytes,
            "lds_wave_reads_per_staged_row": reads, "ms_per_step": ham_ms / nbd,
            "fp64_tflops": flops / (avg * 1e-3) / 1e12,
            "note": "LDS peak = 256 CUs x 128 B/clock x 2.4 GHz; FP64 work (no MFMA) for reference"}
        if opts.get("sigma_exchange") == 1:
            # exchange-symmetric batch: the alpha-alpha DGEMM alone; the
            # grouped sigma adds its result and its transpose
            rooflines["hamiltonian"]["kernel"] += ", beta-beta part from the alpha-alpha DGEMM read transposed"
            gflops = 2.0 * na * na * nb * 64 * math.ceil(args.nvec / 64)
            rooflines["same_spin_dgemm"] = fp64(
                "rocBLAS dgemm_strided_batched x1 (dense alpha-alpha rows; exchange symmetry gives the beta-beta part)",
                gflops, gemm_ms / max(gemm_n, 1), gemm_n, gemm_ms / nbd, "dense FLOPs of the GEMM on FP64 MFMA")
        else:
            gflops = (2.0 * na * na * nb * 64 * math.ceil(args.nvec / 64) +
                      2.0 * nb * nb * 64 * na * math.ceil(args.nvec / 64))
            rooflines["same_spin_dgemm"] = fp64("rocBLAS dgemm_strided_batched x2 (dense same-spin rows)", gflops,
                                                gemm_ms / max(gemm_n, 1), gemm_n, gemm_ms / nbd,
                                                "dense FLOPs of both GEMMs on FP64 MFMA")
    elif ham_kind == 3:
        r = fp64("k_mf_sigma (matrix-free sigma)", 2.0 * nnz * args.nvec / max(1.0, ham_n / nbd),
                 ham_ms / max(ham_n, 1), ham_n, ham_ms / nbd, "FP64 vector-ALU work priced against the FP64 peak")
        r["bound"] = "valu"
        rooflines["hamiltonian"] = r
    else:
        # explicit rows: 12 B per stored entry (8 B value + 4 B column) per
        # launch + 16 B per amplitude per vector (x read, y written)
        vpl = args.nvec / max(1.0, ham_n / nbd)
        rooflines["hamiltonian"] = hbm("k_ell (Hamiltonian rows x %d vectors)" % vpl, 12.0 * nnz + 16.0 * N * vpl,
                                       ham_ms / max(ham_n, 1), ham_n, ham_ms / nbd)
    # HBM traffic per launch where the counter passes recorded it (profiles/)
    for r in rooflines.values():
        t = measured_traffic(r["kernel"], args.config, args.nvec)
        if t is not None:
            r["traffic"] = t["bytes_per_launch"]
            r["traffic_detail"] = t
    dominant = max(rooflines, key=lambda k: rooflines[k]["ms_per_step"])
    roof = dict(rooflines[dominant])
    roof["class"] = dominant
    result = {
        "metric": "TUPS E+grad evals/s at %.2fM dets" % (N / 1e6) if args.config != 3 else
                  "TUPS E+grad evals/s at 0.85M dets",
        "value": value,
        "unit": "E+grad evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded integrals and angles)",
        "config": {
            "workload": "%d orbitals / %d electrons, Sz=0, %d determinants, TUPS L=%d (%d parameters, %d operators), "
                        "%d angle vectors per step, %s (%d entries)"
                        % (w["n"], w["n"], N, w["layers"], P, M, args.nvec,
                           "matrix-free sigma build" if ham_kind == 3 else "explicit Hamiltonian rows", nnz),
            "hamiltonian": "matrix-free" if ham_kind == 3 else "explicit rows",
            "batch": args.nvec,
            "determinants": N,
            "parallelism": "single GPU",
        },
        "roofline": roof,
        "roofline_by_class": rooflines,
        "breakdown_ms_per_step": {"hamiltonian": ham_ms / nbd, "same_spin_dgemm": gemm_ms / nbd,
                                  "forward_sweep": fwd_ms / nbd, "adjoint_sweep": rev_ms / nbd},
        "engine_options": opts,
        "sweep_amplitudes_per_vector_step": amp,
        "autotune": tune,
        "single_vector_egrad_per_s": single_rate,
        "single_vector_breakdown_ms": single_breakdown,
        "single_vector_bound": single_bound,
        "setup_s": setup_s,
        "parity": parity,
        "batch_vs_single_max_diff": batch_consistency,
        "cpu_baseline": None,
    }
    if not args.no_cpu_baseline and os.path.exists(os.path.join(ROOT, "oracle", "_ref", "libref_ansatz.so")):
        threads = args.cpu_threads or host_cpu_share()
        print("[bench] reference CPU baseline on %d threads" % threads, file=sys.stderr, flush=True)
        try:
            cb = cpu_baseline(args.config, threads)
            med = float(np.median(cb["times"]))
            result["cpu_baseline"] = {
                "value": 1.0 / med,
                "unit": "E+grad evals/s",
                "cores": cb["threads"],
                "kind": "reference",
                "sample": "median of %d E+grads (FusedEvolve::evolve + evolveDerivative) at theta_0 of the same "
                          "workload after one warm-up E+grad, reference CPU build, its matrix-free Hamiltonian "
                          "path, %d threads (host logical CPUs: %d): %s s; set-up %.1f s and warm-up %.1f s "
                          "not counted"
                          % (len(cb["times"]), cb["threads"], os.cpu_count() or 0,
                             ", ".join("%.2f" % t for t in cb["times"]), cb["setup"], cb["first"]),
                "times_s": cb["times"],
                "setup_s": cb["setup"],
                "warmup_s": cb["first"],
                "host_logical_cpus": os.cpu_count(),
                "dE_vs_gpu": abs(cb["E"] - a.energy(w["theta"])),
            }
        except Exception as ex:  # report, do not hide the GPU number
            result["cpu_baseline"] = {"error": str(ex)[:300]}
    if not args.no_extras:
        print("[bench] Hessian timings and the multi-instance launch", file=sys.stderr, flush=True)
        try:
            result.update(extra_measurements(a, w))
        except Exception as ex:
            result["extras_error"] = str(ex)[:300]
    a.close()
    result["warm_start"] = None
    if not args.no_warm_start:
        # the metric's second half, after the timed region: solve time to
        # 1e-6 Ha from the learned warm start vs the standard guess, on a
        # short sample of the per-instance seeded family (the full
        # measurement: 8, 10 and 12 orbitals, 160 training / 64 held out)
        print("[bench] learned warm start vs plain init (8 orbitals, 24 + 8 seeded instances)", file=sys.stderr,
              flush=True)
        try:
            from ansatzevolve_amd.warmstart import sample
            ws = sample(n=8, ntrain=24, nheld=8)
            s_all = ws["summary"]
            result["warm_start"] = {
                "sample": "8 orbitals, TUPS L=4, per-instance seeded integrals and site spacings U(1.0, 2.5); "
                          "%d training instances (L-BFGS targets, at most %d evaluations, canonical form), %d held out "
                          "(at most %d evaluations per run); width-256 network, %d epochs; %.0f s"
                          % (ws["train"], ws["generate_max_evals"], ws["held_out"], ws["eval_max_evals"], ws["epochs"],
                             ws["seconds"]),
                "median_s_to_own_final_1e-6": {"plain": s_all["median_t_own_plain"],
                                               "learned": s_all["median_t_own_mlp"]},
                "speedup_own": s_all["speedup_own"],
                "median_s_to_best_minimum_1e-6": {"plain": s_all["median_t_best_plain"],
                                                  "learned": s_all["median_t_best_mlp"]},
                "reach_best_minimum": {"plain": s_all["count"] - s_all["best_misses_plain"],
                                       "learned": s_all["count"] - s_all["best_misses_mlp"]},
                "proposals_used": s_all["proposals_used"],
                "held_out": s_all["count"],
                "full_measurement": "profiles/warmstart_time_to_1e-6_seeded_n8_n10_n12.md",
            }
        except Exception as ex:
            result["warm_start"] = {"error": str(ex)[:300]}
    print(json.dumps(_finite(result)))
    # a throughput figure only counts when the results are right
    bad = []
    if not batch_consistency <= SAME_TOL:
        bad.append("batch vs single-vector results differ by %g" % batch_consistency)
    if parity is not None and not (parity["dE"] <= E_TOL and parity["max_dgrad"] <= G_TOL):
        bad.append("results differ from the reference CPU build's golden vector: dE %g, dgrad %g"
                   % (parity["dE"], parity["max_dgrad"]))
    if bad:
        print("[bench] FAILED: " + "; ".join(bad), file=sys.stderr, flush=True)
        sys.exit(3)


def _finite(o):
    """JSON-safe copy: non-finite floats (a run that never reaches its
    target) become null."""
    if isinstance(o, float):
        return o if math.isfinite(o) else None
    if isinstance(o, dict):
        return {k: _finite(v) for k, v in o.items()}
    if isinstance(o, (list, tuple)):
        return [_finite(v) for v in o]
    return o


def w_row_width(w):
    n, ku, kd = w["n"], w["nup"], w["ndown"]
    c = math.comb
    return (1 + kd * (n - kd) + ku * (n - ku) + c(kd, 2) * c(n - kd, 2) + c(ku, 2) * c(n - ku, 2)
            + ku * (n - ku) * kd * (n - kd))


if __name__ == "__main__":
    main()
