"""Grouped-target programs for the lane-layout sparse sigma at 12 orbitals.

The sparse sigma applies, for one staged source row (alpha string fixed), a
beta operator of the form

    y[b] += sum_{r in b, q not in b} G[r][q] (-1)^bp(b, r, q) x[b - r + q]
          + (sum_{e in b} G[e][e]) x[b]

(``k_beta_table`` entries and the occupancy-sum diagonal of ``k_sigw_sparse``;
the reference's alpha-beta double and single terms,
src/hamiltonianmatrix.cpp:541-640), where bp(b, r, q) is the parity of the
occupied orbitals of b strictly between r and q.

Here the 12 orbitals are split into L = 0..5 and H = 6..11, and a beta string
b is the pair (bL, h).  All targets that share h form one group: C(6, m)
strings, m = K - |h|, one group per 16-lane row of a wave (64 rows = 2^6
values of h).  In ascending (colex) order rank(b) = rankL_m(bL) + RH(h, m),
so every source of a group's targets sits at a row base (one of at most 16
per row) plus a static offset rankL * 128 bytes, and every coefficient is a
per-row value times a static sign:

* inside sources (bL', h): within-L singles, sign from bL only, value
  G[r][q] (r, q in L), plus the diagonal of the target itself;
* within-H sources (bL, h - r + q): value G[r][q] (-1)^bp(h, r, q);
* cross sources (bL - r, h + q) and (bL + q, h - r): value G[r][q] times the
  parity of h below the H orbital, static sign the parity of bL above the L
  orbital.

One LDS read of a source then feeds every target of the group that it
reaches (1 to 10 FMAs), instead of one FMA per read.  This module builds
the static per-row tables and the unrolled per-m programs and writes them
as ``csrc/sigg_prog.inc``; ``apply_model`` interprets the same data on the
CPU for the unit tests.
"""
from __future__ import annotations

import itertools
import os
from math import comb

NL = 6  # L orbitals 0..5
NH = 6  # H orbitals 6..11
N = NL + NH
SLOT = 128  # bytes per staged beta (16 vectors of doubles)
NCHUNK = 6  # coefficient chunks of 16 per row
NBASE = 16


def subsets(nbits, k):
    """k-subsets of nbits as masks, ascending (colex)."""
    return [s for s in range(1 << nbits) if bin(s).count("1") == k]


def bits(mask):
    return [i for i in range(32) if (mask >> i) & 1]


def rank_l(mask, m):
    return subsets(NL, m).index(mask)


def rh(h, mL):
    """Rank contribution of the H part h (6-bit) when mL electrons sit in L."""
    return sum(comb(NL + p, mL + j + 1) for j, p in enumerate(bits(h)))


def full_rank(b, K):
    return subsets(N, K).index(b)


def between_parity(s, r, q):
    lo, hi = min(r, q), max(r, q)
    return bin(s & (((1 << hi) - 1) & ~((1 << (lo + 1)) - 1))).count("1") & 1


def row_layout(K, h):
    """Per-row base list and coefficient-slot layout for row h (beta K electrons)."""
    a = bin(h).count("1")
    m = K - a
    occ = bits(h)  # occupied H positions (relative)
    emp = [p for p in range(NH) if not (h >> p) & 1]
    wh = [(r, q) for r in occ for q in emp]  # within-H swaps, target has r, source q
    bases = [rh(h, m) * SLOT]
    for r, q in wh:
        bases.append(rh(h ^ (1 << r) ^ (1 << q), m) * SLOT)
    for q in emp:  # cross-down sources (bL - r, h + q)
        bases.append(rh(h | (1 << q), m - 1) * SLOT if m >= 1 else 0)
    for r in occ:  # cross-up sources (bL + q, h - r)
        bases.append(rh(h & ~(1 << r), m + 1) * SLOT if m + 1 <= NL else 0)
    assert len(bases) <= NBASE
    bases += [0] * (NBASE - len(bases))
    return dict(a=a, m=m, occ=occ, emp=emp, wh=wh, bases=bases)


def slot_layout(K, m):
    """Coefficient slot indices of the program for m L electrons (a = K - m)."""
    a = K - m
    wl = [(r, q) for r in range(NL) for q in range(NL) if r != q]
    c = 0
    idx = {}
    for rq in wl:
        idx[("wl",) + rq] = c
        c += 1
    for i in range(a * (NH - a)):
        idx[("wh", i)] = c
        c += 1
    for r in range(NL):
        for j in range(NH - a):
            idx[("xd", r, j)] = c
            c += 1
    for j in range(a):
        for q in range(NL):
            idx[("xu", j, q)] = c
            c += 1
    for t in range(comb(NL, m)):
        idx[("dg", t)] = c
        c += 1
    assert c <= 16 * NCHUNK
    return idx, c


def valid_m(K, m):
    return 0 <= m <= NL and 0 <= K - m <= NH


def row_entries(K, h, n=N, nn=N * N):
    """Coefficient-slot entries of row h: a gs byte offset (sign as +nn), or
    0x80000000 | rank << 12 | string for a diagonal slot; unused slots point
    at the zero slot 2 nn."""
    lay = row_layout(K, h)
    a, m = lay["a"], lay["m"]
    idx, cnt = slot_layout(K, m)
    zero = 2 * nn * 8
    ent = [zero] * (16 * NCHUNK)

    def gsoff(r, q, neg):
        return ((r * n + q) + (nn if neg else 0)) * 8

    for key, c in idx.items():
        kind = key[0]
        if kind == "wl":
            _, r, q = key
            ent[c] = gsoff(r, q, False)
        elif kind == "wh":
            r, q = lay["wh"][key[1]]
            neg = between_parity(h, r, q)
            ent[c] = gsoff(NL + r, NL + q, neg)
        elif kind == "xd":
            _, r, j = key
            q = lay["emp"][j]
            neg = bin(h & ((1 << q) - 1)).count("1") & 1
            ent[c] = gsoff(r, NL + q, neg)
        elif kind == "xu":
            _, j, q = key
            r = lay["occ"][j]
            neg = bin(h & ((1 << r) - 1)).count("1") & 1
            ent[c] = gsoff(NL + r, q, neg)
        else:
            t = key[1]
            bl = subsets(NL, m)[t]
            s = bl | (h << NL)
            rk = rh(h, m) + t
            ent[c] = 0x80000000 | (rk << 12) | s
    return ent


def program(K, m):
    """Reads and FMAs of the program for rows with m L electrons.

    Returns a list of reads (base k, byte offset, [(target t, slot c, neg)])."""
    a = K - m
    idx, _ = slot_layout(K, m)
    tl = subsets(NL, m)
    reads_in, reads_wh, reads_xd, reads_xu = [], [], [], []
    # inside sources: the target itself (diagonal) and within-L singles
    for lam in tl:
        f = []
        for t, bl in enumerate(tl):
            if bl == lam:
                f.append((t, idx[("dg", t)], False))
            else:
                d = bl ^ lam
                if bin(d).count("1") == 2:
                    r = bits(bl & d)[0]  # target has r
                    q = bits(lam & d)[0]  # source has q
                    f.append((t, idx[("wl", r, q)], bool(between_parity(bl, r, q))))
        reads_in.append((0, rank_l(lam, m) * SLOT, f))
    # within-H sources: one per (target, swap)
    for t in range(len(tl)):
        for i in range(a * (NH - a)):
            reads_wh.append((1 + i, t * SLOT, [(t, idx[("wh", i)], False)]))
    # cross-down: source L part lam (m - 1), H part h + q_j
    if m >= 1:
        for j in range(NH - a):
            for lam in subsets(NL, m - 1):
                f = []
                for r in range(NL):
                    if not (lam >> r) & 1:
                        bl = lam | (1 << r)
                        t = tl.index(bl)
                        above = bin(bl >> (r + 1)).count("1") & 1
                        f.append((t, idx[("xd", r, j)], bool(above)))
                reads_xd.append((1 + a * (NH - a) + j, rank_l(lam, m - 1) * SLOT, f))
    # cross-up: source L part lam (m + 1), H part h - r_j
    if m + 1 <= NL:
        for j in range(a):
            for lam in subsets(NL, m + 1):
                f = []
                for q in bits(lam):
                    bl = lam & ~(1 << q)
                    t = tl.index(bl)
                    above = bin(bl >> (q + 1)).count("1") & 1
                    f.append((t, idx[("xu", j, q)], bool(above)))
                reads_xu.append((1 + a * (NH - a) + (NH - a) + j, rank_l(lam, m + 1) * SLOT, f))
    # interleave: high-fanout reads between the single-FMA within-H reads
    out = []
    groups = [reads_in, reads_xd, reads_xu, reads_wh]
    heavy = list(itertools.chain(reads_in, reads_xd, reads_xu))
    light = reads_wh
    hi = li = 0
    while hi < len(heavy) or li < len(light):
        if hi < len(heavy):
            out.append(heavy[hi])
            hi += 1
        take = 2 if hi < len(heavy) else len(light)
        for _ in range(take):
            if li < len(light):
                out.append(light[li])
                li += 1
    del groups
    return out


def wave_plan(K):
    """Per wave of 16: list of tasks (m, acc offset, [4 rows h or -1])."""
    rows_by_m = {}
    for h in range(1 << NH):
        m = K - bin(h).count("1")
        if valid_m(K, m):
            rows_by_m.setdefault(m, []).append(h)
    tasks = []
    for m in sorted(rows_by_m, key=lambda mm: -comb(NL, mm)):
        rows = rows_by_m[m]
        for i in range(0, len(rows), 4):
            chunk = rows[i:i + 4]
            tasks.append((m, chunk + [-1] * (4 - len(chunk))))
    # largest programs one wave each, the small ones paired
    waves = []
    big = [t for t in tasks if comb(NL, t[0]) >= 15]
    small = [t for t in tasks if comb(NL, t[0]) < 15]
    for t in big:
        waves.append([(t[0], 0, t[1])])
    while small:
        t1 = small.pop(0)
        if small:
            t2 = small.pop(0)
            waves.append([(t1[0], 0, t1[1]), (t2[0], 10, t2[1])])
        else:
            waves.append([(t1[0], 0, t1[1])])
    return waves


def apply_model(G, x, K):
    """CPU interpretation of the tables and programs for one vector: x, y
    indexed by beta rank (ascending strings with K of 12 bits)."""
    import numpy as np

    nn = N * N
    gs = np.zeros(2 * nn + 8)
    gs[:nn] = G.reshape(-1)
    gs[nn:2 * nn] = -G.reshape(-1)
    dt = np.zeros(128)
    for t in range(64):
        dt[t] = sum(G[b][b] for b in range(6) if (t >> b) & 1)
        dt[64 + t] = sum(G[6 + b][6 + b] for b in range(6) if (t >> b) & 1)
    y = np.zeros_like(x)
    xb = x.reshape(-1)
    progs = {}
    for wave in wave_plan(K):
        for (m, _, rows) in wave:
            if m not in progs:
                progs[m] = program(K, m)
            for h in rows:
                if h < 0:
                    continue
                lay = row_layout(K, h)
                ent = row_entries(K, h)
                cv = np.zeros(16 * NCHUNK)
                for c, e in enumerate(ent):
                    if e & 0x80000000:
                        s = e & 0xFFF
                        cv[c] = dt[s & 63] + dt[64 + ((s >> 6) & 63)]
                    else:
                        cv[c] = gs[e // 8]
                acc = np.zeros(comb(NL, m))
                for (k, off, fl) in progs[m]:
                    addr = lay["bases"][k] + off
                    assert addr % SLOT == 0
                    v = xb[addr // SLOT]
                    for (t, c, neg) in fl:
                        acc[t] += (-cv[c] if neg else cv[c]) * v
                for t in range(len(acc)):
                    y[rh(h, m) + t] += acc[t]
    return y


def _pipelined(prog, batch):
    """Reads in batches of `batch`, software-pipelined: batch j + 1 is issued
    before the FMAs of batch j (two register sets), with scheduling barriers
    so that the compiler keeps the order."""
    out = []
    nb = (len(prog) + batch - 1) // batch
    out.append("    double " + ", ".join(f"x{s}_{i}" for s in range(2) for i in range(batch)) + ";")

    def loads(j):
        r = []
        for i, (k, off, _) in enumerate(prog[j * batch:(j + 1) * batch]):
            r.append(f"    x{j % 2}_{i} = lds_at(ad[{k}] + {off}u);")
        return r

    out += loads(0)
    for j in range(nb):
        if j + 1 < nb:
            out += loads(j + 1)
        out.append("    __builtin_amdgcn_sched_barrier(0);")
        for i, (_, _, fl) in enumerate(prog[j * batch:(j + 1) * batch]):
            for (t, c, neg) in fl:
                fn = "fma_row_bcast_neg" if neg else "fma_row_bcast"
                out.append(f"    {fn}<{c % 16}>(acc[AO + {t}], cv[{c // 16}], x{j % 2}_{i});")
        out.append("    __builtin_amdgcn_sched_barrier(0);")
    return out


def emit(K_list=(6,), path=None, batch=4):
    """Write csrc/sigg_prog.inc: per K, the row tables, the wave plan and one
    unrolled device function per m."""
    if path is None:
        path = os.path.join(os.path.dirname(__file__), "csrc", "sigg_prog.inc")
    L = []
    L.append("// Generated by ansatzevolve_amd/sigg.py (grouped-target sparse sigma, 12 orbitals).")
    L.append("// Do not edit: rerun the generator (build.py does).")
    L.append("#pragma once")
    for K in K_list:
        plan = wave_plan(K)
        assert len(plan) <= 16, len(plan)
        base = []
        ents = []
        for h in range(1 << NH):
            m = K - bin(h).count("1")
            if valid_m(K, m):
                base += row_layout(K, h)["bases"]
                ents += row_entries(K, h)
            else:
                base += [0] * NBASE
                ents += [2 * N * N * 8] * (16 * NCHUNK)
        L.append(f"// K = {K} beta electrons")
        L.append(f"__device__ const uint32_t kSiggBase{K}[{len(base)}] = {{")
        for i in range(0, len(base), 16):
            L.append("    " + ", ".join(f"{v}u" for v in base[i:i + 16]) + ",")
        L.append("};")
        L.append(f"__device__ const uint32_t kSiggEnt{K}[{len(ents)}] = {{")
        for i in range(0, len(ents), 16):
            L.append("    " + ", ".join(f"0x{v:08x}u" for v in ents[i:i + 16]) + ",")
        L.append("};")
        # wave plan: per wave two tasks of (m, acc offset, 4 rows); m = -1 none
        wp = []
        for w in range(16):
            tk = plan[w] if w < len(plan) else []
            for ti in range(2):
                if ti < len(tk):
                    m, off, rows = tk[ti]
                    wp += [m, off] + rows
                else:
                    wp += [-1, 0, -1, -1, -1, -1]
        L.append(f"__device__ const int kSiggPlan{K}[{len(wp)}] = {{")
        for i in range(0, len(wp), 12):
            L.append("    " + ", ".join(str(v) for v in wp[i:i + 12]) + ",")
        L.append("};")
        for m in range(NL + 1):
            if not valid_m(K, m):
                continue
            prog = program(K, m)
            nt = comb(NL, m)
            L.append(f"// m = {m}: {nt} targets, {len(prog)} reads, "
                     f"{sum(len(f) for _, _, f in prog)} FMAs")
            L.append(f"template <int AO> __device__ __forceinline__ void sigg_prog_{K}_{m}("
                     "double (&acc)[kSiggAcc], const uint32_t (&ad)[16], const double (&cv)[6])")
            L.append("{")
            L += _pipelined(prog, batch)
            L.append("}")
    text = "\n".join(L) + "\n"
    old = None
    if os.path.exists(path):
        with open(path) as f:
            old = f.read()
    if old != text:
        with open(path, "w") as f:
            f.write(text)
    return path


if __name__ == "__main__":
    print(emit())
