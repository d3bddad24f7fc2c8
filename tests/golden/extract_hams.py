"""Extract the reference's H-chain data sets (Hams/H*/L*) into hams.npz.

Data only: integrals (_oneEInts.bin / _twoEInts.bin, raw FP64), the TUPS
operator list (_Operators.dat), the order file with scales (_Order.dat),
the start state (_Initial.dat), the nuclear energy (_Nuclear_Energy.dat)
and every parameter set of _Parameters.dat with its "Energy of minimum"
(electronic + nuclear).  Formats follow the reference's loaders
(src/TUPSLoadingUtils.cpp:10-296, Manual.md:129-291); sets whose energy is
the placeholder 1 are dropped.

Run in the build container:  python tests/golden/extract_hams.py
"""
import os
import re
import sys

import numpy as np

REF = os.environ.get("ANSATZ_REFERENCE", "/root/reference") + "/Hams"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hams.npz")


def read_params(path):
    """Sets start with a count line followed by an "Energy of minimum" line;
    the values run to the next set (some placeholder sets hold fewer values
    than their count, so the count is not trusted)."""
    lines = [l.strip() for l in open(path) if l.strip()]
    starts = [i for i in range(len(lines) - 1)
              if re.fullmatch(r"\d+", lines[i]) and lines[i + 1].startswith("Energy")]
    sets = []
    for s, e in zip(starts, starts[1:] + [len(lines)]):
        energy = float(re.search(r"=\s*([-0-9.eE+]+)", lines[s + 1]).group(1))
        vals = [float(x) for x in lines[s + 2:e]]
        if len(vals) == int(lines[s]):
            sets.append((energy, vals))
    return sets


def read_order(path):
    orders, scales = [], []
    for line in open(path):
        line = line.strip()
        if not line:
            continue
        if "," in line:
            o, s = line.split(",")
            orders.append(int(o))
            scales.append(float(s))
        else:
            orders.append(int(line))
            scales.append(1.0)
    return np.array(orders) - 1, np.array(scales)


def read_initial(path):
    idx, co = [], []
    for line in open(path):
        line = line.strip()
        if not line:
            continue
        bits, re_, im_ = line.split(",")
        idx.append(int(bits, 2))  # rightmost character is bit 0
        co.append(float(re_))
    return idx, co, len(bits)


def main():
    data = {}
    for mol in sorted(os.listdir(REF)):
        if not re.match(r"H\d+$", mol):
            continue
        for lay in sorted(os.listdir(os.path.join(REF, mol))):
            d = os.path.join(REF, mol, lay)
            pre = os.path.join(d, mol)
            if not os.path.exists(pre + "_Parameters.dat"):
                continue
            sets = [(e, v) for e, v in read_params(pre + "_Parameters.dat") if e != 1.0]
            if not sets:
                continue
            key = "%s_%s_" % (mol, lay)
            h = np.fromfile(pre + "_oneEInts.bin", dtype=np.float64)
            g = np.fromfile(pre + "_twoEInts.bin", dtype=np.float64)
            ops = np.loadtxt(pre + "_Operators.dat", dtype=np.int32).reshape(-1, 4)
            order, scale = read_order(pre + "_Order.dat")
            idx, co, nq = read_initial(pre + "_Initial.dat")
            nuc = float(open(pre + "_Nuclear_Energy.dat").read().split()[0])
            npar = max(len(sets[0][1]), int(order.max()) + 1)
            data[key + "h"] = h
            data[key + "g"] = g
            data[key + "ops1"] = ops
            data[key + "order0"] = order
            data[key + "scale"] = scale
            data[key + "init_idx0"] = np.array(idx)
            data[key + "init_coeff"] = np.array(co)
            data[key + "nq"] = np.array(nq)
            data[key + "nuclear"] = np.array(nuc)
            data[key + "energies"] = np.array([e for e, _ in sets])
            data[key + "params"] = np.array([v + [0.0] * (npar - len(v)) for _, v in sets])
            print(key, "nq", nq, "ops", len(ops), "params", npar, "sets", len(sets),
                  [e for e, _ in sets], file=sys.stderr)
    np.savez_compressed(OUT, **data)
    print("wrote", OUT, file=sys.stderr)


if __name__ == "__main__":
    main()
