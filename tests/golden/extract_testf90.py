"""Extract the known-answer data held in the reference's Fortran smoke test.

Reads ``src/test.F90`` of the reference (CC0-licensed, lines 1-2) as text and
writes its numeric inputs and printed expectations to ``testf90_kat.json``:
the 2239 one-based Hamiltonian triplets (test.F90:25-27), the 15 operators and
order (:29-30), the initial basis index 171 (:46-48), the two angle sets
(:32, :96) and the expected energies / gradient norms (:67-70, :74-78,
:98-101, :105-109).  Only numbers are kept; no source text is copied.

Run in the build container (the reference is not present on the GPU host):
    python tests/golden/extract_testf90.py
"""
import json
import os
import re
import sys

SRC = os.environ.get("ANSATZ_REFERENCE", "/root/reference") + "/src/test.F90"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "testf90_kat.json")


def _array_after(text, marker, start=0):
    i = text.index(marker, start)
    j = text.index("[", i)
    k = text.index("]", j)
    body = text[j + 1:k].replace("&", " ").replace("_dp", "")
    return [t for t in re.split(r"[,\s]+", body) if t], k


def main():
    text = open(SRC).read()
    hi, _ = _array_after(text, "HamiIndexes =")
    hj, _ = _array_after(text, "HamjIndexes =")
    hc, _ = _array_after(text, "HamCoeffs =")
    ops, _ = _array_after(text, "operators =")
    order, _ = _array_after(text, "order =")
    ang0, _ = _array_after(text, "angles =")
    ang1, _ = _array_after(text, "angles = [0.18")
    expected = [float(x.replace("_dp", "")) for x in
                re.findall(r'"Expected:",\s*([-0-9.Ee+_dp]+)', text)]
    gexp = [float(x.replace("_dp", "").replace("E", "e")) for x in
            re.findall(r'"gradient Norm Expected:",\s*([-0-9.Ee+_dp]+)', text)]
    ops = [int(x) for x in ops]
    nops = len(ops) // 4
    kat = {
        "num_qubits": 8,
        "initial_index_1based": 171,
        "initial_coeff": 1.0,
        "ham_i_1based": [int(x) for x in hi],
        "ham_j_1based": [int(x) for x in hj],
        "ham_coeffs": [float(x) for x in hc],
        # one row per operator, one-based spin-orbital indices, 0 = absent
        "operators_1based": [ops[4 * i:4 * i + 4] for i in range(nops)],
        "order_1based": [int(x) for x in order],
        "angles_hf": [float(x) for x in ang0],
        "angles_min": [float(x) for x in ang1],
        # real-mode expectations only (the complex repeat follows them)
        "energy_hf": expected[0],
        "energy_min": expected[1],
        "grad_norm_hf": gexp[0],
        "grad_norm_min": gexp[1],
    }
    assert len(kat["ham_i_1based"]) == 2239 == len(kat["ham_coeffs"])
    assert nops == 15 and len(kat["order_1based"]) == 15
    assert len(kat["angles_min"]) == 9
    with open(OUT, "w") as f:
        json.dump(kat, f, indent=1)
    print("wrote", OUT, file=sys.stderr)


if __name__ == "__main__":
    main()
