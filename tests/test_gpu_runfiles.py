"""The run-file front end on the device: energies of the recorded parameter
sets of the reference's input files (Hams/H4/L1, Hams/H6/L1: integrals;
Hams/TestFermionic: a text Hamiltonian) read through loadRunFiles, against
the recorded energies, the oracle, and the same set-up made through the
array entry points; the command-line program on the same files."""
import os
import re
import subprocess

import numpy as np
import pytest

import ansatzevolve_amd as ae
from ansatzevolve_amd import AnsatzSynth, read_parameter_sets
from oracle.oracle import OracleEngine

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
RUN = os.path.join(HERE, "golden", "runfiles")
PREFIX = {"H4": os.path.join(RUN, "H4", "L1", "H4"), "H6": os.path.join(RUN, "H6", "L1", "H6"),
          "TestFermionic": os.path.join(RUN, "TestFermionic", "TestFermionic")}
CLI = os.path.join(os.path.dirname(ae.__file__), "ansatzsynth")
E_TOL = 1e-10


@pytest.mark.parametrize("name", ["H4", "H6"])
def test_integral_run_files_recorded_minima(gpu, name):
    """Recorded minima (electronic energies) through the files; the nuclear
    energy is added with with_nuclear; the run-path Hamiltonian equals the
    integrals passed as arrays."""
    pre = PREFIX[name]
    E_rec, P = read_parameter_sets(pre + "_Parameters.dat")
    nuc = float(open(pre + "_Nuclear_Energy.dat").read().split()[0])
    d = np.load(os.path.join(HERE, "golden", "hams.npz"), allow_pickle=False)
    key = name + "_L1_"
    with AnsatzSynth() as a, AnsatzSynth() as b, AnsatzSynth() as c:
        a.load_run_files(pre)
        b.load_run_files(pre, with_nuclear=True)
        nq = int(d[key + "nq"])
        c.set_initial_state(nq, d[key + "init_idx0"] + 1, d[key + "init_coeff"])
        c.set_integrals(d[key + "h"].reshape(nq // 2, nq // 2), d[key + "g"])
        c.set_excitation(d[key + "ops1"], d[key + "order0"] + 1, d[key + "scale"])
        npar = a.dimensions()[2]
        for E_r, p in zip(E_rec, P):
            Ea = a.energy(p[:npar])
            assert abs(b.energy(p[:npar]) - (Ea + nuc)) <= 1e-12
            assert abs(Ea - c.energy(p[:npar])) <= 1e-12
            if E_r != 1.0:  # placeholder sets carry energy 1
                assert abs(Ea - E_r) <= E_TOL, (name, Ea, E_r)
                Eg, g = a.energy_gradient(p[:npar])
                assert abs(Eg - E_r) <= E_TOL and np.linalg.norm(g) < 1e-5


def test_text_hamiltonian_run_files(gpu):
    """TestFermionic: a text Hamiltonian (_Ham_Coeff.dat, _Ham_Index.dat) of
    8 qubits; energies of its two parameter sets against the oracle given
    the triplets read here with numpy, and against the recorded energies.
    The first set's recorded energy (-2.5715) lies below the lowest
    eigenvalue of the two-electron block (-2.3772), so no state of the
    start's sector has it (no test of the reference loads this set); the
    second set's (-2.266587154543422) is reproduced."""
    pre = PREFIX["TestFermionic"]
    E_rec, P = read_parameter_sets(pre + "_Parameters.dat")
    v = np.loadtxt(pre + "_Ham_Coeff.dat", delimiter=",", ndmin=2)[:, 0]
    ij = np.loadtxt(pre + "_Ham_Index.dat", dtype=np.int64, ndmin=2)
    ops = np.loadtxt(pre + "_Operators.dat", dtype=np.int64, ndmin=2) - 1
    order = np.loadtxt(pre + "_Order.dat", dtype=np.int64, ndmin=1) - 1
    bits = open(pre + "_Initial.dat").read().split(",")[0].strip()
    start = int(bits, 2)
    o = OracleEngine(len(bits), [start], [1.0], ops, order)
    o.set_triplets(ij[:, 0] - 1, ij[:, 1] - 1, v)
    H = np.zeros((1 << len(bits),) * 2)
    np.add.at(H, (ij[:, 0] - 1, ij[:, 1] - 1), v)
    sector = [i for i in range(len(H)) if bin(i).count("1") == bin(start).count("1")]
    lowest = np.linalg.eigvalsh(((H + H.T) / 2)[np.ix_(sector, sector)])[0]
    compared = 0
    with AnsatzSynth() as a:
        a.load_run_files(pre)
        npar = a.dimensions()[2]
        for E_r, p in zip(E_rec, P):
            E = a.energy(p[:npar])
            assert abs(E - o.energy(p[:npar])) <= E_TOL
            assert E >= lowest - 1e-12
            if E_r >= lowest - 1e-9:
                assert abs(E - E_r) <= E_TOL, (E, E_r)
                compared += 1
    assert compared == 1 and E_rec[0] < lowest - 0.1


def _cli(*args):
    r = subprocess.run([CLI] + list(args), capture_output=True, text=True, timeout=240)
    assert r.returncode == 0, r.stderr
    return r.stdout


def test_cli_paths_and_optimise(gpu):
    """The command-line program: the energy of every path (path 0 = zero
    angles) equals the library's and the recorded one; optimise from path 1
    ends at or below its start with a small gradient."""
    pre = PREFIX["H4"]
    out = _cli("filepath", pre)
    rows = re.findall(r"path (\d+): energy (\S+) \(with nuclear (\S+)\)(?:, recorded (\S+))?", out)
    E_rec, P = read_parameter_sets(pre + "_Parameters.dat")
    assert len(rows) == 1 + len(E_rec)
    with AnsatzSynth() as a:
        a.load_run_files(pre)
        npar = a.dimensions()[2]
        for k, E, _, rec in rows:
            p = np.zeros(npar) if int(k) == 0 else P[int(k) - 1][:npar]
            assert abs(float(E) - a.energy(p)) <= 1e-12
            if rec and float(rec) != 1.0:
                assert abs(float(E) - float(rec)) <= E_TOL
        start = a.energy(P[0][:npar])
    out = _cli("filepath", pre, "optimise")
    E_opt = float(re.search(r"optimised energy (\S+)", out).group(1))
    th = np.array([float(x) for x in re.search(r"angles (.*)", out).group(1).split()])
    assert E_opt <= start + 1e-12
    with AnsatzSynth() as a:
        a.load_run_files(pre)
        assert abs(a.energy(th) - E_opt) <= 1e-12
        assert np.linalg.norm(a.gradient(th)) < 1e-8
    out = _cli("filepath", pre, "benchmark", "5")
    assert "Benchmark Derivative Time per derivative (us)" in out
    out = _cli("filepath", pre, "onlyevolve")
    assert abs(float(re.search(r"norm (\S+)", out).group(1)) - 1.0) <= 1e-12


def test_cli_iterative_subspace_properties_extracted(gpu, tmp_path):
    """The command-line modes iterativeoptimise, subspacediag,
    writeproperties and loadExtractedMin against the library's entry
    points on the same run files."""
    import shutil
    pre = PREFIX["H6"]
    E_rec, P = read_parameter_sets(pre + "_Parameters.dat")
    with AnsatzSynth() as a:
        a.load_run_files(pre)
        npar = a.dimensions()[2]
        paths = [np.zeros(npar)] + [p[:npar] for p in P]
        out = _cli("filepath", pre, "iterativeoptimise")
        E_it = float(re.search(r"iterative energy (\S+)", out).group(1))
        th = np.array([float(x) for x in re.search(r"angles (.*)", out).group(1).split()])
        assert abs(a.energy(th) - E_it) <= 1e-12
        assert np.linalg.norm(a.gradient(th)) < 1e-4
        out = _cli("filepath", pre, "subspacediag")
        m = re.search(r"null space rank (\d+), eigenvalues (.*)", out)
        ev_cli = np.array([float(x) for x in m.group(2).split()])
        ev, _, _, _ = a.subspace_diagonalise(np.array(paths[1:10]))
        assert np.max(np.abs(ev_cli - ev)) <= 1e-12
        out = _cli("filepath", pre, "writeproperties", "NOverlap", "2")
        rows, _, _, low = a.properties(np.array(paths), noverlaps=2)
        line = [ln for ln in out.splitlines() if ln.startswith("Elec. Energy")][0]
        assert np.max(np.abs(np.array([float(x) for x in line.split()[2:]]) - [r["energy"] for r in rows])) <= 1e-12
        line = [ln for ln in out.splitlines() if ln.startswith("NegativeHessianEValues")][0]
        assert [int(x) for x in line.split()[1:]] == [r["hess_neg"] for r in rows]
        low_cli = [float(x) for x in re.search(r"lowest eigenvalues of H: (.*)", out).group(1).split()]
        assert np.max(np.abs(np.array(low_cli) - low)) <= 1e-12
        # parameter sets from an _extractedmin file (one value per parameter)
        dst = str(tmp_path / "X")
        for suffix in ("_Initial.dat", "_Operators.dat", "_Order.dat", "_Nuclear_Energy.dat", "_oneEInts.bin",
                       "_twoEInts.bin"):
            shutil.copy(pre + suffix, dst + suffix)
        sets = np.stack([paths[1], 0.5 * paths[1]])
        with open(dst + "_extractedmin", "w") as f:
            for s in sets:
                f.write(" ".join("%.17g" % x for x in s) + "\n")
        out = _cli("filepath", dst, "loadExtractedMin")
        got = dict((int(k), float(E)) for k, E in re.findall(r"path (\d+): energy (\S+)", out))
        assert len(got) == 3
        for k in (1, 2):
            assert abs(got[k] - a.energy(sets[k - 1])) <= 1e-12
