"""The run-file front end without a GPU: the parameter-set reader against
the values extracted independently into hams.npz, the set-up reader's
error codes (construction, which needs the device, is not reached), and
the command-line program's argument handling."""
import os
import subprocess

import numpy as np
import pytest

import ansatzevolve_amd as ae
from ansatzevolve_amd import AnsatzError, AnsatzSynth, read_parameter_sets

HERE = os.path.dirname(os.path.abspath(__file__))
RUN = os.path.join(HERE, "golden", "runfiles")
H4 = os.path.join(RUN, "H4", "L1", "H4")
CLI = os.path.join(os.path.dirname(ae.__file__), "ansatzsynth")


def test_parameter_sets_match_extracted_values():
    d = np.load(os.path.join(HERE, "golden", "hams.npz"), allow_pickle=False)
    for key, prefix in (("H4_L1_", H4), ("H6_L1_", os.path.join(RUN, "H6", "L1", "H6"))):
        E, P = read_parameter_sets(prefix + "_Parameters.dat")
        # the files also hold a placeholder set (energy 1), which hams.npz drops
        real = E != 1.0
        assert np.array_equal(E[real], d[key + "energies"])
        assert np.array_equal(P[real], d[key + "params"])
    E, P = read_parameter_sets(os.path.join(RUN, "TestFermionic", "TestFermionic_Parameters.dat"))
    assert np.array_equal(E, [-2.571528425984818, -2.266587154543422])
    assert np.array_equal(P, [[0.3927, 0.0], [0.388888355016995, -1.50714697857731]])


def test_parameter_sets_errors(tmp_path):
    with pytest.raises(AnsatzError) as e:
        read_parameter_sets(str(tmp_path / "missing_Parameters.dat"))
    assert e.value.code == 3
    # a set whose value count differs from its count line is skipped
    f = tmp_path / "x_Parameters.dat"
    f.write_text("3\nEnergy of minimum 1= -1.5 first\n0.1\n0.2\n2\nEnergy of minimum 2= -2.5 x\n1\n2\n")
    E, P = read_parameter_sets(str(f))
    assert np.array_equal(E, [-2.5]) and np.array_equal(P, [[1.0, 2.0]])


def test_load_run_files_codes(tmp_path):
    with AnsatzSynth() as a:
        a.load_run_files(H4)  # stores only; the Hamiltonian is read at construction
    with AnsatzSynth() as a:
        with pytest.raises(AnsatzError) as e:
            a.load_run_files(str(tmp_path / "none"))
        assert e.value.code == 3
        with pytest.raises(AnsatzError) as e:
            a.load_run_files("")
        assert e.value.code == 2
        with pytest.raises(AnsatzError) as e:
            a.set_run_path("")
        assert e.value.code == 2
    # operator and order counts differ
    for name in ("Initial.dat", "Operators.dat", "Nuclear_Energy.dat"):
        (tmp_path / ("Y_" + name)).write_bytes(open(H4 + "_" + name, "rb").read())
    (tmp_path / "Y_Order.dat").write_text("1\n2\n")
    with AnsatzSynth() as a:
        with pytest.raises(AnsatzError) as e:
            a.load_run_files(str(tmp_path / "Y"))
        assert e.value.code == 3


def test_cli_arguments(tmp_path):
    r = subprocess.run([CLI, "help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "filepath" in r.stdout
    r = subprocess.run([CLI, "nonsense"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "unknown argument" in r.stderr
    r = subprocess.run([CLI, "optimise"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "No file given" in r.stderr
    r = subprocess.run([CLI, "filepath", str(tmp_path / "none")], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "loadRunFiles failed (3)" in r.stderr
