"""The reference's Python module mirror (ansatzevolve_amd.PyAnsatzEvolve)
without a GPU: the file functions against the values extracted into
hams.npz, and the set-up methods' return values before construction."""
import os

import numpy as np
import pytest

from ansatzevolve_amd import PyAnsatzEvolve as pae

HERE = os.path.dirname(os.path.abspath(__file__))
H4 = os.path.join(HERE, "golden", "runfiles", "H4", "L1", "H4")


def test_file_functions():
    d = np.load(os.path.join(HERE, "golden", "hams.npz"), allow_pickle=False)
    ops = pae.loadOperators(H4 + "_Operators.dat")
    assert np.array_equal(np.array(ops), d["H4_L1_ops1"])
    assert pae.loadPath(H4 + "_Operators.dat") == [(0, 0.0)] * len(ops)
    assert pae.loadNuclearEnergy(H4) == float(d["H4_L1_nuclear"])
    assert pae.loadNuclearEnergy(H4 + "_missing") == 0.0
    state = pae.readCsvState(H4 + "_Initial.dat")
    assert len(state) == 256 and np.flatnonzero(state).tolist() == d["H4_L1_init_idx0"].tolist()
    paths, order, P = pae.loadParameters(H4 + "_Order.dat", H4 + "_Parameters.dat", pae.loadPath(H4 + "_Operators.dat"))
    assert [p for p, _ in order] == d["H4_L1_order0"].tolist() and [s for _, s in order] == d["H4_L1_scale"].tolist()
    assert P == 9 and len(paths) == 3  # template (zero angles), placeholder set, the minimum
    want = d["H4_L1_scale"] * d["H4_L1_params"][0][d["H4_L1_order0"]]
    assert np.array_equal([a for _, a in paths[2]], want)
    assert all(a == 0.0 for _, a in paths[0])


def test_setup_methods():
    m = pae.stateAnsatzManager()
    with pytest.raises(ValueError):
        m.storeOperators([[1, 2, 0]])
    assert m.storeOperators(pae.loadOperators(H4 + "_Operators.dat"))
    assert m.storeInitial(8, [int(np.flatnonzero(pae.readCsvState(H4 + "_Initial.dat"))[0])], [1.0])
    assert m.storeRunPath(H4)
    assert not m.storeRunPath("")
    assert m.storeNuclearEnergy(1.5)


def test_file_function_edge_cases(tmp_path):
    # a bitstring alone means coefficient 1 (as the reference's reader assumes)
    f = tmp_path / "S_Initial.dat"
    f.write_text("0011\n1100,-0.5\n")
    st = pae.readCsvState(str(f))
    assert len(st) == 16 and st[3] == 1.0 and st[12] == -0.5 and sum(x != 0 for x in st) == 2
    # no parameter file: only the template path; the count comes from the order file
    (tmp_path / "S_Order.dat").write_text("1\n2,-1\n2\n")
    paths, order, P = pae.loadParameters(str(tmp_path / "S_Order.dat"), str(tmp_path / "none.dat"),
                                         [(0, 0.0)] * 3)
    assert paths == [[(0, 0.0)] * 3] and order == [(0, 1.0), (1, -1.0), (1, 1.0)] and P == 2
