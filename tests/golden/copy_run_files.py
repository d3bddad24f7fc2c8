"""Copy three of the reference's input data sets (Hams/H4/L1, Hams/H6/L1,
Hams/TestFermionic) into tests/golden/runfiles/ as fixtures of the file
front end.  Data only: start state (_Initial.dat), operators
(_Operators.dat), order file (_Order.dat), parameter sets with their
energies (_Parameters.dat), nuclear energy (_Nuclear_Energy.dat), the
integrals (_oneEInts.bin, _twoEInts.bin) or the text Hamiltonian
(_Ham_Coeff.dat, _Ham_Index.dat), byte for byte.

Run in the build container:  python tests/golden/copy_run_files.py
"""
import os
import shutil

REF = os.environ.get("ANSATZ_REFERENCE", "/root/reference") + "/Hams"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "runfiles")
SETS = [("H4/L1", "H4"), ("H6/L1", "H6"), ("TestFermionic", "TestFermionic")]


def main():
    for sub, prefix in SETS:
        src = os.path.join(REF, sub)
        dst = os.path.join(OUT, sub)
        os.makedirs(dst, exist_ok=True)
        for name in sorted(os.listdir(src)):
            if name.startswith(prefix + "_"):
                shutil.copyfile(os.path.join(src, name), os.path.join(dst, name))
                os.chmod(os.path.join(dst, name), 0o644)
        print(sub, sorted(os.listdir(dst)))


if __name__ == "__main__":
    main()
