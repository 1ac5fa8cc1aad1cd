"""The reference's fitted example potential (example/Fe.pbe.in) on the CPU.

golden/fe_example.npz holds its parameters and the reference libace's outputs on three
systems (tests/golden/make_fe_example.py): the notebook's 2-atom cell, whose energy the
notebook reports as -15.7479878560111 eV (example/basic_lammps.ipynb:272), and 16- and
128-atom non-collinear systems.  Here: the regenerated file parses into the product's
basis with the counts the reference prints (src/util.F90:1831,1929,2235,2596), and the
oracle reproduces the reference's outputs from it (tlist0: the fixture holds the oracle's
value, checked there against central differences, because the reference's own is corrupted
for this header, src/b_basis.F90:50,198,210).
"""
import numpy as np

import oracle as O
from helpers import FE_CASES, PKG, fe_case, fe_example, relerr


def test_regenerated_file_parses_with_the_reference_counts(tmp_path):
    path, z = fe_example(tmp_path)
    PKG.capi.init_file(path, rcut=float(z["rcut"]), rcut_magnetic=float(z["rcut_magnetic"]),
                       smear=float(z["smear"]))
    assert PKG.capi.basis_counts() == [16, 720, 4932, 495, 0, 0]


def test_oracle_reproduces_reference_outputs(tmp_path):
    path, z = fe_example(tmp_path)
    o = O.Oracle.load(path, rcut=float(z["rcut"]), rcut_magnetic=float(z["rcut_magnetic"]), smear=float(z["smear"]))
    for name in FE_CASES:
        a, ref = fe_case(z, name)
        got = o.acelib(a)
        for k in ("eout", "flist", "tlist", "tlist0"):
            assert relerr(got[k], ref[k]) < 1e-12, (name, k)
        if name == "notebook":
            assert abs(got["eout"].sum() - float(z["notebook_energy"])) < 1e-11
        else:
            # the reference's own tlist0 differs (out-of-bounds accumulator)
            assert relerr(z[name + "_tlist0_reference"], ref["tlist0"]) > 1e-3
