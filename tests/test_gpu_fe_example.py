"""The reference's fitted example potential (example/Fe.pbe.in) through the HIP path.

The potential is regenerated from golden/fe_example.npz (its parameters; the reference libace
gives bit-identical outputs on the original and the regenerated file).  Bars:
  * FP64: eout, flist, tlist within 1e-10 (norm-wise relative) of the reference libace;
    tlist0 within 1e-10 of the oracle (for this header, l3max_magnetic > l3max_atomic, the
    reference's own tlist0 is corrupted by an out-of-bounds accumulator,
    src/b_basis.F90:50,198,210; the oracle's agrees with central differences);
  * the notebook's total energy, -15.7479878560111 eV (example/basic_lammps.ipynb:272), to
    1e-10 relative;
  * mixed precision: per-atom energy, assembled force and magnetic force within 1e-4 (eV,
    eV/A, eV/muB) of the reference.
"""
import sys

import numpy as np
import pytest

from helpers import FE_CASES, PKG, fe_case, fe_example, per_atom_errors, relerr

capi = PKG.capi
torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no GPU visible", allow_module_level=True)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def fe(tmp_path_factory):
    path, z = fe_example(tmp_path_factory.mktemp("fe"))
    return path, z


def init(path, z):
    capi.init_file(path, rcut=float(z["rcut"]), rcut_magnetic=float(z["rcut_magnetic"]), smear=float(z["smear"]))


@pytest.mark.parametrize("name", FE_CASES)
def test_fp64_matches_reference(fe, name):
    path, z = fe
    say("example potential, fp64,", name)
    init(path, z)
    capi.set_precision(capi.FP64)
    a, ref = fe_case(z, name)
    got = capi.acelib(a)
    errs = {k: relerr(got[k], ref[k]) for k in ("eout", "flist", "tlist", "tlist0")}
    say("  rel errors", errs)
    for k, e in errs.items():
        assert e <= 1e-10, (k, e)
    if name == "notebook":
        e = got["eout"].sum()
        want = float(z["notebook_energy"])
        say("  notebook energy %.13f (published %.13f)" % (e, want))
        assert abs(e - want) <= 1e-10 * abs(want)


@pytest.mark.parametrize("name", FE_CASES)
def test_mixed_within_targets(fe, name):
    path, z = fe
    say("example potential, mixed,", name)
    init(path, z)
    capi.set_precision(capi.MIXED)
    a, ref = fe_case(z, name)
    got = capi.acelib(a)
    capi.set_precision(capi.FP64)
    de, df, dm = per_atom_errors(a, got, ref)
    say("  max|dE_i| %.3e eV  max|dF_i| %.3e eV/A  max|dfm_i| %.3e eV/muB" % (de, df, dm))
    assert de <= 1e-4 and df <= 1e-4 and dm <= 1e-4
