"""Host-side checks of the fitting support: coefficient get/set and the parameter-file
writer (record layout of src/util.F90:1110-1362), read back by the library, by the oracle
and, where it is built here, by the reference libace itself."""
import os

import numpy as np
import pytest

import oracle as O
from helpers import HEADERS, PKG

capi, potgen, systems, nbr = PKG.capi, PKG.potgen, PKG.systems, PKG.neighbors


def test_set_get_write_roundtrip(tmp_path):
    h = HEADERS["H500"]
    pot = str(tmp_path / "Fe.pbe.in")
    with open(pot, "w") as f:
        f.write(potgen.render(h, capi.enumerate_basis(h), 2050))
    capi.init_file(pot)
    c0 = capi.fit_get_coefficients()
    nb = capi.basis_counts()
    assert len(c0) == sum(nb)
    idx0, w0 = capi.model_terms()
    c1 = c0 * 1.5 + np.random.default_rng(1).normal(scale=1e-3, size=len(c0))
    capi.fit_set_coefficients(c1)
    idx1, w1 = capi.model_terms()
    assert np.array_equal(idx0, idx1)  # the monomial set does not depend on the coefficients
    assert not np.allclose(w0, w1)
    out = str(tmp_path / "written" / "Fe.pbe.in")
    os.makedirs(os.path.dirname(out))
    capi.write_potential(out)
    capi.init_file(out)
    assert np.array_equal(capi.fit_get_coefficients(), c1)  # 17 significant digits round trip
    idx2, w2 = capi.model_terms()
    assert np.array_equal(idx2, idx1) and np.array_equal(w2, w1)
    with pytest.raises(capi.MagaceError):
        capi.fit_set_coefficients(c1[:-1])


def test_written_file_reads_in_the_oracle_and_the_reference(tmp_path):
    h = HEADERS["H500"]
    pot = str(tmp_path / "Fe.pbe.in")
    with open(pot, "w") as f:
        f.write(potgen.render(h, capi.enumerate_basis(h), 2051))
    capi.init_file(pot)
    c = capi.fit_get_coefficients() * -0.7
    capi.fit_set_coefficients(c)
    out = str(tmp_path / "w.pbe.in")
    capi.write_potential(out)
    # the oracle parses the written file like the original with the same coefficients scaled
    s = systems.make_system(2, 3, "uniform")
    a = systems.pack_acelib(s["pos"], s["cell"], s["sdir"], s["smag"], 5.0)
    e_w = O.Oracle.load(out).acelib(a)["eout"]
    e_0 = O.Oracle.load(pot).acelib(a)["eout"]
    # E_i = c.B_i (linear embedding, no core): scaling every coefficient scales E
    np.testing.assert_allclose(e_w, -0.7 * e_0, rtol=1e-12, atol=1e-14)
    if os.path.exists(O.REF_DRIVER):
        ref = O.RefDriver(out).run(a)
        np.testing.assert_allclose(ref["eout"], e_w, rtol=1e-10, atol=1e-14)


def test_two_densities_write_roundtrip(tmp_path):
    """ndensity 2: per-density coefficients and embedding parameters survive the writer, and
    the oracle and the reference read the written file to the same energies."""
    h = HEADERS["H500ND2"]
    pot = str(tmp_path / "Fe.pbe.in")
    with open(pot, "w") as f:
        f.write(potgen.render(h, capi.enumerate_basis(h), 2052))
    capi.init_file(pot)
    c = capi.fit_get_coefficients()
    assert len(c) == 2 * sum(capi.basis_counts())
    out = str(tmp_path / "w.pbe.in")
    capi.write_potential(out)
    capi.init_file(out)
    assert np.array_equal(capi.fit_get_coefficients(), c)
    s = systems.make_system(2, 4, "uniform")
    a = systems.pack_acelib(s["pos"], s["cell"], s["sdir"], s["smag"], 5.0)
    e0 = O.Oracle.load(pot).acelib(a)["eout"]
    np.testing.assert_allclose(O.Oracle.load(out).acelib(a)["eout"], e0, rtol=1e-13, atol=1e-14)
    if os.path.exists(O.REF_DRIVER):
        np.testing.assert_allclose(O.RefDriver(out).run(a)["eout"], e0, rtol=1e-10, atol=1e-14)
