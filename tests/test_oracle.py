"""Pins of the oracle (oracle/magace_oracle.c), the CPU checker of the GPU path.

Every oracle function is checked against something that does not restate it:
  * golden vectors written by the reference libace itself (tests/golden/make_golden.py);
  * the reference binary directly, when it is built here (oracle/_ref);
  * closed forms: sympy Clebsch-Gordan, scipy spherical harmonics, Chebyshev identities,
    Hermite-spline node identities;
  * central finite differences of the energy for forces and magnetic gradients;
  * invariances: rotation, time reversal;
  * the enumeration counts printed by the reference for example/Fe.pbe.in and the
    published notebook energy (both only where the reference tree is present).
"""
import os

import numpy as np
import pytest

import oracle as O
from helpers import ORACLE_CASES, PKG, load_case, relerr, write_potential

REF_TREE = "/root/reference"


@pytest.mark.parametrize("name", ORACLE_CASES)
def test_oracle_matches_reference_golden(name, tmp_path):
    c = load_case(name, tmp_path, enum="oracle")
    got = O.Oracle.load(c["pot"]).acelib(c["a"])
    for k in ("eout", "flist", "tlist", "tlist0"):
        assert relerr(got[k], c["ref"][k]) < 1e-12, k


@pytest.mark.skipif(not os.path.exists(O.REF_DRIVER), reason="reference driver not built here")
def test_reference_driver_reproduces_golden(tmp_path):
    c = load_case("tiny_16", tmp_path, enum="oracle")
    out = O.RefDriver(c["pot"]).run(c["a"])
    for k in ("eout", "flist", "tlist", "tlist0"):
        assert relerr(out[k], c["ref"][k]) < 1e-14, k


def test_reference_self_torque_defect_is_recorded(tmp_path):
    """hfe_16: l3max_magnetic > l3max_atomic makes the reference's dB3/dA0 accumulator run
    out of bounds (src/b_basis.F90:50,198,210); its tlist0 then differs from the gradient."""
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "hfe_16.npz"))
    assert bool(z["tlist0_from_oracle"])
    assert relerr(z["tlist0_reference"], z["tlist0"]) > 1e-2


def test_reference_force_defect_beyond_l3max_magnetic_is_recorded():
    """l4m_16 (l4max_magnetic 2 > l3max_magnetic 1): the reference keeps A zero beyond
    l3max_magnetic (src/a_basis.F90:61,248-251) but its force pass differentiates those
    entries, so its forces and magnetic forces miss the central differences of its own
    energy stored in the fixture by 0.1-1%; the library is held to the differences."""
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "l4m_16.npz"))
    a = PKG.systems.pack_acelib(z["pos"], z["cell"], z["sdir"], z["smag"], float(z["rlist"]))
    _, f, fm, _ = PKG.systems.assemble(a, z["eout"], z["flist"], z["tlist"], z["tlist0"])
    assert relerr(f, z["fd_force"]) > 1e-3
    assert relerr(fm, z["fd_mforce"]) > 1e-4


@pytest.mark.parametrize("hname,counts", [("H500", [3, 120, 180, 186, 0, 0]), ("H2000", [3, 210, 735, 1095, 0, 0]),
                                           ("HTINY", [2, 36, 30, 186, 0, 0])])
def test_enumeration_counts(hname, counts):
    from helpers import HEADERS
    assert O.Oracle.from_header(HEADERS[hname]).counts() == counts


@pytest.mark.skipif(not os.path.isdir(REF_TREE), reason="reference tree not mounted")
def test_fe_pbe_enumeration_and_notebook_energy():
    """example/Fe.pbe.in: the reference prints 16/720/4932/495 basis functions
    (src/util.F90:1831,1929,2235,2596); the notebook reports E = -15.7479878560111 eV for
    2-atom bcc Fe, a = 2.83 A, 2.22 muB along z (example/basic_lammps.ipynb:272)."""
    path = os.path.join(REF_TREE, "example", "Fe.pbe.in")
    o = O.Oracle.load(path, rcut=4.5, rcut_magnetic=4.0, smear=0.2)
    assert o.counts() == [16, 720, 4932, 495, 0, 0]
    pos = np.array([[0.0, 0.0, 0.0], [1.415, 1.415, 1.415]])
    sdir = np.array([[0.0, 0.0, 1.0], [0.0, 0.0, 1.0]])
    a = PKG.systems.pack_acelib(pos, np.eye(3) * 2.83, sdir, np.array([2.22, 2.22]), 4.5)
    e = o.acelib(a)["eout"].sum()
    assert abs(e - (-15.7479878560111)) < 1e-11
    if os.path.exists(O.REF_DRIVER):
        r = O.RefDriver(path, rcut=4.5).run(a)
        assert abs(r["eout"].sum() - e) < 1e-12


def test_clebsch_gordan_vs_sympy():
    from sympy import S
    from sympy.physics.quantum.cg import CG
    rng = np.random.default_rng(0)
    for _ in range(300):
        j1, j2 = rng.integers(0, 7, size=2)
        J = rng.integers(abs(j1 - j2), j1 + j2 + 1)
        m1 = rng.integers(-j1, j1 + 1)
        m2 = rng.integers(-j2, j2 + 1)
        if abs(m1 + m2) > J:
            continue
        ref = float(CG(S(int(j1)), S(int(m1)), S(int(j2)), S(int(m2)), S(int(J)), S(int(m1 + m2))).doit())
        assert abs(O.cg(int(j1), int(j2), int(J), int(m1), int(m2), int(m1 + m2)) - ref) < 1e-12
    # orthogonality sum_{m1,m2} <j1 m1 j2 m2|J M><j1 m1 j2 m2|J' M> = delta_JJ'
    j1, j2 = 3, 2
    for J in range(1, 6):
        for Jp in range(1, 6):
            s = sum(O.cg(j1, j2, J, m1, 1 - m1, 1) * O.cg(j1, j2, Jp, m1, 1 - m1, 1)
                    for m1 in range(-j1, j1 + 1) if abs(1 - m1) <= j2)
            assert abs(s - (J == Jp)) < 1e-12


def test_ylm_vs_scipy_and_gradient():
    from scipy.special import sph_harm_y
    rng = np.random.default_rng(1)
    L = 6
    for _ in range(20):
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        th, ph = np.arccos(v[2]), np.arctan2(v[1], v[0])
        y, dy = O.ylm(L, v)
        for l in range(L + 1):
            for m in range(-l, l + 1):
                ref = np.sqrt(4 * np.pi) * sph_harm_y(l, m, th, ph)
                assert abs(y[l * l + l + m] - ref) < 1e-12 * max(1, abs(ref))
        # tangential gradient: directional derivative along a tangent vector
        t = np.cross(v, rng.normal(size=3))
        t /= np.linalg.norm(t)
        h = 1e-6
        yp, _ = O.ylm(L, (v + h * t) / np.linalg.norm(v + h * t))
        ym, _ = O.ylm(L, (v - h * t) / np.linalg.norm(v - h * t))
        fd = (yp - ym) / (2 * h)
        an = t @ dy
        assert np.abs(fd - an).max() < 1e-6 * max(1, np.abs(an).max())
        # the gradient is tangential
        assert np.abs(v @ dy).max() < 1e-12 * max(1, np.abs(dy).max())


def test_radial_basis_closed_forms():
    # radtype 7: g_n = T_n(1 - 2 (r/rc)^2) = cos(n acos x)
    rc = 4.0
    for r in (0.3, 1.7, 3.1):
        g, dg = O.radbase(7, 5.25, 6, rc, 0.01, r)
        x = 1 - 2 * (r / rc) ** 2
        assert np.allclose(g, np.cos(np.arange(1, 7) * np.arccos(x)), atol=1e-13)
        gp, _ = O.radbase(7, 5.25, 6, rc, 0.01, r + 1e-6)
        gm, _ = O.radbase(7, 5.25, 6, rc, 0.01, r - 1e-6)
        assert np.allclose(dg, (gp - gm) / 2e-6, atol=1e-7)
    # radtype 3: g_1 = env, g_n = (1 - T_{n-1}(x))/2 * env with the exponential variable
    rc, lam = 5.0, 5.25
    for r in (1.1, 2.5, 4.2, 4.995):
        g, dg = O.radbase(3, lam, 8, rc, 0.01, r)
        x = 1 - 2 * (np.exp(-lam * r / rc) - np.exp(-lam)) / (1 - np.exp(-lam))
        env = 0.5 * (1 + np.cos(np.pi * r / rc))
        ref = np.concatenate([[1.0], 0.5 - 0.5 * np.cos(np.arange(1, 8) * np.arccos(x))]) * env
        if r > rc - 0.01:
            ref = ref * 0.5 * (1 + np.cos(np.pi * (r - (rc - 0.01)) / 0.01))
        assert np.allclose(g, ref, atol=1e-13)
        gp, _ = O.radbase(3, lam, 8, rc, 0.01, r + 1e-7)
        gm, _ = O.radbase(3, lam, 8, rc, 0.01, r - 1e-7)
        assert np.allclose(dg, (gp - gm) / 2e-7, atol=2e-5)
    g, dg = O.radbase(3, lam, 8, rc, 0.01, rc)
    assert not np.any(g) and not np.any(dg)
    # core repulsion: pre e^{-lam r^2}/r * cosine envelope
    cr, dcr = O.radcore(1.3, 2.0, 0.5, 5.0)
    assert abs(cr - 2.0 * np.exp(-0.5 * 1.69) / 1.3 * 0.5 * (1 + np.cos(np.pi * 1.3 / 5.0))) < 1e-14


def test_spline_tables_interpolate_node_values(tmp_path):
    """c0 of every row is the function at the node (src/radial.F90:245): at r = n h the
    lookup returns the analytic radial basis; between nodes the cubic Hermite error is
    tiny for these smooth functions."""
    path = write_potential(str(tmp_path / "p.in"), "H500", 3, enum="oracle")
    o = O.Oracle.load(path)
    from helpers import HEADERS
    h = HEADERS["H500"]
    step = 5.0 / 10000
    for n in (137, 2000, 4567, 9000):
        r = n * step
        val, dval = o.spline(0, r)
        g, dg = O.radbase(3, h["lambda_a"], h["nb_a"], h["cut_a"], h["dcut_a"], r)
        assert np.allclose(val, g, atol=1e-14) and np.allclose(dval, dg, atol=1e-10)
        rm = r + 0.37 * step
        val, dval = o.spline(0, rm)
        g, dg = O.radbase(3, h["lambda_a"], h["nb_a"], h["cut_a"], h["dcut_a"], rm)
        assert np.allclose(val, g, atol=1e-10) and np.allclose(dval, dg, atol=1e-6)
    assert not np.any(o.spline(0, 5.0)[0])  # beyond the table: zero


def _total_energy(o, pos, cell, sdir, smag, rcut=5.0):
    a = PKG.systems.pack_acelib(pos, cell, sdir, smag, rcut)
    return o.acelib(a)["eout"].sum()


@pytest.mark.parametrize("hname", ["HTINY", "H500"])
def test_forces_and_magnetic_gradients_by_central_differences(hname, tmp_path):
    path = write_potential(str(tmp_path / "p.in"), hname, 21, enum="oracle")
    o = O.Oracle.load(path)
    s = PKG.systems.make_system(2, 21, "uniform")
    pos, cell, sdir, smag = s["pos"], s["cell"], s["sdir"], s["smag"]
    a = PKG.systems.pack_acelib(pos, cell, sdir, smag, 5.0)
    r = o.acelib(a)
    _, f, fm, vir = PKG.systems.assemble(a, r["eout"], r["flist"], r["tlist"], r["tlist0"])
    h = 1e-5
    scale_f, scale_m = np.abs(f).max(), np.abs(fm).max()
    for k in (0, 5):
        for c in range(3):
            p1, p2 = pos.copy(), pos.copy()
            p1[k, c] += h
            p2[k, c] -= h
            fd = -(_total_energy(o, p1, cell, sdir, smag) - _total_energy(o, p2, cell, sdir, smag)) / (2 * h)
            assert abs(fd - f[k, c]) < 1e-6 * scale_f + 1e-9
            m = sdir * smag[:, None]
            e = []
            for sg in (1, -1):
                mm = m.copy()
                mm[k, c] += sg * h
                nm = np.linalg.norm(mm, axis=1)
                e.append(_total_energy(o, pos, cell, mm / nm[:, None], nm))
            fd = -(e[0] - e[1]) / (2 * h)
            assert abs(fd - fm[k, c]) < 1e-6 * scale_m + 1e-9
    # virial: dE/dF_ab under a homogeneous deformation u_p -> F u_p of every pair vector
    # equals sum_p F_p,a u_p,b (F_p = dE/du_p), the tally of pair_spin_ace.cpp:635-636
    eps = 1e-6
    u = a["elist"].reshape(-1, 3) * a["dlist"][:, None]
    for a_, b_ in ((0, 0), (0, 1), (2, 1)):
        e = []
        for sg in (1, -1):
            F = np.eye(3)
            F[a_, b_] += sg * eps
            up = u @ F.T
            d = np.linalg.norm(up, axis=1)
            b = dict(a, dlist=d, elist=(up / d[:, None]).ravel())
            e.append(o.acelib(b)["eout"].sum())
        dE = (e[0] - e[1]) / (2 * eps)
        assert abs(dE - vir[b_, a_]) < 1e-6 * max(1.0, np.abs(vir).max())


def test_rotation_and_time_reversal_invariance(tmp_path):
    path = write_potential(str(tmp_path / "p.in"), "H500", 5, enum="oracle")
    o = O.Oracle.load(path)
    s = PKG.systems.make_system(2, 5, "uniform")
    e0 = _total_energy(o, s["pos"], s["cell"], s["sdir"], s["smag"])
    from scipy.spatial.transform import Rotation
    R = Rotation.from_rotvec([0.3, -1.1, 0.7]).as_matrix()
    a = PKG.systems.pack_acelib(s["pos"], s["cell"], s["sdir"], s["smag"], 5.0)
    a_rot = dict(a)
    a_rot["elist"] = (a["elist"].reshape(-1, 3) @ R.T).ravel()
    a_rot["elist_mom"] = (a["elist_mom"].reshape(-1, 3) @ R.T).ravel()
    a_rot["elist_momi"] = (a["elist_momi"].reshape(-1, 3) @ R.T).ravel()
    e_rot = o.acelib(a_rot)["eout"].sum()
    assert abs(e_rot - e0) < 1e-11 * max(1, abs(e0))
    e_tr = _total_energy(o, s["pos"], s["cell"], -s["sdir"], s["smag"])
    assert abs(e_tr - e0) < 1e-11 * max(1, abs(e0))
