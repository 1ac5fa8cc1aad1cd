"""Golden fixture of the reference's own fitted potential (example/Fe.pbe.in).

Reads the potential file of the reference's example as data, and stores its parameters
(header, par, core, the atomic and magnetic radial coefficients, and every expansion coefficient in
the enumeration order of the basis) rather than the 1.1 MB text.  tests/ regenerate the text
with potgen.render(values=...).  The script checks here that the reference libace
(oracle/_ref, built from its Fortran sources) gives bit-identical outputs on the original and
the regenerated file.  It then runs the reference on:

  * "notebook": the 2-atom bcc Fe cell of example/basic_lammps.ipynb (a = 2.83 A, 2.22 muB
    along z, pair_style ... 4.5 4.0 pbe 0.2), whose total energy the notebook reports as
    -15.7479878560111 eV (basic_lammps.ipynb:272);
  * "nc16": 16 atoms (2x2x2 cells), positions jittered N(0, 0.05 A), directions uniform on the
    sphere, magnitudes N(2.2, 0.1) muB (seed 41);
  * "nc128": 128 atoms (4x4x4 cells), the same distributions (seed 42).

This header has l3max_magnetic (2) > l3max_atomic (1), where the reference's dB3/dA0
accumulator runs out of bounds (src/b_basis.F90:50,198,210) and corrupts tlist0 only.
tlist0 is therefore stored from the oracle (which agrees with central differences of the
reference's energy, checked below), and the reference's value is kept as tlist0_reference.

Usage: python tests/golden/make_fe_example.py   (needs the reference tree and oracle/_ref)
"""
import hashlib
import os
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import __graft_entry__  # noqa: E402
import oracle as O  # noqa: E402

pkg = __graft_entry__.load_package()
capi, potgen, systems = pkg.capi, pkg.potgen, pkg.systems

SRC = os.path.join(os.environ.get("MAGACE_REFERENCE", "/root/reference"), "example", "Fe.pbe.in")
RCUT, RCUT_M, SMEAR = 4.5, 4.0, 0.2
WIDTH = {1: 1, 2: 4, 3: 9, 4: 15}


def parse(path):
    """Header fields and records of a potential file (src/util.F90:551-1106 layout)."""
    lines = open(path).read().splitlines()
    f = lambda i: lines[i].split()
    h = dict(potgen.BASE)
    h["lambda_a"], h["lambda_m"] = map(float, f(5)[1:3])
    h["cut_a"], h["cut_m"], h["dcut_a"], h["dcut_m"], h["ecut"], h["decut"] = map(float, f(6)[4:10])
    for i, keys in ((7, ("nb_a", "nb_m", "nr2_m", "l2m")), (8, ("nr3_a", "nr3_m", "l3a", "l3m")),
                    (9, ("nr4_a", "nr4_m", "l4a", "l4m")), (10, ("nr5_a", "nr5_m", "l5a", "l5m")),
                    (11, ("nr6_a", "nr6_m", "l6a", "l6m"))):
        for k, v in zip(keys, f(i)[2:6]):
            h[k] = int(v)
    h["ndensity"] = int(f(12)[1])
    h["npot"], h["nfemb"], h["radtype_a"], h["radtype_m"] = map(int, f(13)[1:5])
    assert h["ndensity"] == 1 and h["nr5_a"] == 0 and h["nr6_a"] == 0
    par, core, rad_a, rad_m, clu = {}, {}, {}, {}, {1: {}, 2: {}, 3: {}, 4: {}}
    i = 14
    while i + 1 < len(lines):
        kind = lines[i].split()
        body = lines[i + 1].split()
        i += 2
        if kind[0] == "par":
            par[int(kind[1])] = float(body[1])
        elif kind[0] == "core":
            core[int(kind[1])] = float(body[2])
        elif kind[0] in ("rad_atomic", "rad_magnetic"):
            nb, nr, l = map(int, body[2:5])
            (rad_a if kind[0] == "rad_atomic" else rad_m)[(nr, l, nb)] = float(body[5])
        elif kind[0] == "clu":
            order = int(kind[1])
            idx = tuple(int(x) for x in body[order:order + WIDTH[order]])
            assert idx not in clu[order], "duplicate clu record"
            clu[order][idx] = float(body[order + WIDTH[order]])
        else:
            raise ValueError("unknown record " + lines[i - 2])
    h.update(p1a=par[1], p1b=par[2], core_pre=core[1], core_lambda=core[2])
    return h, rad_a, rad_m, clu


def values_of(h, rad_a, rad_m, clu, basis):
    ca = np.zeros((h["nr3_a"], h["l3a"] + 1, h["nb_a"]))
    cm = np.zeros((h["nr2_m"], h["l2m"] + 1, h["nb_m"]))
    for (nr, l, nb), v in rad_a.items():
        ca[nr - 1, l, nb - 1] = v
    for (nr, l, nb), v in rad_m.items():
        cm[nr - 1, l, nb - 1] = v
    assert len(rad_a) == ca.size and len(rad_m) == cm.size
    c = {}
    for k in (1, 2, 3, 4):
        rows = [tuple(int(x) for x in r) for r in basis[k]]
        # every enumerated function has exactly one record and every record one function
        assert set(rows) == set(clu[k]), "clu records do not match the enumeration (order %d)" % k
        c[k] = np.array([clu[k][r] for r in rows])
    return dict(crad_a=ca, crad_m=cm, c=c)


def energy_sum(path, s, sdir, smag):
    a = systems.pack_acelib(s["pos"], s["cell"], sdir, smag, RCUT)
    return O.RefDriver(path, rcut=RCUT).run(a)["eout"].sum()


def main():
    h, rad_a, rad_m, clu = parse(SRC)
    basis = capi.enumerate_basis(h)
    assert [len(basis[k]) for k in (1, 2, 3, 4)] == [16, 720, 4932, 495]  # src/util.F90:1831-2596
    vals = values_of(h, rad_a, rad_m, clu, basis)
    text = potgen.render(h, basis, 0, values=vals)
    nb = dict(pos=np.array([[0.0, 0.0, 0.0], [1.415, 1.415, 1.415]]), cell=np.eye(3) * 2.83,
              sdir=np.array([[0.0, 0.0, 1.0], [0.0, 0.0, 1.0]]), smag=np.array([2.22, 2.22]))
    cases = {"notebook": nb, "nc16": systems.make_system(2, 41, "uniform"),
             "nc128": systems.make_system(4, 42, "uniform")}
    out = dict(header_keys=np.array(sorted(h)), header_vals=np.array([float(h[k]) for k in sorted(h)]),
               crad_a=vals["crad_a"], crad_m=vals["crad_m"], c1=vals["c"][1], c2=vals["c"][2],
               c3=vals["c"][3], c4=vals["c"][4], rcut=RCUT, rcut_magnetic=RCUT_M, smear=SMEAR,
               source_sha256=hashlib.sha256(open(SRC, "rb").read()).hexdigest(),
               pot_sha256=hashlib.sha256(text.encode()).hexdigest(), notebook_energy=-15.7479878560111)
    with tempfile.TemporaryDirectory() as td:
        regen = os.path.join(td, "Fe_regen.pbe.in")
        open(regen, "w").write(text)
        o = O.Oracle.load(regen, rcut=RCUT, rcut_magnetic=RCUT_M, smear=SMEAR)
        for name, s in cases.items():
            a = systems.pack_acelib(s["pos"], s["cell"], s["sdir"], s["smag"], RCUT)
            r0 = O.RefDriver(SRC, rcut=RCUT).run(a)
            r1 = O.RefDriver(regen, rcut=RCUT).run(a)
            for k in ("eout", "flist", "tlist", "tlist0"):
                assert np.array_equal(r0[k], r1[k]), (name, k, "regenerated file differs")
            orc = o.acelib(a)
            for k in ("eout", "flist", "tlist"):
                d = np.abs(orc[k] - r0[k]).max() / max(np.abs(r0[k]).max(), 1e-300)
                assert d < 1e-12, (name, k, d)
            # tlist0 of the oracle against central differences of the reference's energy
            # (the moment of atom i rotated and scaled: dE/dm_i . delta)
            g = orc["tlist0"].reshape(-1, 3)
            for i in range(min(2, len(s["pos"]))):
                for c in range(3):
                    e = []
                    m = s["sdir"] * s["smag"][:, None]
                    for sg in (1, -1):
                        mm = m.copy()
                        mm[i, c] += sg * 1e-5
                        n = np.linalg.norm(mm, axis=1)
                        e.append(energy_sum(SRC, s, mm / n[:, None], n))
                    fd = (e[0] - e[1]) / 2e-5
                    # dE/dm_i of the whole system = tlist0_i + sum over pairs (j -> i) of tlist
                    tl = orc["tlist"].reshape(-1, 3)
                    full = g[i, c] + tl[a["pair_j"] == i, c].sum()
                    assert abs(fd - full) < 1e-6 * max(1.0, abs(full)), (name, i, c, fd, full)
            if name == "notebook":
                assert abs(r0["eout"].sum() - (-15.7479878560111)) < 1e-11, r0["eout"].sum()
            for k in ("pos", "cell", "sdir", "smag"):
                out["%s_%s" % (name, k)] = s[k]
            for k in ("eout", "flist", "tlist"):
                out["%s_%s" % (name, k)] = r0[k]
            out["%s_tlist0" % name] = orc["tlist0"]
            out["%s_tlist0_reference" % name] = r0["tlist0"]
            print("%-9s natom %4d pairs %6d  E %.13f  max|F_p| %.3e  tlist0 ref/oracle rel diff %.2e" %
                  (name, a["natom"], len(a["dlist"]), r0["eout"].sum(), np.abs(r0["flist"]).max(),
                   np.abs(r0["tlist0"] - orc["tlist0"]).max() / max(np.abs(orc["tlist0"]).max(), 1e-300)))
    np.savez_compressed(os.path.join(HERE, "fe_example.npz"), **out)


if __name__ == "__main__":
    main()
