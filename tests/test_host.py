"""CPU tests of the product's host side (no GPU): the C-ABI library, the C++ basis
enumeration, the merged monomial table, potential-file handling and error paths."""
import hashlib
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
from helpers import CASES, GOLDEN, HEADERS, HIGHER_HEADERS, PKG, potential_text, write_potential

capi = PKG.capi
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    header = open(os.path.join(ROOT, "include", "magace.h")).read()
    declared = set(re.findall(r"^\s*(?:void|int|const char \*)\s*\*?\s*(\w+)\s*\(", header, re.M))
    assert declared == set(capi.EXPORTS)
    out = subprocess.run(["nm", "-D", "--defined-only", capi.LIBPATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    missing = declared - exported
    assert not missing, missing
    L = capi.lib()
    for name in declared:
        assert hasattr(L, name)


@pytest.mark.parametrize("hname", sorted(h for h in HEADERS if h not in HIGHER_HEADERS))
def test_enumeration_matches_oracle(hname):
    h = HEADERS[hname]
    mine = capi.enumerate_basis(h)
    o = O.Oracle.from_header(h)
    for k in (1, 2, 3, 4, 5, 6):
        assert np.array_equal(mine[k], o.basis(k)), (hname, k)


@pytest.mark.parametrize("name", ["o5_16", "o56_16", "l4m_16"])
def test_higher_order_enumeration_counts_match_reference(name):
    """Body orders 5 and 6: the numbers of B functions of orders 1..6 equal the counts the
    reference printed when it wrote the fixture (src/util.F90:1831-3755), and the fixture's
    potential, whose clu records carry this enumeration's index tuples, regenerates bit for
    bit (the reference matched every record to one of its functions, or its outputs, checked
    on the GPU, would differ)."""
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    b = capi.enumerate_basis(HEADERS[str(z["header"])])
    assert [len(b[k]) for k in range(1, 7)] == [int(x) for x in z["reference_counts"]]


def test_higher_order_structure():
    """5-body tuples (b5index columns): radial-only functions couple every l to 0; the
    6-body ordering keys use the 5-body maxima, as the reference's do."""
    b = capi.enumerate_basis(HEADERS["H56"])
    assert b[5].shape[1] == 21 and b[6].shape[1] == 27
    assert np.all(b[5][:, 9:] == 0) and np.all(b[6][:, 11:] == 0)
    # combined (atomic, magnetic) keys are non-increasing along each tuple
    for row in b[5]:
        keys = [(row[i], row[5 + i]) for i in range(4)]
        assert keys == sorted(keys, reverse=True)


@pytest.mark.parametrize("extra,msg", [
    (dict(nr5_a=1, nr5_m=1, l5a=1), "pairs 2 atomic with 1 magnetic"),
    (dict(nr5_a=1, nr5_m=1, l5m=1), "pairs 1 atomic with 2 magnetic"),
    (dict(nr6_a=1, nr6_m=1), "without 5-body"),
    (dict(nr4_m=2), "above nradial3_magnetic"),
])
def test_header_refusals(extra, msg):
    """Headers the reference cannot evaluate are refused with the reason: higher-order
    couplings it pairs with zero indices, 6-body terms without 5-body ones (its prepB skips
    them), and radial indices of A beyond its array."""
    h = dict(HEADERS["H500"], nr2_m=2, **extra)
    with pytest.raises(capi.MagaceError, match=msg):
        capi.enumerate_basis(h)


def test_fourth_order_magnetic_l_beyond_third():
    """l4max_magnetic above l3max_magnetic: the reference enumerates 4-body functions up to
    l4max_magnetic but fills A only up to l3max_magnetic (src/a_basis.F90:248-251), so the
    functions with a factor of higher l' vanish (golden case l4m_16 checks the outputs on
    the GPU)."""
    b = capi.enumerate_basis(HEADERS["HL4"])
    assert HEADERS["HL4"]["l3m"] == 1 and b[4][:, 11:14].max() == 2


@pytest.mark.parametrize("name", CASES)
def test_potential_regenerates_bitwise(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"))
    text = potential_text(str(z["header"]), int(z["pot_seed"]), enum="product")
    assert hashlib.sha256(text.encode()).hexdigest() == str(z["pot_sha256"])


def test_reference_example_potential_counts():
    path = "/root/reference/example/Fe.pbe.in"
    if not os.path.exists(path):
        pytest.skip("reference tree not mounted")
    capi.init_file(path, rcut=4.5, rcut_magnetic=4.0, smear=0.2)
    assert capi.basis_counts() == [16, 720, 4932, 495, 0, 0]


def _clu1(text):
    lines = text.splitlines()
    return [float(lines[i + 1].split()[2]) for i, l in enumerate(lines) if l.split()[:2] == ["clu", "1"]]


def _zvector(o, header, a, i, nz, oA2, oA0):
    """Z = [A | A2 | A0] of atom i from the oracle's splines and spherical harmonics."""
    h = dict(PKG.potgen.BASE, **header)
    lsa, lsm, lsm3 = (h["l3a"] + 1) ** 2, (h["l2m"] + 1) ** 2, (h["l3m"] + 1) ** 2
    nm3 = h["nr3_m"] * lsm3
    Z = np.zeros(nz + 1, complex)
    Z[nz] = 1.0
    lo, hi = a["nstart"][i] - 1, a["nstop"][i]
    for p in range(lo, hi):
        g, _ = o.spline(0, a["dlist"][p])
        R, _ = o.spline(1, a["dlist"][p])          # [n*(l3a+1) + l]
        M, _ = o.spline(4, a["dlist_mom"][p])      # [n'*(l2m+1) + l']
        Y, _ = O.ylm(h["l3a"], a["elist"][3 * p:3 * p + 3])
        Ys, _ = O.ylm(h["l2m"], a["elist_mom"][3 * p:3 * p + 3])
        for n in range(h["nr3_a"]):
            for l in range(h["l3a"] + 1):
                for m in range(-l, l + 1):
                    lm = l * l + l + m
                    ry = R[n * (h["l3a"] + 1) + l] * Y[lm]
                    for npr in range(h["nr3_m"]):
                        for lp in range(h["l3m"] + 1):
                            for mp in range(-lp, lp + 1):
                                q = npr * lsm3 + lp * lp + lp + mp
                                Z[(n * lsa + lm) * nm3 + q] += ry * M[npr * (h["l2m"] + 1) + lp] * Ys[lp * lp + lp + mp]
        for n in range(h["nb_a"]):
            for npr in range(h["nr2_m"]):
                for lp in range(h["l2m"] + 1):
                    for mp in range(-lp, lp + 1):
                        Z[oA2 + (n * h["nr2_m"] + npr) * lsm + lp * lp + lp + mp] += \
                            g[n] * M[npr * (h["l2m"] + 1) + lp] * Ys[lp * lp + lp + mp]
    M, _ = o.spline(4, a["dlist_momi"][i])
    Ys, _ = O.ylm(h["l2m"], a["elist_momi"][3 * i:3 * i + 3])
    for npr in range(h["nr2_m"]):
        for lp in range(h["l2m"] + 1):
            for mp in range(-lp, lp + 1):
                Z[oA0 + npr * lsm + lp * lp + lp + mp] = M[npr * (h["l2m"] + 1) + lp] * Ys[lp * lp + lp + mp]
    return Z


@pytest.mark.parametrize("name", ["tiny_16", "h2000_16"])
def test_monomial_table_reproduces_the_reference_energy(name, tmp_path):
    """The merged table rho = Re sum_t w_t prod Z replaces the B-function loops of
    src/b_basis.F90:91-307; with linear embedding and no core E_i = rho_i exactly."""
    from helpers import load_case
    c = load_case(name, tmp_path)
    header = HEADERS[c["header"]]
    capi.init_file(c["pot"])
    info = capi.model_info()
    idx, w = capi.model_terms()
    assert info["terms"] == len(w) and info["terms"] < info["raw"]
    o = O.Oracle.load(c["pot"])
    c1 = _clu1(open(c["pot"]).read())
    for i in (0, 7):
        Z = _zvector(o, header, c["a"], i, info["nz"], info["oA2"], info["oA2"] + header["nb_a"] * header["nr2_m"] * (header["l2m"] + 1) ** 2)
        rho = np.real(np.sum(w * np.prod(Z[idx], axis=1)))
        gm, _ = o.spline(3, c["a"]["dlist_momi"][i])
        rho += np.dot(c1, gm - 1.0)
        assert abs(rho - c["ref"]["eout"][i]) < 1e-12 * max(1.0, abs(c["ref"]["eout"][i]))


def test_init_error_paths(tmp_path):
    pot = write_potential(str(tmp_path / "p.in"), "HTINY", 1)
    with pytest.raises(capi.MagaceError, match="smearing"):
        capi.init_file(pot, smear=0.3)
    with pytest.raises(capi.MagaceError, match="element"):
        capi.init_file(pot, element="Co")
    with pytest.raises(capi.MagaceError, match="cutoff"):
        capi.init_file(pot, rcut=4.0)
    with pytest.raises(capi.MagaceError, match="not found"):
        capi.init_file(str(tmp_path / "missing.in"))
    # a 5-body header with angular couplings the reference cannot pair
    bad = open(pot).read().replace("nradial5 l5max 0 0 0 0", "nradial5 l5max 1 1 1 0")
    open(tmp_path / "b.in", "w").write(bad)
    with pytest.raises(capi.MagaceError, match="5-body basis pairs"):
        capi.init_file(str(tmp_path / "b.in"))
    capi.init_file(pot)  # a good file still loads after failures


def test_acelibini_reads_cwd_and_enforces_single_element(tmp_path):
    """acelibini_ semantics in a fresh process: <El>.<xc>.in from the CWD, nele == 1."""
    write_potential(str(tmp_path / "Fe.pbe.in"), "HTINY", 1)
    code = r'''
import sys
sys.path.insert(0, %r)
import __graft_entry__
capi = __graft_entry__.load_package().capi
try:
    capi.acelibini(2, 5.0, 4.0, 0.2, 1, [27, 28])
    print("no error")
except capi.MagaceError as e:
    print("multi:", e)
capi.acelibini(1, 5.0, 4.0, 0.2, 1, [27])
print("counts", capi.basis_counts())
''' % ROOT
    p = subprocess.run([sys.executable, "-c", code], cwd=str(tmp_path), capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    assert "multi: Multiple elements" in p.stdout
    assert "counts [2, 36, 30, 186, 0, 0]" in p.stdout


@pytest.mark.parametrize("hname", ["HTINY", "H500", "H2000", "H500ND2", "H56"])
def test_gradient_records_equal_the_monomial_derivative(hname, tmp_path):
    """The adjoint's transposed table reproduces dRho/dZ[x] of the monomial table for a
    random complex Z, entry by entry: the brute-force derivative differentiates every slot
    of every monomial (products Y of body orders 5 and 6 are formed from Z, and a Y target's
    gradient goes back to its two factors).  Each chunk header holds its group's length: the
    64 chunks of a group have equal lengths, and records past a chunk's end carry weight 0; a
    chunk is flagged sole exactly when it is its target's only one (the kernel then stores
    instead of adding)."""
    pot = write_potential(str(tmp_path / "Fe.pbe.in"), hname, 31)
    capi.init_file(pot)
    nz = capi.model_info()["nz"]
    idx, w = capi.model_terms()
    ch, ridx, rw = capi.model_records()
    prod = capi.model_products()
    ny = len(prod)
    rng = np.random.default_rng(5)
    Z = np.zeros(nz + 1 + len(prod), complex)
    Z[:nz] = rng.normal(size=nz) + 1j * rng.normal(size=nz)
    Z[nz] = 1.0
    for i, (a, b) in enumerate(prod):
        assert max(a, b) < nz + 1 + i  # factors come before the product
        Z[nz + 1 + i] = Z[a] * Z[b]
    nzx = nz + 1 + ny
    assert (ny > 0) == (hname == "H56")
    want = np.zeros(nzx, complex)
    F = Z[idx]
    for s in range(4):
        np.add.at(want, idx[:, s], w * np.prod(np.delete(F, s, axis=1), axis=1))
    y = ch[:, 1].astype(np.int64) & 0xFFFFFFFF
    first, x, cnt, sole, slot = ch[:, 0], y & 0xFFF, (y >> 12) & 63, (y >> 18) & 1, y >> 19
    assert np.all(x < nzx) and np.all(cnt % 4 == 0) and np.all((cnt >= 4) & (cnt <= 32))
    assert np.all(ridx[:, :3] < nzx)
    nchunk_of = np.bincount(x, minlength=nz)
    assert np.array_equal(sole == 1, nchunk_of[x] == 1)  # flagged exactly when the target has one chunk
    # runs (a lane's chunks in consecutive rounds with one target) of targets with several
    # chunks: one slot per run, and a target's slots consecutive
    L = 512
    runs = {}
    for c in range(len(ch)):
        if sole[c]:
            continue
        cont = c >= L and x[c - L] == x[c]
        if cont:
            assert slot[c] == slot[c - L]
        else:
            assert slot[c] not in runs
            runs[slot[c]] = x[c]
    slots = np.array(sorted(runs))
    assert np.array_equal(slots, np.arange(len(slots)))
    tg = np.array([runs[k] for k in slots])
    assert np.all(np.diff(tg) >= 0)  # consecutive per target
    for g0 in range(0, len(ch), 64):
        assert np.all(cnt[g0:g0 + 64] == cnt[g0]) and np.all(first[g0:g0 + 64] == first[g0] + np.arange(len(ch[g0:g0 + 64])))
    got = np.zeros(nzx, complex)
    for c in range(len(ch)):
        r = first[c] + 64 * np.arange(cnt[c])
        got[x[c]] += np.sum(rw[r] * Z[ridx[r, 0]] * Z[ridx[r, 1]] * Z[ridx[r, 2]])
    for i in range(ny):  # a Y target's gradient goes back to its factors
        a, b = prod[i]
        for t in ((a, b), (b, a)):
            want[t[0]] += want[nz + 1 + i] * Z[t[1]]
            got[t[0]] += got[nz + 1 + i] * Z[t[1]]
    scale = np.abs(want[:nz]).max()
    assert np.abs(got[:nz] - want[:nz]).max() <= 1e-12 * scale
    if hname == "H2000":
        # LDS passes per 16-byte read that the record order gives the adjoint: per step of a
        # group of 64 chunks, per 16-lane group of the read (gfx950: in each half wave, the
        # quads 4q..4q+3 with an even number of set bits in q, and those with an odd number)
        # and factor slot, the largest number of distinct entries in one bank class x mod 16
        # (padding, index nz, is a broadcast and not counted)
        lane = np.arange(64)
        q = (lane >> 2) & 7
        grp = ((lane >> 5) << 1) | ((q ^ (q >> 1) ^ (q >> 2)) & 1)
        passes = reads = 0
        for g0 in range(0, len(ch), 64):
            nl = min(64, len(ch) - g0)
            for k in range(cnt[g0]):
                r = ridx[first[g0] + 64 * k + np.arange(nl), :3]
                for g in range(4):
                    sel = r[grp[:nl] == g]
                    for f in range(3):
                        e = np.unique(sel[:, f])
                        e = e[e != nz]
                        if len(e):
                            passes += np.bincount(e % 16).max()
                            reads += 1
        assert passes / reads < 1.35, passes / reads


def test_tuning_table_applied_at_init(tmp_path):
    """The tuning table compiled into the library (csrc/tuning_table.inc, written by
    mag-ace_amd/tune.py) sets the density variants of a listed header at init; an unlisted
    header keeps the built-in rule; magace_set_tuning validates its values."""
    import re
    from helpers import HEADERS
    inc = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mag-ace_amd", "csrc",
                       "tuning_table.inc")
    rows = {}
    for line in open(inc):
        m = re.match(r"\{\{([-\d, ]+)\}, \{(-?\d+), (\d+)\}, \{(-?\d+), (\d+)\}\},\s*//\s*(\w+)", line.strip())
        if m:
            rows[m.group(6)] = (tuple(int(v) for v in m.group(1).split(",")), (int(m.group(2)), int(m.group(3))),
                                (int(m.group(4)), int(m.group(5))))
    keys = ("nb_a", "nb_m", "nr2_m", "l2m", "nr3_a", "nr3_m", "l3a", "l3m", "nr4_a", "nr4_m", "l4a", "l4m")
    for name, (hdr, v64, v32) in rows.items():
        assert hdr == tuple(HEADERS[name][k] for k in keys), name
        write_potential(str(tmp_path / (name + ".in")), name, 3)
        capi.init_file(str(tmp_path / (name + ".in")))
        assert capi.get_tuning() == (v64, v32), name
    write_potential(str(tmp_path / "m0.in"), "H500M0", 3)  # not a tuned header
    capi.init_file(str(tmp_path / "m0.in"))
    assert capi.get_tuning() == ((-1, 0), (-1, 0))
    capi.set_tuning(fp64=(1, 2), mixed=(2, 3))
    assert capi.get_tuning() == ((1, 2), (2, 3))
    with pytest.raises(capi.MagaceError):
        capi.set_tuning(fp64=(3, 0))
