"""Python access to the oracle (TEST INFRASTRUCTURE ONLY).

Two checkers live here, both FP64 and both CPU-only:

* ``RefDriver`` runs the reference libace itself, compiled from the reference Fortran
  sources by ``oracle/Makefile`` into ``oracle/_ref/ace_ref_driver``.  It is called the way
  LAMMPS pair_spin_ace calls it (lammps/pair_spin_ace.cpp:288 acelibini_, :545 acelib_).
* ``Oracle`` loads ``oracle/liboracle.so``, the C restatement of the same algorithm
  (oracle/magace_oracle.c), which also exposes building blocks (CG, Y_lm, radial basis,
  splines, basis enumeration) for closed-form tests.

Nothing in the product imports this module; only tests/, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of bench.py do.
"""
from __future__ import annotations

import ctypes as C
import os
import shutil
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_DRIVER = os.path.join(HERE, "_ref", "ace_ref_driver")
REF_NEILIST = os.path.join(HERE, "_ref", "neilist_ref")
LIBORACLE = os.path.join(HERE, "liboracle.so")

_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_I = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class OrcHeader(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "nb_a", "nb_m", "nr2_m", "l2m", "nr3_a", "nr3_m", "l3a", "l3m",
        "nr4_a", "nr4_m", "l4a", "l4m", "nr5_a", "nr5_m", "l5a", "l5m",
        "nr6_a", "nr6_m", "l6a", "l6m", "ndensity", "npot", "nfemb", "radtype_a", "radtype_m")] + \
        [(n, C.c_double) for n in ("lambda_a", "lambda_m", "cut_a", "cut_m", "dcut_a", "dcut_m",
                                    "ecut", "decut")]


def build(force: bool = False) -> None:
    """Compile the C restatement (and the reference driver when the reference tree exists)."""
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    if os.path.isdir(os.environ.get("MAGACE_REFERENCE", "/root/reference")):
        subprocess.run(["make", "-s", "-C", HERE, "ref",
                        "REF=" + os.environ.get("MAGACE_REFERENCE", "/root/reference")], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBORACLE):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle oracle`")
        L = C.CDLL(LIBORACLE)
        L.orc_error.restype = C.c_char_p
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.POINTER(OrcHeader)]
        L.orc_load.restype = C.c_void_p
        L.orc_load.argtypes = [C.c_char_p, C.c_char_p, C.c_double, C.c_double, C.c_double, C.c_int]
        L.orc_free.argtypes = [C.c_void_p]
        L.orc_counts.argtypes = [C.c_void_p, _I]
        L.orc_basis_index.argtypes = [C.c_void_p, C.c_int, C.c_int, _I]
        L.orc_acelib.argtypes = [C.c_void_p, C.c_int, C.c_int, _I, _I, _D, _D, _D, _D, _D, _D,
                                 _D, _D, _D, _D, C.c_int]
        L.orc_bvectors.argtypes = [C.c_void_p, C.c_int, C.c_int, _I, _I, _D, _D, _D, _D, _D, _D, _D]
        L.orc_cg.restype = C.c_double
        L.orc_cg.argtypes = [C.c_int] * 6
        L.orc_ylm.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, _D, _D, _D, _D]
        L.orc_radbase.argtypes = [C.c_int, C.c_double, C.c_int, C.c_double, C.c_double, C.c_double, _D, _D]
        L.orc_radcore.argtypes = [C.c_double] * 4 + [C.POINTER(C.c_double)] * 2
        L.orc_spline_eval.argtypes = [C.c_void_p, C.c_int, C.c_double, _D, _D]
        L.orc_cutoff.restype = C.c_double
        L.orc_cutoff.argtypes = [C.c_void_p, C.c_int]
        L.orc_neilist.argtypes = [C.c_int, C.c_int, C.c_double, _D, _D, _D, C.c_int, C.c_int, _I,
                                  _I, _I, _I, _I, _I, _D, _D, _D, _D, _D, _D, _D]
        _lib = L
    return _lib


def header_struct(h: dict) -> OrcHeader:
    s = OrcHeader()
    for k, v in h.items():
        if hasattr(s, k):
            setattr(s, k, v)
    return s


class Oracle:
    """C restatement of libace (oracle/magace_oracle.c)."""

    def __init__(self, handle):
        self.h = handle

    @classmethod
    def from_header(cls, header: dict) -> "Oracle":
        L = lib()
        s = header_struct(header)
        h = L.orc_create(C.byref(s))
        if not h:
            raise RuntimeError(L.orc_error().decode())
        return cls(h)

    @classmethod
    def load(cls, path, element="Fe", rcut=5.0, rcut_magnetic=4.0, smear=0.2, xc=1) -> "Oracle":
        L = lib()
        h = L.orc_load(os.fsencode(path), element.encode(), rcut, rcut_magnetic, smear, xc)
        if not h:
            raise RuntimeError(L.orc_error().decode())
        return cls(h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_free(self.h)
            self.h = None

    def counts(self):
        """Numbers of B functions of orders 1..6."""
        n = np.zeros(6, np.int32)
        lib().orc_counts(self.h, n)
        return [int(x) for x in n]

    def basis(self, order):
        n = self.counts()[order - 1]
        width = {1: 1, 2: 4, 3: 9, 4: 15, 5: 21, 6: 27}[order]
        out = np.zeros((n, width), np.int32)
        buf = np.zeros(width, np.int32)
        for k in range(n):
            lib().orc_basis_index(self.h, order, k, buf)
            out[k] = buf
        return out

    def acelib(self, a: dict, nthreads: int = 0):
        natom = len(a["nstart"])
        nei = len(a["dlist"])
        eout = np.zeros(natom)
        flist = np.zeros(3 * nei + 1)
        tlist = np.zeros(3 * nei + 1)
        tlist0 = np.zeros(3 * natom)
        rc = lib().orc_acelib(self.h, natom, nei, a["nstart"], a["nstop"], a["dlist"], a["elist"],
                              a["dlist_momi"], a["elist_momi"], a["dlist_mom"], a["elist_mom"],
                              eout, flist, tlist, tlist0, nthreads)
        if rc != 0:
            raise RuntimeError(lib().orc_error().decode())
        return dict(eout=eout, flist=flist[:3 * nei], tlist=tlist[:3 * nei], tlist0=tlist0)

    def bvectors(self, a: dict):
        natom = len(a["nstart"])
        nb = sum(self.counts())
        out = np.zeros(natom * nb)
        rc = lib().orc_bvectors(self.h, natom, len(a["dlist"]), a["nstart"], a["nstop"], a["dlist"],
                                a["elist"], a["dlist_momi"], a["elist_momi"], a["dlist_mom"],
                                a["elist_mom"], out)
        if rc != 0:
            raise RuntimeError(lib().orc_error().decode())
        return out.reshape(natom, nb)

    def spline(self, which, r):
        val = np.zeros(4096)
        dval = np.zeros(4096)
        n = lib().orc_spline_eval(self.h, which, float(r), val, dval)
        if n < 0:
            raise RuntimeError("spline lookup failed (%d)" % n)
        return val[:n].copy(), dval[:n].copy()


def cg(j1, j2, J, m1, m2, M):
    return lib().orc_cg(j1, j2, J, m1, m2, M)


def ylm(lmax, v):
    n = (lmax + 1) ** 2
    yr, yi = np.zeros(n), np.zeros(n)
    dr, di = np.zeros(3 * n), np.zeros(3 * n)
    lib().orc_ylm(lmax, float(v[0]), float(v[1]), float(v[2]), yr, yi, dr, di)
    return yr + 1j * yi, (dr + 1j * di).reshape(3, n)


def radbase(radtype, lam, nb, cut, dcut, r):
    g, dg = np.zeros(nb), np.zeros(nb)
    lib().orc_radbase(radtype, lam, nb, cut, dcut, float(r), g, dg)
    return g, dg


def radcore(r, pre, lam, cut):
    cr, dcr = C.c_double(), C.c_double()
    lib().orc_radcore(float(r), pre, lam, cut, C.byref(cr), C.byref(dcr))
    return cr.value, dcr.value


# ----------------------------------------------------------------- reference driver
class RefDriver:
    """Runs the reference libace (oracle/_ref/ace_ref_driver) on acelib_ arrays."""

    MAGIC = 0x4D414345

    def __init__(self, potential_path, rcut=5.0, rcut_magnetic=4.0, smear=0.2, xc=1,
                 elenumber=27, element="Fe"):
        if not os.path.exists(REF_DRIVER):
            raise RuntimeError("reference driver not built (oracle/_ref/ace_ref_driver)")
        self.pot = os.path.abspath(potential_path)
        self.rcut, self.rcut_m, self.smear, self.xc = rcut, rcut_magnetic, smear, xc
        self.elenumber, self.element = elenumber, element

    def run(self, a: dict, nreps: int = 1, timeout: float = 3600.0):
        natom = len(a["nstart"])
        nei = len(a["dlist"])
        nall = int(a.get("nall", natom))
        with tempfile.TemporaryDirectory(prefix="magace_ref_") as td:
            xcname = {1: "pbe", 2: "lda"}[self.xc]
            shutil.copy(self.pot, os.path.join(td, "%s.%s.in" % (self.element, xcname)))
            fin = os.path.join(td, "in.bin")
            fout = os.path.join(td, "out.bin")
            with open(fin, "wb") as f:
                f.write(struct.pack("<7i", self.MAGIC, natom, nall, nei, nreps, self.xc, self.elenumber))
                f.write(struct.pack("<3d", self.rcut, self.rcut_m, self.smear))
                for k in ("nstart", "nstop", "nlist", "occ", "occlist"):
                    f.write(np.ascontiguousarray(a[k], dtype="<i4").tobytes())
                for k in ("dlist", "elist", "dlist_momi", "elist_momi", "dlist_mom", "elist_mom"):
                    f.write(np.ascontiguousarray(a[k], dtype="<f8").tobytes())
            p = subprocess.run([REF_DRIVER, td, fin, fout], capture_output=True, text=True,
                               timeout=timeout)
            if p.returncode != 0 or not os.path.exists(fout):
                raise RuntimeError("reference driver failed (%d): %s %s" %
                                   (p.returncode, p.stdout[-2000:], p.stderr[-2000:]))
            self.stdout = p.stdout  # the reference's own messages (basis counts)
            raw = np.fromfile(fout, dtype="<f8")
        o = 0
        out = {}
        for k, n in (("eout", natom), ("flist", 3 * nei), ("tlist", 3 * nei), ("tlist0", 3 * natom),
                     ("seconds", nreps)):
            out[k] = raw[o:o + n].copy()
            o += n
        out["stdout"] = p.stdout
        return out


_NL_INT = ("nstart", "nstop", "nlist", "origin", "nshift")


def _nl_unpack(natom, nbox, nei, ints, dbls):
    out = {"natom": natom, "natombox": nbox, "nei": nei}
    o = 0
    for k, n in (("nstart", natom), ("nstop", natom), ("nlist", nei), ("origin", nbox),
                 ("nshift", 3 * nbox)):
        out[k] = ints[o:o + n].copy()
        o += n
    o = 0
    for k, n in (("dist", nei), ("eist", 3 * nei), ("dist_mom", nei), ("eist_mom", 3 * nei),
                 ("dist_momi", natom), ("eist_momi", 3 * natom), ("pos", 3 * nbox)):
        out[k] = dbls[o:o + n].copy()
        o += n
    return out


def neilist(pos, mag, cell, rcut, periodic=True):
    """C restatement of the reference neighbour list (neilist/neilist.F90:148-796).

    pos, mag: (natom, 3); cell: (3, 3) with ROWS = lattice vectors.  Returns the reference's
    module arrays (1-based indices, eist = (x_j - x_i)/d) as a dict of flat arrays."""
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
    mag = np.ascontiguousarray(mag, dtype=np.float64).reshape(-1)
    natom = len(pos) // 3
    cellc = np.ascontiguousarray(np.asarray(cell, dtype=np.float64).reshape(3, 3)).reshape(-1)
    L = lib()
    capb, capp = max(64, 64 * natom), max(64, 80 * natom)
    for _ in range(3):
        counts = np.zeros(2, np.int32)
        I = {k: np.zeros(max(1, n), np.int32) for k, n in (("nstart", natom), ("nstop", natom),
             ("nlist", capp), ("origin", capb), ("nshift", 3 * capb))}
        D = {k: np.zeros(max(1, n)) for k, n in (("dist", capp), ("eist", 3 * capp), ("dist_mom", capp),
             ("eist_mom", 3 * capp), ("dist_momi", natom), ("eist_momi", 3 * natom), ("pos", 3 * capb))}
        rc = L.orc_neilist(natom, int(bool(periodic)), float(rcut), cellc, pos, mag, capb, capp, counts,
                           I["nstart"], I["nstop"], I["nlist"], I["origin"], I["nshift"], D["dist"],
                           D["eist"], D["dist_mom"], D["eist_mom"], D["dist_momi"], D["eist_momi"], D["pos"])
        if rc < 0:
            raise ValueError("orc_neilist: input rejected (%d)" % rc)
        if rc == 0:
            break
        capb, capp = int(counts[0]), int(counts[1])
    nbox, nei = int(counts[0]), int(counts[1])
    out = {"natom": natom, "natombox": nbox, "nei": nei}
    for k, n in (("nstart", natom), ("nstop", natom), ("nlist", nei), ("origin", nbox), ("nshift", 3 * nbox)):
        out[k] = I[k][:n].copy()
    for k, n in (("dist", nei), ("eist", 3 * nei), ("dist_mom", nei), ("eist_mom", 3 * nei),
                 ("dist_momi", natom), ("eist_momi", 3 * natom), ("pos", 3 * nbox)):
        out[k] = D[k][:n].copy()
    return out


def ref_neilist(pos, mag, cell, rcut, periodic=True, timeout=600.0):
    """Runs the reference neilist itself (oracle/_ref/neilist_ref, built from
    neilist/neilist.F90 by oracle/Makefile).  Same inputs and outputs as neilist()."""
    if not os.path.exists(REF_NEILIST):
        raise RuntimeError("reference neighbour list not built (oracle/_ref/neilist_ref)")
    pos = np.ascontiguousarray(pos, dtype="<f8").reshape(-1)
    mag = np.ascontiguousarray(mag, dtype="<f8").reshape(-1)
    natom = len(pos) // 3
    cellc = np.ascontiguousarray(np.asarray(cell, dtype="<f8").reshape(3, 3)).reshape(-1)
    with tempfile.TemporaryDirectory(prefix="magace_nl_") as td:
        fin, fout = os.path.join(td, "in.bin"), os.path.join(td, "out.bin")
        with open(fin, "wb") as f:
            f.write(struct.pack("<2id", natom, int(bool(periodic)), float(rcut)))
            f.write(cellc.tobytes() + pos.tobytes() + mag.tobytes())
        p = subprocess.run([REF_NEILIST, fin, fout], capture_output=True, text=True, timeout=timeout)
        if p.returncode != 0 or not os.path.exists(fout):
            raise RuntimeError("reference neilist failed (%d): %s %s" % (p.returncode, p.stdout[-2000:],
                                                                         p.stderr[-2000:]))
        raw = open(fout, "rb").read()
    natom, nbox, nei = struct.unpack("<3i", raw[:12])
    ni = 2 * natom + nei + 4 * nbox
    ints = np.frombuffer(raw[12:12 + 4 * ni], dtype="<i4")
    dbls = np.frombuffer(raw[12 + 4 * ni:], dtype="<f8")
    return _nl_unpack(natom, nbox, nei, ints, dbls)


def packstructures(structures, rcut, builder=neilist):
    """Restatement of packstructures (src/neiutil.F90:5-264): runs the neighbour list per
    structure and concatenates.  structures: list of dicts with pos (n,3), mag (n,3), cell
    (3,3, rows = lattice vectors), periodic.  Returns the packed arrays with the reference's
    names: n2startpack/n2stoppack (per-structure local, 1-based), n2listpack = origin(nlist)
    (local, 1-based), d2listpack, e2listpack, d2listpack_mom, e2listpack_mom,
    d2listpack_momi, e2listpack_momi, strucaddress (nstruc, 2) 1-based atom and pair starts,
    originpack, natomboxpackstart/stop.  (The reference's strucaddress has 4 columns of which
    it fills 2, neiutil.F90:73,146,178.)"""
    keys = ("n2startpack", "n2stoppack", "n2listpack", "d2listpack", "e2listpack", "d2listpack_mom",
            "e2listpack_mom", "d2listpack_momi", "e2listpack_momi", "originpack")
    acc = {k: [] for k in keys}
    strucaddress, boxstart, boxstop = [], [], []
    natompack = nei2pack = natomboxpack = 0
    for s in structures:
        nl = builder(s["pos"], s["mag"], s["cell"], rcut, s.get("periodic", True))
        natom, nei = nl["natom"], nl["nei"]
        boxstart.append(natomboxpack + 1)
        boxstop.append(natomboxpack + nl["natombox"])
        acc["originpack"].append(nl["origin"])
        acc["d2listpack_momi"].append(nl["dist_momi"])
        acc["e2listpack_momi"].append(nl["eist_momi"])
        strucaddress.append((natompack + 1, nei2pack + 1))
        acc["n2startpack"].append(nl["nstart"])
        acc["n2stoppack"].append(nl["nstop"])
        acc["n2listpack"].append(nl["origin"][nl["nlist"] - 1] if nei else np.zeros(0, np.int32))
        for k, src in (("d2listpack", "dist"), ("e2listpack", "eist"), ("d2listpack_mom", "dist_mom"),
                       ("e2listpack_mom", "eist_mom")):
            acc[k].append(nl[src])
        natompack += natom
        nei2pack += nei
        natomboxpack += nl["natombox"]
    out = {k: (np.concatenate(v) if v else np.zeros(0)) for k, v in acc.items()}
    out["strucaddress"] = np.array(strucaddress, np.int32).reshape(-1, 2)
    out["natomboxpackstart"] = np.array(boxstart, np.int32)
    out["natomboxpackstop"] = np.array(boxstop, np.int32)
    return out
