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
        n = np.zeros(4, np.int32)
        lib().orc_counts(self.h, n)
        return [int(x) for x in n]

    def basis(self, order):
        n = self.counts()[order - 1]
        width = {1: 1, 2: 4, 3: 9, 4: 15}[order]
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
            raw = np.fromfile(fout, dtype="<f8")
        o = 0
        out = {}
        for k, n in (("eout", natom), ("flist", 3 * nei), ("tlist", 3 * nei), ("tlist0", 3 * natom),
                     ("seconds", nreps)):
            out[k] = raw[o:o + n].copy()
            o += n
        out["stdout"] = p.stdout
        return out
