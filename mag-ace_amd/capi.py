"""ctypes binding of libmagace.so (the C-ABI declared in include/magace.h).

There is no CPU fallback: if the shared library is missing, or no GPU is visible when a
compute entry point is called, an exception is raised.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libmagace.so")

FP64, MIXED, MIXED_ACC = 0, 1, 2

_D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_I = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
_IP = C.POINTER(C.c_int)
_DP = C.POINTER(C.c_double)

#: every symbol include/magace.h declares
EXPORTS = ("acelibini_", "acelib_", "magace_last_error", "magace_error_string", "magace_init_file",
           "magace_finalize", "magace_set_precision", "magace_get_precision", "magace_compute_device",
           "magace_check_errors", "magace_basis_counts", "magace_enumerate", "magace_model_info",
           "magace_pair_compute", "magace_model_terms", "magace_profile", "magace_profile_read",
           "magace_neighbors_device", "magace_packstructures", "magace_md_create", "magace_md_run",
           "magace_md_read", "magace_md_destroy", "magace_fit_ncoef", "magace_fit_get_coefficients",
           "magace_fit_set_coefficients", "magace_fit_forward", "magace_fit_backward", "magace_write_potential",
           "magace_model_records", "magace_compute_device_noalloc", "magace_set_tuning", "magace_get_tuning",
           "magace_tune", "magace_model_products")

ERR_CAPACITY = 4


class MagaceError(RuntimeError):
    pass


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise MagaceError("libmagace.so is not built (run __graft_entry__.build() or "
                          "python mag-ace_amd/build.py)")
    L = C.CDLL(LIBPATH, mode=C.RTLD_GLOBAL)
    L.magace_error_string.restype = C.c_char_p
    L.magace_init_file.argtypes = [C.c_char_p, C.c_char_p, C.c_double, C.c_double, C.c_double, C.c_int]
    L.magace_set_precision.argtypes = [C.c_int]
    L.acelibini_.argtypes = [_IP, _DP, _DP, _DP, _IP, _IP]
    L.acelibini_.restype = None
    L.acelib_.argtypes = [_IP, _IP, _IP, _I, _I, _I, _I, _I, _D, _D, _D, _D, _D, _D, _D, _D, _D, _D]
    L.acelib_.restype = None
    L.magace_compute_device.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 12
    L.magace_compute_device_noalloc.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 12
    L.magace_set_tuning.argtypes = [_I, C.c_int]
    L.magace_get_tuning.argtypes = [_I, C.c_int]
    L.magace_tune.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 11 + [C.c_int, _D, _I, C.c_void_p]
    L.magace_check_errors.argtypes = [C.c_void_p]
    L.magace_basis_counts.argtypes = [_I, C.c_int]
    L.magace_enumerate.argtypes = [_I, C.c_int, C.c_void_p, C.c_int]
    L.magace_model_info.argtypes = [np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS"), C.c_int]
    L.magace_neighbors_device.argtypes = [C.c_int, _I, _I, _D, C.c_double, C.c_void_p, C.c_void_p,
                                          C.c_longlong] + [C.c_void_p] * 9 + [_IP, C.c_void_p, C.c_void_p]
    L.magace_fit_get_coefficients.argtypes = [_D, C.c_int]
    L.magace_fit_set_coefficients.argtypes = [_D, C.c_int]
    L.magace_fit_forward.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 9 + [C.c_int, C.c_void_p, C.c_void_p,
                                                                              C.c_void_p]
    L.magace_fit_backward.argtypes = [C.c_int, C.c_int] + [C.c_void_p] * 12
    L.magace_write_potential.argtypes = [C.c_char_p]
    L.magace_md_create.argtypes = [C.c_int, _D, C.c_double, C.c_double, C.c_double, C.c_int, _D, _D, _D, _D,
                                   C.POINTER(C.c_void_p)]
    L.magace_md_run.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.magace_md_read.argtypes = [C.c_void_p] + [C.c_void_p] * 9 + [C.c_void_p]
    L.magace_md_destroy.argtypes = [C.c_void_p]
    L.magace_packstructures.argtypes = [C.c_int, _I, _I, _D, C.c_double, _D, _D, C.c_longlong,
                                        _I, _I, _I, _I, _D, _D, _D, _D, _D, _D, _I, _IP]
    _lib = L
    return L


def _check(rc=None):
    L = lib()
    code = L.magace_last_error()
    if (rc is not None and rc != 0) or code != 0:
        raise MagaceError(L.magace_error_string().decode() or ("error %s" % rc))


def init_file(path, element="Fe", rcut=5.0, rcut_magnetic=4.0, smear=0.2, xc=1):
    rc = lib().magace_init_file(os.fsencode(path), element.encode(), rcut, rcut_magnetic, smear, xc)
    _check(rc)


def acelibini(nele, rcut, rcut_magnetic, smear, xcnumber, elenumbers):
    """libace acelibini_ (reads <El>.<xc>.in from the current directory)."""
    L = lib()
    n, xc = C.c_int(nele), C.c_int(xcnumber)
    rc_, rm, sm = C.c_double(rcut), C.c_double(rcut_magnetic), C.c_double(smear)
    els = (C.c_int * len(elenumbers))(*elenumbers)
    L.acelibini_(C.byref(n), C.byref(rc_), C.byref(rm), C.byref(sm), C.byref(xc), els)
    _check()


def set_precision(mode):
    _check(lib().magace_set_precision(int(mode)))


def finalize():
    lib().magace_finalize()


def acelib(a: dict, out: dict | None = None):
    """Call acelib_ on host arrays in the pair_spin_ace layout (see systems.pack_acelib).

    `out`, if given, is a dict returned by an earlier call on a list of the same size; its
    arrays are written in place (a caller like LAMMPS keeps its output arrays between steps)."""
    L = lib()
    natom = int(len(a["nstart"]))
    nei = int(len(a["dlist"]))
    nall = int(a.get("nall", natom))
    if out is not None:
        eout, flist, tlist, tlist0 = out["_buffers"]
        if len(eout) != max(natom, 1) or len(flist) != max(3 * nei, 1):
            raise ValueError("out was made for a list of another size")
    else:
        eout = np.zeros(max(natom, 1))
        flist = np.zeros(max(3 * nei, 1))
        tlist = np.zeros(max(3 * nei, 1))
        tlist0 = np.zeros(max(3 * natom, 1))
    ci = lambda v: C.byref(C.c_int(v))
    pad_i = lambda x: np.ascontiguousarray(x if len(x) else np.zeros(1), dtype=np.int32)
    pad_d = lambda x: np.ascontiguousarray(x if len(x) else np.zeros(1), dtype=np.float64)
    L.acelib_(ci(natom), ci(nall), ci(nei), pad_i(a["nlist"]), pad_i(a["occ"]), pad_i(a["occlist"]),
              pad_i(a["nstart"]), pad_i(a["nstop"]), eout, pad_d(a["dlist"]), pad_d(a["elist"]), flist,
              pad_d(a["dlist_momi"]), pad_d(a["elist_momi"]), pad_d(a["dlist_mom"]), pad_d(a["elist_mom"]),
              tlist, tlist0)
    _check()
    return dict(eout=eout[:natom], flist=flist[:3 * nei], tlist=tlist[:3 * nei], tlist0=tlist0[:3 * natom],
                _buffers=(eout, flist, tlist, tlist0))


def compute_device(natom, nei, offsets, dlist, elist, dmomi, emomi, dmom, emom, eout, flist, tlist,
                   tlist0, stream=0):
    """acelib_ on device pointers (ints).  Enqueued on `stream` (int handle, 0 = library stream)."""
    rc = lib().magace_compute_device(natom, nei, offsets, dlist, elist, dmomi, emomi, dmom, emom, eout,
                                     flist, tlist, tlist0, stream or None)
    _check(rc)


def compute_device_noalloc(natom, nei, offsets, dlist, elist, dmomi, emomi, dmom, emom, eout, flist, tlist,
                           tlist0, stream=0):
    """compute_device with the per-atom intermediates in the library's fixed device scratch:
    no device memory is allocated (a size beyond the scratch raises MagaceError)."""
    rc = lib().magace_compute_device_noalloc(natom, nei, offsets, dlist, elist, dmomi, emomi, dmom, emom, eout,
                                             flist, tlist, tlist0, stream or None)
    _check(rc)


TUNE_CANDIDATES = ((0, 0), (1, 1), (1, 2), (1, 3), (1, 4), (2, 1), (2, 2), (2, 3), (2, 4))


def set_tuning(fp64=(-1, 0), mixed=(-1, 0)):
    """Density kernel variant and neighbour split per precision (see magace_set_tuning)."""
    _check(lib().magace_set_tuning(np.array([*fp64, *mixed], np.int32), 4))


def get_tuning():
    v = np.zeros(4, np.int32)
    _check(lib().magace_get_tuning(v, 4))
    return (int(v[0]), int(v[1])), (int(v[2]), int(v[3]))


def tune(natom, nei, offsets, dlist, elist, dmomi, emomi, dmom, emom, eout, flist, tlist, tlist0, reps=5,
         stream=0):
    """Times the density variants on device arrays for the current precision and keeps the
    fastest.  Returns (best (variant, split), {candidate: ms per evaluation or None})."""
    ms = np.zeros(len(TUNE_CANDIDATES))
    best = np.zeros(2, np.int32)
    rc = lib().magace_tune(natom, nei, offsets, dlist, elist, dmomi, emomi, dmom, emom, eout, flist, tlist, tlist0,
                           int(reps), ms, best, stream or None)
    _check(rc)
    return (int(best[0]), int(best[1])), {c: (float(t) if t >= 0 else None) for c, t in zip(TUNE_CANDIDATES, ms)}


def check_errors(stream=0):
    _check(lib().magace_check_errors(stream or None))


def basis_counts():
    """Numbers of B functions of orders 1..6 of the loaded potential."""
    n = np.zeros(6, np.int32)
    _check(lib().magace_basis_counts(n, 6))
    return [int(x) for x in n]


HDR_KEYS = ("nb_a", "nb_m", "nr2_m", "l2m", "nr3_a", "nr3_m", "l3a", "l3m", "nr4_a", "nr4_m", "l4a", "l4m",
            "nr5_a", "nr5_m", "l5a", "l5m", "nr6_a", "nr6_m", "l6a", "l6m")


def enumerate_basis(header: dict):
    """Index tuples of the basis of `header` (order -> int array), from the C++ enumeration."""
    L = lib()
    h = np.array([int(header.get(k, 0)) for k in HDR_KEYS], np.int32)
    out = {}
    for order, width in ((1, 1), (2, 4), (3, 9), (4, 15), (5, 21), (6, 27)):
        n = L.magace_enumerate(h, order, None, 0)
        if n < 0:
            _check(n)
        buf = np.zeros(max(n * width, 1), np.int32)
        L.magace_enumerate(h, order, buf.ctypes.data, buf.size)
        out[order] = buf[:n * width].reshape(n, width)
    return out


def model_info():
    v = np.zeros(13, np.int64)
    _check(lib().magace_model_info(v, 13))
    keys = ("nz", "terms", "raw", "raw2", "raw3", "raw4", "shape_class", "oA2", "contribs", "chunks", "raw5", "raw6",
            "products")
    return dict(zip(keys, (int(x) for x in v)))


def pair_compute(x, sp, ilist, neighbors, nlocal=None, newton_pair=1):
    """PairSpinACE::compute through the C++ caller (csrc/pair_host.cpp).

    x: (nall, 3); sp: (nall, 4) [unit direction, magnitude]; neighbors: list of int arrays
    (full list, indices into x).  Returns (energy, f (nall,3), fm (nall,3), virial[6], eatom).
    """
    L = lib()
    L.magace_pair_compute.argtypes = [C.c_int, C.c_int, _D, _D, _I, _I, C.c_int, _I, _I, _I, _I, C.c_int,
                                      _D, _D, _D, _D, _D]
    x = np.ascontiguousarray(x, np.float64)
    sp = np.ascontiguousarray(sp, np.float64)
    nall = len(x)
    nlocal = nall if nlocal is None else nlocal
    numneigh = np.array([len(n) for n in neighbors] + [0] * (nall - len(neighbors)), np.int32)
    offs = np.zeros(nall, np.int32)
    offs[1:] = np.cumsum(numneigh)[:-1]
    flat = np.ascontiguousarray(np.concatenate([np.asarray(n, np.int32) for n in neighbors])
                                if len(neighbors) else np.zeros(1, np.int32), np.int32)
    typ = np.ones(nall, np.int32)
    mp = np.array([-1, 0], np.int32)
    f = np.zeros((nall, 3))
    fm = np.zeros((nall, 3))
    e = np.zeros(1)
    vir = np.zeros(6)
    eatom = np.zeros(nall)
    rc = L.magace_pair_compute(nlocal, nall - nlocal, x, sp, typ, mp, len(ilist),
                               np.ascontiguousarray(ilist, np.int32), numneigh, offs, flat, int(newton_pair),
                               f, fm, e, vir, eatom)
    _check(rc)
    return float(e[0]), f, fm, vir, eatom


def model_terms():
    L = lib()
    L.magace_model_terms.argtypes = [C.c_void_p, C.c_void_p, C.c_int]
    n = L.magace_model_terms(None, None, 0)
    if n < 0:
        _check(n)
    idx = np.zeros((max(n, 1), 4), np.uint16)
    w = np.zeros(max(n, 1))
    L.magace_model_terms(idx.ctypes.data, w.ctypes.data, n)
    return idx[:n].astype(np.int64), w[:n]


def model_records():
    """The adjoint's gradient table: (chunks (nchunk, 2) int32, idx (nrec, 4), w (nrec,))."""
    L = lib()
    L.magace_model_records.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, _IP]
    nch = C.c_int(0)
    n = L.magace_model_records(None, 0, None, None, 0, C.byref(nch))
    if n < 0:
        _check(n)
    ch = np.zeros((max(nch.value, 1), 2), np.int32)
    idx = np.zeros((max(n, 1), 4), np.uint16)
    w = np.zeros(max(n, 1))
    L.magace_model_records(ch.ctypes.data, nch.value, idx.ctypes.data, w.ctypes.data, n, C.byref(nch))
    return ch[:nch.value], idx[:n].astype(np.int64), w[:n]


def model_products():
    """Products Y of the extended vector Z' (body orders 5, 6): (n, 2) factor pairs, entry
    nz + 1 + i = Z[pairs[i, 0]] Z[pairs[i, 1]]."""
    L = lib()
    L.magace_model_products.argtypes = [C.c_void_p, C.c_int]
    n = L.magace_model_products(None, 0)
    if n < 0:
        _check(n)
    p = np.zeros((max(n, 1), 2), np.int32)
    L.magace_model_products(p.ctypes.data, n)
    return p[:n].astype(np.int64)


def profile(enable=True):
    _check(lib().magace_profile(1 if enable else 0))


def profile_read():
    """(ms per kernel summed over evaluations [density, adjoint, forces], evaluations)."""
    L = lib()
    L.magace_profile_read.argtypes = [_D, C.POINTER(C.c_int)]
    ms = np.zeros(3)
    n = C.c_int(0)
    _check(L.magace_profile_read(ms, C.byref(n)))
    return ms, n.value


def _structs(natoms, periodic, cells):
    natoms = np.ascontiguousarray(natoms, dtype=np.int32).reshape(-1)
    periodic = np.ascontiguousarray(periodic, dtype=np.int32).reshape(-1)
    cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1)
    if not (len(periodic) == len(natoms) and len(cells) == 9 * len(natoms)):
        raise MagaceError("natoms, periodic and cells (9 per structure) disagree in length")
    return natoms, periodic, cells


def neighbors_device(natoms, periodic, cells, rcut, pos, mag, capacity, offsets, nlist, dlist, elist,
                     dlist_mom, elist_mom, dlist_momi, elist_momi, stream=0, shifts=0):
    """magace_neighbors_device: device pointers (ints) for pos/mag and the outputs; nlist = 0
    counts only.  cells: (nstruc, 3, 3) with ROWS = lattice vectors.  Returns (rc, npairs,
    natombox) where rc is 0 or ERR_CAPACITY; other errors raise."""
    natoms, periodic, cells = _structs(natoms, periodic, cells)
    npairs = C.c_int(0)
    nbox = np.zeros(max(1, len(natoms)), np.int32)
    rc = lib().magace_neighbors_device(len(natoms), natoms, periodic, cells, float(rcut), pos, mag,
                                       int(capacity), offsets, nlist or None, dlist, elist, dlist_mom,
                                       elist_mom, dlist_momi, elist_momi, shifts or None, C.byref(npairs),
                                       nbox.ctypes.data, stream or None)
    if rc not in (0, ERR_CAPACITY):
        _check(rc)
    return rc, npairs.value, nbox[:len(natoms)]


def packstructures(natoms, periodic, cells, rcut, pos, mag, capacity=None):
    """magace_packstructures on host arrays; returns the packed arrays with the reference's
    names (src/neiutil.F90:5-264).  Grows the capacity and retries once if needed."""
    natoms, periodic, cells = _structs(natoms, periodic, cells)
    pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1)
    mag = np.ascontiguousarray(mag, dtype=np.float64).reshape(-1)
    ntot = int(natoms.sum())
    if len(pos) != 3 * ntot or len(mag) != 3 * ntot:
        raise MagaceError("pos and mag need 3 values per atom")
    cap = int(capacity if capacity is not None else 96 * max(ntot, 1))
    S = len(natoms)
    for _ in range(2):
        o = dict(strucaddress=np.zeros(2 * max(S, 1), np.int32), n2startpack=np.zeros(max(ntot, 1), np.int32),
                 n2stoppack=np.zeros(max(ntot, 1), np.int32), n2listpack=np.zeros(max(cap, 1), np.int32),
                 d2listpack=np.zeros(max(cap, 1)), e2listpack=np.zeros(3 * max(cap, 1)),
                 d2listpack_mom=np.zeros(max(cap, 1)), e2listpack_mom=np.zeros(3 * max(cap, 1)),
                 d2listpack_momi=np.zeros(max(ntot, 1)), e2listpack_momi=np.zeros(3 * max(ntot, 1)),
                 natombox=np.zeros(max(S, 1), np.int32))
        npairs = C.c_int(0)
        rc = lib().magace_packstructures(S, natoms, periodic, cells, float(rcut), pos, mag, cap,
                                         o["strucaddress"], o["n2startpack"], o["n2stoppack"], o["n2listpack"],
                                         o["d2listpack"], o["e2listpack"], o["d2listpack_mom"],
                                         o["e2listpack_mom"], o["d2listpack_momi"], o["e2listpack_momi"],
                                         o["natombox"], C.byref(npairs))
        if rc == ERR_CAPACITY:
            cap = npairs.value
            continue
        _check(rc)
        P = npairs.value
        sizes = dict(strucaddress=2 * S, n2startpack=ntot, n2stoppack=ntot, n2listpack=P, d2listpack=P,
                     e2listpack=3 * P, d2listpack_mom=P, e2listpack_mom=3 * P, d2listpack_momi=ntot,
                     e2listpack_momi=3 * ntot, natombox=S)
        out = {k: v[:sizes[k]].copy() for k, v in o.items()}
        out["strucaddress"] = out["strucaddress"].reshape(S, 2)
        out["npairs"] = P
        return out
    raise MagaceError("packstructures: capacity retry failed")


def fit_ncoef():
    n = lib().magace_fit_ncoef()
    if n < 0:
        _check(-n)
    return n


def fit_get_coefficients():
    c = np.zeros(fit_ncoef())
    _check(lib().magace_fit_get_coefficients(c, len(c)))
    return c


def fit_set_coefficients(c):
    c = np.ascontiguousarray(c, dtype=np.float64)
    _check(lib().magace_fit_set_coefficients(c, len(c)))


def fit_forward(natom, nei, offsets, nlist, dlist, elist, dmomi, emomi, dmom, emom, atom_struct, nstruc,
                estruct, forces, stream=0):
    """magace_fit_forward on device pointers."""
    _check(lib().magace_fit_forward(natom, nei, offsets, nlist, dlist, elist, dmomi, emomi, dmom, emom,
                                    atom_struct or None, nstruc, estruct or None, forces, stream or None))


def fit_backward(natom, nei, offsets, nlist, dlist, elist, dmomi, dmom, emom, atom_struct, gE, gF, grad,
                 stream=0):
    """magace_fit_backward on device pointers (gE may be 0)."""
    _check(lib().magace_fit_backward(natom, nei, offsets, nlist, dlist, elist, dmomi, dmom, emom,
                                     atom_struct or None, gE or None, gF, grad, stream or None))


def write_potential(path):
    _check(lib().magace_write_potential(os.fsencode(path)))
