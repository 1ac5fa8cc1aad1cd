"""acelib_ on host arrays: the atoms go through the kernels in chunks of whole atoms, staged
through pinned buffers so that copies overlap the kernels (MAGACE_ACELIB_CHUNK_PAIRS sets the
chunk size and is read at every call).  Chunks cut at even atoms and the adjoint's walk
depends only on an atom's index in the whole system, so chunking changes no arithmetic: every
chunk size must agree with the device-pointer entry point on the whole list to 1e-12 (FP64),
including lists with atoms that have no neighbours and a list with no pairs at all.  The device entry point's own partitions are checked the same way:
evaluations of 2048+ atoms run as two halves on two streams, and the FP32 adjoint pairs atoms
two per workgroup (odd counts included).
"""
import os
import sys

import numpy as np
import pytest

from helpers import PKG, relerr

capi, potgen, systems, nbr = PKG.capi, PKG.potgen, PKG.systems, PKG.neighbors
torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no GPU visible", allow_module_level=True)

DEV = torch.device("cuda:0")
KEYS = ("eout", "flist", "tlist", "tlist0")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def lists(tmp_path_factory):
    pot = str(tmp_path_factory.mktemp("chunks") / "Fe.pbe.in")
    h = potgen.H500
    with open(pot, "w") as f:
        f.write(potgen.render(h, capi.enumerate_basis(h), 2031))
    s = systems.make_system(8, 2031, spins="uniform")  # 1024 atoms
    a = nbr.neilist(s["pos"], s["sdir"] * s["smag"][:, None], s["cell"], 5.0)
    return pot, a


def strip(a, drop):
    """The same list with every pair of the atoms in `drop` removed (empty pair ranges)."""
    drop = set(int(i) for i in drop)
    n = len(a["nstart"])
    keep = np.zeros(len(a["dlist"]), bool)
    nstart, nstop = np.zeros(n, np.int32), np.zeros(n, np.int32)
    nxt = 1
    for i in range(n):
        lo, hi = a["nstart"][i] - 1, a["nstop"][i]
        nstart[i] = nxt
        if i not in drop:
            keep[lo:hi] = True
            nxt += hi - lo
        nstop[i] = nxt - 1
    b = dict(a)
    b["nstart"], b["nstop"] = nstart, nstop
    for k in ("dlist", "dlist_mom", "nlist", "occlist"):
        b[k] = np.ascontiguousarray(a[k][keep])
    for k in ("elist", "elist_mom"):
        b[k] = np.ascontiguousarray(a[k].reshape(-1, 3)[keep].reshape(-1))
    return b


def device_entry(a):
    t = lambda x, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(x), dtype=dt, device=DEV)
    off = t(np.append(a["nstart"] - 1, a["nstop"][-1]), torch.int32)
    ins = [t(a[k]) for k in ("dlist", "elist", "dlist_momi", "elist_momi", "dlist_mom", "elist_mom")]
    N, P = len(a["nstart"]), len(a["dlist"])
    outs = [torch.zeros(n, dtype=torch.float64, device=DEV) for n in (N, 3 * P, 3 * P, 3 * N)]
    stream = torch.cuda.current_stream().cuda_stream
    capi.compute_device(N, P, off.data_ptr(), *[x.data_ptr() for x in ins], *[x.data_ptr() for x in outs],
                        stream=stream)
    capi.check_errors(stream)
    return {k: o.cpu().numpy() for k, o in zip(KEYS, outs)}


def host_entry(a, chunk):
    old = os.environ.get("MAGACE_ACELIB_CHUNK_PAIRS")
    if chunk is None:
        os.environ.pop("MAGACE_ACELIB_CHUNK_PAIRS", None)
    else:
        os.environ["MAGACE_ACELIB_CHUNK_PAIRS"] = str(chunk)
    try:
        return capi.acelib(a)
    finally:
        if old is None:
            os.environ.pop("MAGACE_ACELIB_CHUNK_PAIRS", None)
        else:
            os.environ["MAGACE_ACELIB_CHUNK_PAIRS"] = old


@pytest.mark.parametrize("chunk", [None, 1024, 5000, 20000])
def test_chunked_host_path_equals_device_entry(lists, chunk):
    pot, a = lists
    say("acelib_ chunks:", chunk, "pairs per chunk,", len(a["dlist"]), "pairs")
    capi.init_file(pot)
    capi.set_precision(capi.FP64)
    want = device_entry(a)
    got = host_entry(a, chunk)
    for k in KEYS:
        assert relerr(got[k], want[k]) < 1e-12, k


def test_chunks_with_empty_atoms(lists):
    pot, a = lists
    n = len(a["nstart"])
    # isolated atoms at the start, the end, single ones inside, and a run of 40 whose pairs
    # would fill more than one 1024-pair chunk
    drop = [0, 1, 77, 300, n - 1] + list(range(500, 540))
    b = strip(a, drop)
    say("acelib_ chunks with", len(drop), "empty atoms,", len(b["dlist"]), "pairs")
    capi.init_file(pot)
    capi.set_precision(capi.FP64)
    want = device_entry(b)
    for chunk in (None, 1024):
        got = host_entry(b, chunk)
        for k in KEYS:
            assert relerr(got[k], want[k]) < 1e-12, (chunk, k)
    # no pairs at all: only the one-body (B1) terms remain
    e = strip(a, range(n))
    assert len(e["dlist"]) == 0
    got = host_entry(e, 1024)
    want = device_entry(e)
    assert len(got["flist"]) == 0 and len(got["tlist"]) == 0
    assert relerr(got["eout"], want["eout"]) < 1e-12 and relerr(got["tlist0"], want["tlist0"]) < 1e-12


def test_mixed_precision_chunked(lists):
    pot, a = lists
    say("acelib_ chunks, mixed precision")
    capi.init_file(pot)
    capi.set_precision(capi.MIXED)
    try:
        want = device_entry(a)
        got = host_entry(a, 4096)
    finally:
        capi.set_precision(capi.FP64)
    for k in KEYS:
        assert relerr(got[k], want[k]) < 1e-6, k


def test_second_species_rejected_then_recovers(lists):
    """A pair of another species anywhere in occlist is refused (the single-species check of
    lammps/pair_spin_ace.cpp:453); the next valid call works."""
    pot, a = lists
    capi.init_file(pot)
    capi.set_precision(capi.FP64)
    for where in (0, len(a["occlist"]) // 2 + 7, len(a["occlist"]) - 1):
        b = dict(a)
        b["occlist"] = a["occlist"].copy()
        b["occlist"][where] = 2
        with pytest.raises(capi.MagaceError, match="multiple species"):
            capi.acelib(b)
    got = capi.acelib(a)
    want = device_entry(a)
    assert relerr(got["flist"], want["flist"]) < 1e-12


def test_split_halves_equal_one_sequence(tmp_path):
    """From 2048 atoms an evaluation runs as two halves of the atom range on two streams
    (MAGACE_SPLIT_STREAMS=0 keeps one sequence).  Only the order of the adjoint's LDS atomics
    may change: FP64 to 1e-12, mixed to 1e-6, and acelib_ on host arrays agrees as well."""
    pot = str(tmp_path / "Fe.pbe.in")
    h = potgen.H500
    with open(pot, "w") as f:
        f.write(potgen.render(h, capi.enumerate_basis(h), 2032))
    s = systems.make_system(12, 2032, spins="uniform")  # 3456 atoms
    a = nbr.neilist(s["pos"], s["sdir"] * s["smag"][:, None], s["cell"], 5.0)
    assert len(a["nstart"]) >= 2048
    capi.init_file(pot)
    old = os.environ.get("MAGACE_SPLIT_STREAMS")
    try:
        for prec, tol in ((capi.FP64, 1e-12), (capi.MIXED, 1e-6)):
            capi.set_precision(prec)
            say("split halves, precision", prec, len(a["nstart"]), "atoms")
            os.environ["MAGACE_SPLIT_STREAMS"] = "0"
            want = device_entry(a)
            os.environ["MAGACE_SPLIT_STREAMS"] = "1"
            got = device_entry(a)
            again = device_entry(a)
            host = capi.acelib(a)
            for k in KEYS:
                assert relerr(got[k], want[k]) < tol, (prec, k)
                assert relerr(again[k], want[k]) < tol, (prec, k)
                assert relerr(host[k], want[k]) < tol, (prec, k)
    finally:
        capi.set_precision(capi.FP64)
        if old is None:
            os.environ.pop("MAGACE_SPLIT_STREAMS", None)
        else:
            os.environ["MAGACE_SPLIT_STREAMS"] = old


def first_atoms(a, n):
    """The list of the first n atoms (their pairs only); each atom's results do not depend on
    the others."""
    b = dict(a)
    npair = int(a["nstop"][n - 1])
    b["nstart"], b["nstop"] = a["nstart"][:n].copy(), a["nstop"][:n].copy()
    for k in ("dlist", "dlist_mom", "nlist", "occlist"):
        b[k] = np.ascontiguousarray(a[k][:npair])
    for k in ("elist", "elist_mom"):
        b[k] = np.ascontiguousarray(a[k][:3 * npair])
    b["dlist_momi"] = np.ascontiguousarray(a["dlist_momi"][:n])
    b["elist_momi"] = np.ascontiguousarray(a["elist_momi"][:3 * n])
    return b


def compare_prefix(got, full, n, npair, tol, what):
    assert relerr(got["eout"], full["eout"][:n]) < tol, (what, "eout")
    assert relerr(got["tlist0"], full["tlist0"][:3 * n]) < tol, (what, "tlist0")
    assert relerr(got["flist"], full["flist"][:3 * npair]) < tol, (what, "flist")
    assert relerr(got["tlist"], full["tlist"][:3 * npair]) < tol, (what, "tlist")


def test_odd_atom_counts(lists):
    """The FP32 adjoint takes two atoms per workgroup; with an odd count the last atom is paired
    with itself and only its first copy is written.  1023 and 1 atoms against the same atoms in
    the 1024-atom evaluation (mixed to 1e-6; FP64 to 1e-12)."""
    pot, a = lists
    assert "dlist_momi" in a and len(a["dlist_momi"]) == len(a["nstart"])
    capi.init_file(pot)
    try:
        for prec, tol in ((capi.MIXED, 1e-6), (capi.FP64, 1e-12)):
            capi.set_precision(prec)
            full = device_entry(a)
            for n in (1023, 1):
                say("odd atom count", n, "precision", prec)
                b = first_atoms(a, n)
                compare_prefix(device_entry(b), full, n, len(b["dlist"]), tol, (prec, n))
    finally:
        capi.set_precision(capi.FP64)


def test_split_with_odd_half(tmp_path):
    """3455 atoms: halves of 1727 and 1728 atoms, the first odd, against the same atoms in one
    sequence."""
    pot = str(tmp_path / "Fe.pbe.in")
    h = potgen.H500
    with open(pot, "w") as f:
        f.write(potgen.render(h, capi.enumerate_basis(h), 2033))
    s = systems.make_system(12, 2033, spins="uniform")
    a = nbr.neilist(s["pos"], s["sdir"] * s["smag"][:, None], s["cell"], 5.0)
    n = len(a["nstart"]) - 1
    b = first_atoms(a, n)
    capi.init_file(pot)
    old = os.environ.get("MAGACE_SPLIT_STREAMS")
    try:
        for prec, tol in ((capi.MIXED, 1e-6), (capi.FP64, 1e-12)):
            capi.set_precision(prec)
            say("split with an odd half,", n, "atoms, precision", prec)
            os.environ["MAGACE_SPLIT_STREAMS"] = "0"
            full = device_entry(a)
            os.environ["MAGACE_SPLIT_STREAMS"] = "1"
            compare_prefix(device_entry(b), full, n, len(b["dlist"]), tol, prec)
    finally:
        capi.set_precision(capi.FP64)
        if old is None:
            os.environ.pop("MAGACE_SPLIT_STREAMS", None)
        else:
            os.environ["MAGACE_SPLIT_STREAMS"] = old

