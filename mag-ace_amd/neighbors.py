"""Neighbour lists on the GPU with the reference neilist's pair sets, order and values.

Host-side mirror of the reference's neighbour interface, computed by libmagace
(``magace_packstructures`` / ``magace_neighbors_device``):

* ``packstructures(structures, rcut)`` -- src/neiutil.F90:5-264: every structure's
  neighbour list (neilist/neilist.F90:148-796) packed one after another, returned with the
  reference's array names and conventions (1-based, eist = (x_j - x_i)/d).
* ``neilist(pos, mag, cell, rcut, periodic)`` -- one structure, returned as the acelib_
  arrays (what pair_spin_ace would pass, elist = (x_i - x_j)/r), ready for ``capi.acelib``.
* ``DeviceNeighbors`` -- device-resident rebuilds on torch tensors feeding
  ``magace_compute_device`` without a host round trip of the pair arrays.

Cells are (3, 3) arrays whose ROWS are the lattice vectors.  There is no CPU fallback.
"""
from __future__ import annotations

import numpy as np

from . import capi


def _cells(structures):
    return np.stack([np.asarray(s["cell"], dtype=np.float64).reshape(3, 3) for s in structures]) \
        if structures else np.zeros((0, 3, 3))


def packstructures(structures, rcut, capacity=None):
    """structures: list of dicts with ``pos`` (n, 3), ``mag`` (n, 3) moment vectors (muB),
    ``cell`` (3, 3), optional ``periodic`` (default True).  Returns n2startpack, n2stoppack,
    n2listpack, d2listpack, e2listpack, d2listpack_mom, e2listpack_mom, d2listpack_momi,
    e2listpack_momi (flat, reference order), strucaddress (nstruc, 2), natombox, npairs."""
    natoms = [len(np.asarray(s["pos"]).reshape(-1, 3)) for s in structures]
    periodic = [int(bool(s.get("periodic", True))) for s in structures]
    pos = np.concatenate([np.asarray(s["pos"], np.float64).reshape(-1) for s in structures]) \
        if structures else np.zeros(0)
    mag = np.concatenate([np.asarray(s["mag"], np.float64).reshape(-1) for s in structures]) \
        if structures else np.zeros(0)
    return capi.packstructures(natoms, periodic, _cells(structures), rcut, pos, mag, capacity)


def neilist(pos, mag, cell, rcut, periodic=True):
    """One structure -> acelib_ arrays (natom, nstart/nstop 1-based, nlist 1-based, occ,
    occlist, dlist, elist = (x_i - x_j)/r, dlist_mom, elist_mom, dlist_momi, elist_momi,
    offsets, pair_i, pair_j)."""
    pos = np.asarray(pos, np.float64).reshape(-1, 3)
    p = packstructures([dict(pos=pos, mag=mag, cell=cell, periodic=periodic)], rcut)
    n = len(pos)
    P = p["npairs"]
    offsets = np.zeros(n + 1, np.int32)
    offsets[:n] = p["n2startpack"] - 1
    offsets[n] = P
    pair_i = np.repeat(np.arange(n, dtype=np.int32), np.diff(offsets))
    return dict(natom=n, nall=n, nstart=p["n2startpack"], nstop=p["n2stoppack"], nlist=p["n2listpack"],
                occ=np.ones(n, np.int32), occlist=np.ones(P, np.int32), dlist=p["d2listpack"],
                elist=-p["e2listpack"], dlist_mom=p["d2listpack_mom"], elist_mom=p["e2listpack_mom"],
                dlist_momi=p["d2listpack_momi"], elist_momi=p["e2listpack_momi"], offsets=offsets,
                pair_i=pair_i, pair_j=(p["n2listpack"] - 1).astype(np.int32), natombox=int(p["natombox"][0]))


class DeviceNeighbors:
    """Rebuilds neighbour lists of a fixed set of structures on the device.

    ``build(pos, mag, cells=None)`` takes torch FP64 tensors (ntot, 3) on the GPU and returns
    a dict of device tensors in the acelib_ conventions: offsets (ntot+1, int32, global CSR),
    nlist (global 0-based neighbour atom), dlist, elist, dlist_mom, elist_mom, dlist_momi,
    elist_momi, plus npairs.  Capacity grows by 25% when exceeded (one extra pass)."""

    def __init__(self, natoms, periodic, cells, rcut, device="cuda:0", capacity=None):
        import torch
        self.torch = torch
        self.natoms = np.ascontiguousarray(natoms, dtype=np.int32).reshape(-1)
        self.periodic = np.ascontiguousarray(periodic, dtype=np.int32).reshape(-1)
        self.cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 3, 3)
        self.rcut = float(rcut)
        self.device = torch.device(device)
        self.ntot = int(self.natoms.sum())
        self.offsets = torch.zeros(self.ntot + 1, dtype=torch.int32, device=self.device)
        self.dlist_momi = torch.zeros(self.ntot, dtype=torch.float64, device=self.device)
        self.elist_momi = torch.zeros(3 * self.ntot, dtype=torch.float64, device=self.device)
        self.capacity = 0
        self._alloc(int(capacity) if capacity else 64 * max(self.ntot, 1))
        self.npairs = 0
        self.natombox = None

    def _alloc(self, cap):
        t, dev = self.torch, self.device
        self.capacity = cap
        self.nlist = t.zeros(cap, dtype=t.int32, device=dev)
        self.dlist = t.zeros(cap, dtype=t.float64, device=dev)
        self.elist = t.zeros(3 * cap, dtype=t.float64, device=dev)
        self.dlist_mom = t.zeros(cap, dtype=t.float64, device=dev)
        self.elist_mom = t.zeros(3 * cap, dtype=t.float64, device=dev)

    def build(self, pos, mag, cells=None, stream=None):
        t = self.torch
        if cells is not None:
            self.cells = np.ascontiguousarray(cells, dtype=np.float64).reshape(-1, 3, 3)
        if pos.dtype != t.float64 or mag.dtype != t.float64 or not pos.is_contiguous() or not mag.is_contiguous():
            raise capi.MagaceError("pos and mag must be contiguous float64 device tensors")
        if pos.numel() != 3 * self.ntot or mag.numel() != 3 * self.ntot:
            raise capi.MagaceError("pos/mag size does not match the structures")
        s = stream if stream is not None else t.cuda.current_stream(self.device).cuda_stream
        for _ in range(2):
            rc, npairs, nbox = capi.neighbors_device(
                self.natoms, self.periodic, self.cells, self.rcut, pos.data_ptr(), mag.data_ptr(), self.capacity,
                self.offsets.data_ptr(), self.nlist.data_ptr(), self.dlist.data_ptr(), self.elist.data_ptr(),
                self.dlist_mom.data_ptr(), self.elist_mom.data_ptr(), self.dlist_momi.data_ptr(),
                self.elist_momi.data_ptr(), s)
            if rc == 0:
                self.npairs, self.natombox = npairs, nbox
                P = npairs
                return dict(offsets=self.offsets, nlist=self.nlist[:P], dlist=self.dlist[:P],
                            elist=self.elist[:3 * P], dlist_mom=self.dlist_mom[:P], elist_mom=self.elist_mom[:3 * P],
                            dlist_momi=self.dlist_momi, elist_momi=self.elist_momi, npairs=P)
            self._alloc(npairs + npairs // 4 + 64)
        raise capi.MagaceError("neighbour capacity retry failed")
