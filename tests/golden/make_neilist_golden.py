"""Write tests/golden/neilist_batches.npz from the reference neighbour list itself.

The reference neilist (neilist/neilist.F90) is compiled from its own source by
oracle/Makefile into oracle/_ref/neilist_ref (this container only) and run per structure;
the outputs are packed as packstructures packs them (src/neiutil.F90:5-264; restated in
oracle.packstructures because the reference's packstructures needs `cellvol`, which the
reference tree does not contain).  Two batches, one cutoff each:

  A (rc 5.0): bcc 2x2x2 jittered; triclinic 5-atom cell with atoms outside the cell and a
     zero moment; an open 40-atom cluster; one atom in a 2.5 A cube (neighbours exactly at
     rc are excluded); a 2-atom bcc cell strained +3%; bcc 2x2x2 strained -3% with a vacancy
  B (rc 3.7): bcc 2x2x2 jittered; a sheared cell; an open 12-atom cluster

Stored per batch: inputs (natoms, periodic, cells with rows = lattice vectors, pos, mag,
rcut) and the packed reference outputs with the reference's names, plus natombox.

Usage: python tests/golden/make_neilist_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import __graft_entry__  # noqa: E402
import oracle as O  # noqa: E402

systems = __graft_entry__.load_package().systems


def cluster(rng, n, half, dmin):
    pts = []
    while len(pts) < n:
        p = rng.uniform(-half, half, 3)
        if all(np.linalg.norm(p - q) > dmin for q in pts):
            pts.append(p)
    return np.array(pts)


def batch_a(rng):
    out = []
    pos, cell = systems.bcc(2)
    out.append(dict(pos=systems.jitter(pos, rng, 0.05), cell=cell, mag=rng.normal(size=(16, 3)) * 2.2))
    cell = np.array([[2.9, 0.0, 0.0], [0.7, 2.6, 0.0], [-0.4, 0.5, 3.1]])
    m = rng.normal(size=(5, 3))
    m[2] = 0.0
    out.append(dict(pos=rng.normal(size=(5, 3)) * 2.0, cell=cell, mag=m))
    out.append(dict(pos=cluster(rng, 40, 4.0, 1.2), cell=np.eye(3), mag=rng.normal(size=(40, 3)), periodic=False))
    out.append(dict(pos=np.array([[0.3, 0.3, 0.3]]), cell=np.eye(3) * 2.5, mag=np.array([[0.0, 0.0, 2.0]])))
    pos, cell = systems.bcc(1, 2.83 * 1.03)
    out.append(dict(pos=pos, cell=cell, mag=rng.normal(size=(2, 3))))
    pos, cell = systems.bcc(2, 2.83 * 0.97)
    pos = np.delete(systems.jitter(pos, rng, 0.05), 5, axis=0)
    out.append(dict(pos=pos, cell=cell, mag=rng.normal(size=(15, 3)) * 2.2))
    return out


def batch_b(rng):
    out = []
    pos, cell = systems.bcc(2)
    out.append(dict(pos=systems.jitter(pos, rng, 0.08), cell=cell, mag=rng.normal(size=(16, 3))))
    cell = np.array([[5.6, 0.0, 0.0], [1.9, 5.3, 0.0], [0.8, -1.1, 5.9]])
    f = rng.uniform(0, 1, size=(14, 3))
    out.append(dict(pos=f @ cell, cell=cell, mag=rng.normal(size=(14, 3))))
    out.append(dict(pos=cluster(rng, 12, 3.0, 1.5), cell=np.eye(3), mag=rng.normal(size=(12, 3)), periodic=False))
    return out


def pack(structs, rcut):
    p = O.packstructures(structs, rcut, builder=O.ref_neilist)
    c = O.packstructures(structs, rcut, builder=O.neilist)
    for k in p:
        assert np.array_equal(p[k], c[k]), "oracle restatement differs from the reference in " + k
    natombox = np.array([b - a + 1 for a, b in zip(p["natomboxpackstart"], p["natomboxpackstop"])], np.int32)
    inputs = dict(natoms=np.array([len(s["pos"]) for s in structs], np.int32),
                  periodic=np.array([int(s.get("periodic", True)) for s in structs], np.int32),
                  cells=np.stack([np.asarray(s["cell"], np.float64) for s in structs]),
                  pos=np.concatenate([np.asarray(s["pos"], np.float64) for s in structs]),
                  mag=np.concatenate([np.asarray(s["mag"], np.float64) for s in structs]),
                  rcut=np.float64(rcut))
    keep = ("n2startpack", "n2stoppack", "n2listpack", "d2listpack", "e2listpack", "d2listpack_mom",
            "e2listpack_mom", "d2listpack_momi", "e2listpack_momi", "strucaddress")
    return inputs, {k: p[k] for k in keep}, natombox


def main():
    rng = np.random.default_rng(2031)
    arrays = {}
    for name, structs, rcut in (("A", batch_a(rng), 5.0), ("B", batch_b(rng), 3.7)):
        inputs, ref, natombox = pack(structs, rcut)
        for k, v in inputs.items():
            arrays[name + "_in_" + k] = v
        for k, v in ref.items():
            arrays[name + "_ref_" + k] = v
        arrays[name + "_ref_natombox"] = natombox
        print(name, "structures", len(structs), "atoms", int(inputs["natoms"].sum()), "pairs", len(ref["d2listpack"]))
    np.savez_compressed(os.path.join(HERE, "neilist_batches.npz"), **arrays)


if __name__ == "__main__":
    main()
