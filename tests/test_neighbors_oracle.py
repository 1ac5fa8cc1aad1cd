"""CPU checks of the neighbour-list oracle (oracle/neilist_oracle.c, oracle.packstructures).

Pins: the golden batches written by the reference neilist itself (bit-exact), a live run of
the reference build on further geometries when it is present, and a brute-force
enumeration of periodic images that fixes the pair SET (count and distances per atom)
independently of the restatement.
"""
import os

import numpy as np
import pytest

import oracle as O
from helpers import GOLDEN

KEYS = ("n2startpack", "n2stoppack", "n2listpack", "d2listpack", "e2listpack", "d2listpack_mom",
        "e2listpack_mom", "d2listpack_momi", "e2listpack_momi", "strucaddress")


def golden_batches():
    z = np.load(os.path.join(GOLDEN, "neilist_batches.npz"))
    out = {}
    for name in ("A", "B"):
        inp = {k[len(name) + 4:]: z[k] for k in z.files if k.startswith(name + "_in_")}
        ref = {k[len(name) + 5:]: z[k] for k in z.files if k.startswith(name + "_ref_")}
        out[name] = (inp, ref)
    return out


def structures(inp):
    res, a0 = [], 0
    for n, per, cell in zip(inp["natoms"], inp["periodic"], inp["cells"]):
        res.append(dict(pos=inp["pos"][a0:a0 + n], mag=inp["mag"][a0:a0 + n], cell=cell, periodic=bool(per)))
        a0 += n
    return res


@pytest.mark.parametrize("batch", ["A", "B"])
def test_oracle_matches_reference_golden(batch):
    inp, ref = golden_batches()[batch]
    got = O.packstructures(structures(inp), float(inp["rcut"]))
    for k in KEYS:
        assert np.array_equal(got[k], ref[k]), k
    nbox = got["natomboxpackstop"] - got["natomboxpackstart"] + 1
    assert np.array_equal(nbox, ref["natombox"])


def brute_force(pos, cell, rcut, periodic, tie=1e-9):
    """Per atom: sorted distances to every image x_j + n.cell (n integer, j != i or n != 0)
    with |d| < rcut + tie, from raw (unwrapped) positions."""
    pos = np.asarray(pos, np.float64)
    cell = np.asarray(cell, np.float64)
    if periodic:
        # enough images: rcut over the cell thickness along each reciprocal direction, plus
        # the spread of the (unwrapped) fractional coordinates
        rec = np.linalg.inv(cell).T
        thick = 1.0 / np.linalg.norm(rec, axis=1)
        frac = pos @ np.linalg.inv(cell)
        span = frac.max(axis=0) - frac.min(axis=0)
        reps = np.ceil(rcut / thick + span).astype(int) + 1
        g = [np.arange(-r, r + 1) for r in reps]
        shifts = np.stack(np.meshgrid(*g, indexing="ij"), -1).reshape(-1, 3) @ cell
    else:
        shifts = np.zeros((1, 3))
    out = []
    for i in range(len(pos)):
        d = (pos[None, :, :] + shifts[:, None, :] - pos[i][None, None, :]).reshape(-1, 3)
        r = np.linalg.norm(d, axis=1)
        keep = (r < rcut + tie) & (r > 1e-9)
        out.append(np.sort(r[keep]))
    return out


@pytest.mark.parametrize("batch", ["A", "B"])
def test_pair_set_matches_brute_force(batch):
    inp, ref = golden_batches()[batch]
    rcut = float(inp["rcut"])
    for s, st in enumerate(structures(inp)):
        nl = O.neilist(st["pos"], st["mag"], st["cell"], rcut, st["periodic"])
        bf = brute_force(st["pos"], st["cell"], rcut, st["periodic"])
        for i in range(len(st["pos"])):
            d = np.sort(nl["dist"][nl["nstart"][i] - 1:nl["nstop"][i]])
            # away from the cutoff the sets agree exactly; pairs within 1e-9 of rc (exact
            # ties in exact arithmetic) may go either way with the reference's rounding of
            # wrapped coordinates, but must be real ties
            inner, binner = d[d < rcut - 1e-9], bf[i][bf[i] < rcut - 1e-9]
            assert len(inner) == len(binner), (batch, s, i)
            np.testing.assert_allclose(inner, binner, rtol=0, atol=1e-12)
            assert np.sum(d >= rcut - 1e-9) <= np.sum(bf[i] >= rcut - 1e-9), (batch, s, i)


def test_pair_set_symmetric_and_unit_vectors():
    rng = np.random.default_rng(5)
    cell = np.array([[4.1, 0.2, 0.0], [0.3, 3.9, 0.1], [0.0, -0.4, 4.4]])
    pos = rng.uniform(0, 1, (9, 3)) @ cell + rng.normal(size=(9, 3))
    mag = rng.normal(size=(9, 3))
    nl = O.neilist(pos, mag, cell, 5.2, True)
    # every i -> j pair has a j -> i partner (same distance): counts per unordered origin pair
    i = np.repeat(np.arange(9), nl["nstop"] - nl["nstart"] + 1)
    j = nl["origin"][nl["nlist"] - 1] - 1
    c = np.zeros((9, 9), int)
    np.add.at(c, (i, j), 1)
    assert np.array_equal(c, c.T)
    e = nl["eist"].reshape(-1, 3)
    np.testing.assert_allclose(np.linalg.norm(e, axis=1), 1.0, atol=1e-14)
    # eist points from i to the image of j: x_image - x_i = d * eist
    pb = nl["pos"].reshape(-1, 3)
    np.testing.assert_allclose(pb[nl["nlist"] - 1] - pb[i], e * nl["dist"][:, None], atol=1e-12)


@pytest.mark.skipif(not os.path.exists(O.REF_NEILIST), reason="reference neilist not built here")
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_oracle_matches_live_reference(seed):
    rng = np.random.default_rng(100 + seed)
    cell = np.eye(3) * rng.uniform(2.5, 7.0) + rng.normal(scale=0.6, size=(3, 3))
    n = int(rng.integers(1, 30))
    pos = rng.normal(scale=3.0, size=(n, 3))
    mag = rng.normal(size=(n, 3))
    mag[0] = 0.0
    rcut = float(rng.uniform(2.0, 6.0))
    for periodic in (True, False):
        a = O.neilist(pos, mag, cell, rcut, periodic)
        b = O.ref_neilist(pos, mag, cell, rcut, periodic)
        for k in a:
            if isinstance(a[k], np.ndarray):
                assert np.array_equal(a[k], b[k]), (k, periodic)
            else:
                assert a[k] == b[k], (k, periodic)


def test_oracle_rejects_reference_stops():
    with pytest.raises(ValueError):
        O.neilist(np.zeros((2, 3)), np.ones((2, 3)), np.eye(3) * 5, 4.0, True)  # identical positions
    with pytest.raises(ValueError):
        O.neilist(np.zeros((1, 3)), np.ones((1, 3)), np.eye(3) * 5, 200.0, True)  # rcut too large
