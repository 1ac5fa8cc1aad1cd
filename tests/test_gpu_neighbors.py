"""GPU neighbour lists (magace_packstructures / magace_neighbors_device) against the
reference neilist (golden batches written by the reference itself) and the oracle.

Bar: bit-exact.  Pair order, neighbour indices, ranges and every FP64 value (distances,
unit vectors, moments) must equal the reference's, including pairs that sit exactly at the
cutoff in exact arithmetic and that the reference's rounding keeps.
"""
import sys

import numpy as np
import pytest

import oracle as O
from helpers import PKG, load_case, relerr, write_potential
from test_neighbors_oracle import KEYS, golden_batches, structures

capi, nbr, systems = PKG.capi, PKG.neighbors, PKG.systems
torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no GPU visible", allow_module_level=True)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def assert_packed_equal(got, ref, what):
    for k in KEYS:
        g, r = np.asarray(got[k]), np.asarray(ref[k])
        assert g.shape == r.shape, (what, k, g.shape, r.shape)
        bad = np.flatnonzero(g.reshape(-1) != r.reshape(-1))  # FP64 compared exactly (+0 == -0)
        assert bad.size == 0, (what, k, bad[:5], g.reshape(-1)[bad[:5]], r.reshape(-1)[bad[:5]])


@pytest.mark.parametrize("batch", ["A", "B"])
def test_packstructures_matches_reference(batch):
    say("neighbours: golden batch", batch)
    inp, ref = golden_batches()[batch]
    got = capi.packstructures(inp["natoms"], inp["periodic"], inp["cells"], float(inp["rcut"]), inp["pos"],
                              inp["mag"])
    assert_packed_equal(got, ref, batch)
    assert np.array_equal(got["natombox"], ref["natombox"])


def test_config_batch_matches_oracle():
    """The batched-fit workload: 256 periodic configurations of 14-1023 atoms."""
    say("neighbours: 256-configuration batch")
    structs = systems.config_batch()
    ref = O.packstructures(structs, 5.0)
    got = nbr.packstructures(structs, 5.0)
    say("  atoms", sum(len(s["pos"]) for s in structs), "pairs", got["npairs"])
    assert_packed_equal(got, ref, "config_batch")
    assert np.array_equal(got["natombox"], ref["natomboxpackstop"] - ref["natomboxpackstart"] + 1)


def test_device_build_16k_matches_oracle():
    """The 16000-atom benchmark system, built on the device from device-resident inputs."""
    say("neighbours: 16000 atoms on the device")
    s = systems.make_system(20, 2028, spins="cone")
    mag = s["sdir"] * s["smag"][:, None]
    dev = torch.device("cuda:0")
    pos_t = torch.tensor(s["pos"], dtype=torch.float64, device=dev).contiguous()
    mag_t = torch.tensor(mag, dtype=torch.float64, device=dev).contiguous()
    dn = nbr.DeviceNeighbors([len(s["pos"])], [1], [s["cell"]], 5.0, capacity=1000)  # forces a regrow
    out = dn.build(pos_t, mag_t)
    torch.cuda.synchronize()
    ref = O.neilist(s["pos"], mag, s["cell"], 5.0, True)
    P = out["npairs"]
    assert P == ref["nei"]
    off = out["offsets"].cpu().numpy()
    assert np.array_equal(off[:-1] + 1, ref["nstart"]) and np.array_equal(off[1:], ref["nstop"])
    assert np.array_equal(out["nlist"].cpu().numpy() + 1, ref["origin"][ref["nlist"] - 1])
    for k_got, k_ref, sign in (("dlist", "dist", 1), ("elist", "eist", -1), ("dlist_mom", "dist_mom", 1),
                               ("elist_mom", "eist_mom", 1), ("dlist_momi", "dist_momi", 1),
                               ("elist_momi", "eist_momi", 1)):
        g = out[k_got].cpu().numpy()
        assert np.array_equal(g, sign * ref[k_ref]), k_got
    assert int(dn.natombox[0]) == ref["natombox"]


def test_device_lists_drive_the_evaluator(tmp_path):
    """Device-built lists feed magace_compute_device directly; the result equals acelib_ on
    the same lists from the host (FP64, 1e-12 relative) and the h500_128 golden case's
    reference outputs on its own (host-built, differently ordered) list (1e-10)."""
    say("neighbours: device lists -> evaluator (h500_128)")
    c = load_case("h500_128", tmp_path)
    z = c["z"]
    capi.init_file(c["pot"])
    capi.set_precision(capi.FP64)
    pos, cell = z["pos"], z["cell"]
    mag = z["sdir"] * z["smag"][:, None]
    host = nbr.neilist(pos, mag, cell, float(z["rlist"]), True)
    want = capi.acelib(host)
    dev = torch.device("cuda:0")
    dn = nbr.DeviceNeighbors([len(pos)], [1], [cell], float(z["rlist"]))
    L = dn.build(torch.tensor(pos, device=dev), torch.tensor(mag, device=dev))
    N, P = len(pos), L["npairs"]
    outs = [torch.zeros(N, dtype=torch.float64, device=dev), torch.zeros(3 * P, dtype=torch.float64, device=dev),
            torch.zeros(3 * P, dtype=torch.float64, device=dev), torch.zeros(3 * N, dtype=torch.float64, device=dev)]
    capi.compute_device(N, P, L["offsets"].data_ptr(), L["dlist"].data_ptr(), L["elist"].data_ptr(),
                        L["dlist_momi"].data_ptr(), L["elist_momi"].data_ptr(), L["dlist_mom"].data_ptr(),
                        L["elist_mom"].data_ptr(), *[o.data_ptr() for o in outs])
    capi.check_errors()
    got = {k: o.cpu().numpy() for k, o in zip(("eout", "flist", "tlist", "tlist0"), outs)}
    for k in got:
        assert relerr(got[k], want[k]) <= 1e-12, k
    # physical totals against the reference outputs on the case's own list
    e_ref, f_ref, m_ref, _ = systems.assemble(c["a"], c["ref"]["eout"], c["ref"]["flist"], c["ref"]["tlist"],
                                              c["ref"]["tlist0"])
    e_got, f_got, m_got, _ = systems.assemble(host, got["eout"], got["flist"], got["tlist"], got["tlist0"])
    assert abs(e_got - e_ref) <= 1e-10 * abs(e_ref)
    assert relerr(f_got, f_ref) <= 1e-10 and relerr(m_got, m_ref) <= 1e-10


def test_capacity_count_only_and_empty_structures():
    say("neighbours: capacity protocol, count-only, empty structures")
    rng = np.random.default_rng(3)
    pos1, cell1 = systems.bcc(2)
    structs = [dict(pos=pos1, mag=rng.normal(size=(16, 3)), cell=cell1),
               dict(pos=np.zeros((0, 3)), mag=np.zeros((0, 3)), cell=np.eye(3) * 3.0),
               dict(pos=pos1 + 0.1, mag=rng.normal(size=(16, 3)), cell=cell1, periodic=False)]
    ref = O.packstructures([s for s in structs if len(s["pos"])], 5.0)
    got = nbr.packstructures(structs, 5.0, capacity=10)  # too small: the mirror regrows once
    for k in ("d2listpack", "e2listpack", "n2listpack"):
        assert np.array_equal(got[k], ref[k]), k
    assert got["strucaddress"][1][0] == got["strucaddress"][2][0] == 17
    dev = torch.device("cuda:0")
    p = torch.tensor(np.concatenate([s["pos"] for s in structs]), device=dev)
    m = torch.tensor(np.concatenate([s["mag"] for s in structs]), device=dev)
    off = torch.zeros(33, dtype=torch.int32, device=dev)
    rc, npairs, nbox = capi.neighbors_device([16, 0, 16], [1, 1, 0], np.stack([s["cell"] for s in structs]), 5.0,
                                             p.data_ptr(), m.data_ptr(), 0, off.data_ptr(), 0, 0, 0, 0, 0, 0, 0)
    assert rc == 0 and npairs == got["npairs"] and nbox[1] == 0
    o = off.cpu().numpy()
    assert o[0] == 0 and o[-1] == npairs and np.all(np.diff(o) >= 0)
    assert o[16] - o[0] == got["strucaddress"][2][1] - 1


def test_reference_stops_become_errors():
    say("neighbours: error paths")
    with pytest.raises(capi.MagaceError, match="identical"):
        capi.packstructures([2], [1], [np.eye(3) * 6.0], 4.0, np.zeros(6), np.ones(6))
    with pytest.raises(capi.MagaceError, match="rcut"):
        capi.packstructures([1], [1], [np.eye(3) * 6.0], 150.0, np.zeros(3), np.ones(3))
    with pytest.raises(capi.MagaceError, match="Extreme"):
        capi.packstructures([1], [1], [np.diag([5.0, 5.0, 1e-4])], 4.0, np.zeros(3), np.ones(3))
    # the library stays usable after an error
    inp, ref = golden_batches()["B"]
    got = capi.packstructures(inp["natoms"], inp["periodic"], inp["cells"], float(inp["rcut"]), inp["pos"],
                              inp["mag"])
    assert_packed_equal(got, ref, "B after errors")
