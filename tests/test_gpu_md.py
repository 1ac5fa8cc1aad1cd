"""Device-resident spin-lattice dynamics (magace_md_*) on the GPU.

Checks: the assembled forces and magnetic gradients equal the caller-side assembly of the
evaluator's pair outputs (lammps/pair_spin_ace.cpp:571-641 conventions) on the same list;
energy conservation is second order in dt; total momentum is conserved; spin norms stay 1;
steps replayed from the captured HIP graph equal eager steps, replay after replay.
"""
import os
import sys

import numpy as np
import pytest

from helpers import PKG, relerr

capi, potgen, systems, md, nbr = PKG.capi, PKG.potgen, PKG.systems, PKG.md, PKG.neighbors
torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no GPU visible", allow_module_level=True)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def system(tmp_path_factory):
    """lmax-4 header with small coefficients (c ~ N(0, 1e-3)), 128 jittered bcc atoms, cone
    spins, 300 K velocities."""
    pot = str(tmp_path_factory.mktemp("md") / "Fe.pbe.in")
    h = dict(potgen.H500)
    with open(pot, "w") as f:
        f.write(potgen.render(h, capi.enumerate_basis(h), 2030, c_scale=1e-3))
    s = systems.make_system(4, 5, spins="cone")
    v = md.maxwell_velocities(len(s["pos"]), 300.0, np.random.default_rng(5))
    return pot, s, v


def test_md_forces_equal_assembled_evaluator_outputs(system):
    pot, s, v = system
    say("md: assembly vs evaluator outputs")
    capi.init_file(pot)
    capi.set_precision(capi.FP64)
    m = md.SpinLatticeMD(s["pos"], s["cell"], s["sdir"], s["smag"], vel=v, dt=1e-4, skin=0.5)
    st = m.state()
    m.close()
    a = nbr.neilist(s["pos"], s["sdir"] * s["smag"][:, None], s["cell"], 5.5, True)
    out = capi.acelib(a)
    e, f, fm, _ = systems.assemble(a, out["eout"], out["flist"], out["tlist"], out["tlist0"])
    assert abs(st["epot"] - e) <= 1e-10 * max(1.0, abs(e))
    assert relerr(st["f"], f) <= 1e-10
    # systems.assemble returns fm = -dE/dm (the LAMMPS sign); the MD state holds dE/dm
    assert relerr(-st["dEdm"], fm) <= 1e-10


def test_md_energy_conservation_second_order(system):
    pot, s, v = system
    say("md: energy drift vs dt")
    capi.init_file(pot)
    capi.set_precision(capi.FP64)
    drift = {}
    for dt in (4e-4, 2e-4, 1e-4):
        m = md.SpinLatticeMD(s["pos"], s["cell"], s["sdir"], s["smag"], vel=v, dt=dt)
        st = m.state()
        e0 = st["epot"] + st["ekin"]
        p0 = (st["v"]).sum(axis=0)
        n = int(round(0.02 / dt))
        worst = 0.0
        for _ in range(10):
            m.run(n // 10)
            st = m.state()
            worst = max(worst, abs(st["epot"] + st["ekin"] - e0) / len(s["pos"]))
        drift[dt] = worst
        assert np.abs(st["v"].sum(axis=0) - p0).max() < 1e-10  # momentum
        assert np.abs(np.linalg.norm(st["sdir"], axis=1) - 1.0).max() < 1e-13
        assert st["rebuilds"] >= n // 10
        m.close()
        say("  dt %.0e: max |dE|/N %.3e eV" % (dt, worst))
    assert drift[1e-4] < 5e-5
    # second order: halving dt divides the drift by ~4
    assert drift[4e-4] / drift[2e-4] > 3.0 and drift[2e-4] / drift[1e-4] > 3.0


def test_md_graph_replays_equal_eager_steps(system):
    pot, s, v = system
    say("md: graph vs eager")
    capi.init_file(pot)
    capi.set_precision(capi.FP64)
    a = md.SpinLatticeMD(s["pos"], s["cell"], s["sdir"], s["smag"], vel=v, dt=2e-4)
    b = md.SpinLatticeMD(s["pos"], s["cell"], s["sdir"], s["smag"], vel=v, dt=2e-4)
    for block in range(4):  # replay the captured 10-step graph; compare after every replay
        a.run(10, use_graph=True)
        b.run(10, use_graph=False)
        sa, sb = a.state(), b.state()
        assert np.abs(sa["x"] - sb["x"]).max() < 1e-12, block
        assert np.abs(sa["sdir"] - sb["sdir"]).max() < 1e-12, block
        assert abs(sa["epot"] - sb["epot"]) < 1e-10, block
    assert sa["captures"] == 1 and sb["captures"] == 0
    a.close()
    b.close()


def test_md_mixed_precision_runs_and_conserves(system):
    pot, s, v = system
    say("md: mixed precision")
    capi.init_file(pot)
    capi.set_precision(capi.MIXED)
    m = md.SpinLatticeMD(s["pos"], s["cell"], s["sdir"], s["smag"], vel=v, dt=1e-4)
    st = m.state()
    e0 = st["epot"] + st["ekin"]
    m.run(100)
    st = m.state()
    assert abs(st["epot"] + st["ekin"] - e0) / len(s["pos"]) < 1e-4
    m.close()
    capi.set_precision(capi.FP64)


def test_md_graph_recaptured_when_evaluator_state_changes(system):
    """A captured graph bakes in the evaluator's per-atom buffers, precision and coefficients.
    When another evaluation grows those buffers (freeing the old ones) or the precision
    changes, the next replay captures again instead of replaying stale launches; the replayed
    steps equal eager steps before and after."""
    pot, s, v = system
    say("md: graph after the evaluator state changed")
    capi.init_file(pot)
    capi.set_precision(capi.FP64)
    a = md.SpinLatticeMD(s["pos"], s["cell"], s["sdir"], s["smag"], vel=v, dt=2e-4)
    b = md.SpinLatticeMD(s["pos"], s["cell"], s["sdir"], s["smag"], vel=v, dt=2e-4)
    a.run(10, use_graph=True)
    b.run(10, use_graph=False)
    assert a.state()["captures"] == 1
    # a larger evaluation reallocates the evaluator's per-atom buffers
    big = systems.make_system(6, 8, spins="uniform")
    capi.acelib(nbr.neilist(big["pos"], big["sdir"] * big["smag"][:, None], big["cell"], 5.0))
    capi.set_precision(capi.MIXED)
    a.run(10, use_graph=True)
    b.run(10, use_graph=False)
    sa, sb = a.state(), b.state()
    assert sa["captures"] == 2, sa["captures"]
    # both ran the same mixed evaluations, which are bitwise repeatable
    assert np.array_equal(sa["x"], sb["x"])
    assert np.array_equal(sa["sdir"], sb["sdir"])
    assert sa["epot"] == sb["epot"]
    # unchanged state: the second graph is replayed, not captured again
    a.run(10, use_graph=True)
    assert a.state()["captures"] == 2
    a.close()
    b.close()
    capi.set_precision(capi.FP64)


def test_md_full_size_workload(tmp_path):
    """The 128000-atom spin-lattice workload: bcc Fe 40x40x40 cells, lmax-4 header
    "8 3 2 4 / 4 1 4 1 / 2 1 2 1" with c ~ N(0, 1e-3) (seed 2030), displacements N(0, 0.063 A)
    (1000 K), Maxwell velocities at 1000 K, uniform spins, dt 0.1 fs, list cutoff 5.5 A rebuilt
    every 10 steps on the device, 200 steps replayed from the captured graph, mixed precision.

    Step 0: the evaluator's pair outputs on the starting geometry (FP64) equal the oracle's for
    500 sampled central atoms (<= 1e-10 norm-wise); the mixed outputs are within 1e-4 of them;
    the MD's assembled forces, dE/dm, energy and device virial (mixed) are within the mixed bar
    of the caller-side assembly of the FP64 outputs.  After 200 steps: 20 rebuilds, one graph
    capture, momentum conserved, |s| = 1, |dE_total|/N < 2e-5 eV."""
    import oracle as O

    h = dict(potgen.H500)
    pot = str(tmp_path / "Fe_md.pbe.in")
    with open(pot, "w") as f:
        f.write(potgen.render(h, capi.enumerate_basis(h), 2030, c_scale=1e-3))
    rng = np.random.default_rng(2030)
    pos, cell = systems.bcc(40)
    pos = systems.jitter(pos, rng, 0.063)
    sdir, smag = systems.spins_uniform(len(pos), rng)
    vel = md.maxwell_velocities(len(pos), 1000.0, rng)
    n = len(pos)
    capi.init_file(pot)
    say("md full size: %d atoms, neighbour list (5.5 A) on the device" % n)
    a = nbr.neilist(pos, sdir * smag[:, None], cell, 5.5)
    say("  %d pairs; FP64 evaluation" % len(a["dlist"]))
    capi.set_precision(capi.FP64)
    o64 = capi.acelib(a)
    capi.set_precision(capi.MIXED)
    o32 = capi.acelib(a)
    # 500 sampled central atoms against the oracle
    sample = np.sort(np.random.default_rng(7).choice(n, 500, replace=False))
    off = a["offsets"]
    pidx = np.concatenate([np.arange(off[i], off[i + 1]) for i in sample])
    cnt = np.diff(off)[sample]
    sub_off = np.concatenate([[0], np.cumsum(cnt)]).astype(np.int32)
    sub = dict(nstart=(sub_off[:-1] + 1).astype(np.int32), nstop=sub_off[1:].astype(np.int32),
               dlist=a["dlist"][pidx], elist=a["elist"].reshape(-1, 3)[pidx].ravel(),
               dlist_mom=a["dlist_mom"][pidx], elist_mom=a["elist_mom"].reshape(-1, 3)[pidx].ravel(),
               dlist_momi=a["dlist_momi"][sample], elist_momi=a["elist_momi"].reshape(-1, 3)[sample].ravel())
    say("  oracle on %d sampled atoms (%d pairs)" % (len(sample), len(pidx)))
    ref = O.Oracle.load(pot).acelib(sub, nthreads=16)
    pick = dict(eout=lambda o: o["eout"][sample], flist=lambda o: o["flist"].reshape(-1, 3)[pidx].ravel(),
                tlist=lambda o: o["tlist"].reshape(-1, 3)[pidx].ravel(),
                tlist0=lambda o: o["tlist0"].reshape(-1, 3)[sample].ravel())
    for k, fn in pick.items():
        e64 = relerr(fn(o64), ref[k])
        e32 = np.abs(fn(o32) - ref[k]).max()
        say("  %-6s fp64 rel %.2e  mixed abs %.2e" % (k, e64, e32))
        assert e64 <= 1e-10, k
        assert e32 <= 1e-4, k
    e0, f0, fm0, vir0 = systems.assemble(a, o64["eout"], o64["flist"], o64["tlist"], o64["tlist0"])

    say("  MD create (mixed) and 200 graph steps")
    m = md.SpinLatticeMD(pos, cell, sdir, smag, vel=vel, dt=1e-4, rcut=5.0, skin=0.5, rebuild_every=10)
    st0 = m.state()
    df, dg = np.abs(st0["f"] - f0).max(), np.abs(-st0["dEdm"] - fm0).max()
    dv = np.abs(st0["virial"] - vir0).max() / np.abs(vir0).max()
    say("  step 0 vs FP64 assembly: max|dF| %.2e  max|d dE/dm| %.2e  |dE|/N %.2e  virial rel %.2e" %
        (df, dg, abs(st0["epot"] - e0) / n, dv))
    assert df <= 1e-4 and dg <= 1e-4
    assert abs(st0["epot"] - e0) / n <= 1e-6
    assert dv <= 1e-5
    p0 = st0["v"].sum(axis=0)
    etot0 = st0["epot"] + st0["ekin"]
    m.run(200, use_graph=True)
    st = m.state()
    m.close()
    capi.set_precision(capi.FP64)
    drift = abs(st["epot"] + st["ekin"] - etot0) / n
    say("  200 steps: rebuilds %d  captures %d  |dP| %.2e  |dE|/N %.2e eV  max||s|-1| %.2e" %
        (st["rebuilds"] - st0["rebuilds"], st["captures"], np.abs(st["v"].sum(axis=0) - p0).max(), drift,
         np.abs(np.linalg.norm(st["sdir"], axis=1) - 1.0).max()))
    assert st["steps"] == 200
    assert st["rebuilds"] - st0["rebuilds"] == 20
    assert st["captures"] == 1
    assert np.abs(st["v"].sum(axis=0) - p0).max() < 1e-8
    assert np.abs(np.linalg.norm(st["sdir"], axis=1) - 1.0).max() < 1e-12
    assert drift < 2e-5
    assert np.isfinite(st["virial"]).all()
