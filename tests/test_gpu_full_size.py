"""Checks at the benchmark's full size (16000 bcc Fe atoms, lmax-6 header, 2043 B
functions) through properties that do not depend on an oracle run at that size:

* Newton's third law: the assembled forces sum to zero;
* central differences of the total energy for displaced atoms and rotated moments equal the
  assembled force and dE/dm (FP64);
* rotating positions and moments together leaves every per-atom energy unchanged;
* reversing every moment (time reversal) leaves every per-atom energy unchanged;
* the mixed path stays within 1e-4 eV/A and 1e-4 eV/muB of the FP64 path.
"""
import sys

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from helpers import PKG

capi, potgen, systems, nbr = PKG.capi, PKG.potgen, PKG.systems, PKG.neighbors
torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no GPU visible", allow_module_level=True)

DEV = torch.device("cuda:0")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def setup(tmp_path_factory):
    pot = str(tmp_path_factory.mktemp("full") / "Fe.pbe.in")
    h = potgen.H2000
    with open(pot, "w") as f:
        f.write(potgen.render(h, capi.enumerate_basis(h), 2028))
    s = systems.make_system(20, 2028, spins="cone")
    return pot, s


def evaluate(s, pos=None, sdir=None, smag=None, cell=None):
    """Device neighbour build + evaluation of the whole system; returns per-atom energies,
    assembled forces and dE/dm (host arrays)."""
    pos = s["pos"] if pos is None else pos
    sdir = s["sdir"] if sdir is None else sdir
    smag = s["smag"] if smag is None else smag
    cell = s["cell"] if cell is None else cell
    a = nbr.neilist(pos, sdir * smag[:, None], cell, 5.0)
    out = capi.acelib(a)
    e, f, fm, _ = systems.assemble(a, out["eout"], out["flist"], out["tlist"], out["tlist0"])
    return out["eout"], f, -fm


def test_forces_sum_to_zero_and_match_central_differences(setup):
    pot, s = setup
    say("full size: Newton's third law and central differences (16000 atoms)")
    capi.init_file(pot)
    capi.set_precision(capi.FP64)
    e0, f0, g0 = evaluate(s)
    assert np.abs(f0.sum(axis=0)).max() < 1e-9 * np.abs(f0).max() * np.sqrt(len(f0))
    rng = np.random.default_rng(11)
    h = 1e-4
    for i in rng.choice(len(s["pos"]), 3, replace=False):
        for k in range(3):
            p = s["pos"].copy()
            p[i, k] += h
            ep, _, _ = evaluate(s, pos=p)
            p[i, k] -= 2 * h
            em, _, _ = evaluate(s, pos=p)
            fd = -(ep - em).sum() / (2 * h)
            say("  atom %d dir %d: force %.10f  central difference %.10f" % (i, k, f0[i, k], fd))
            assert abs(fd - f0[i, k]) < 1e-6 * max(1.0, abs(f0[i, k]))
    # dE/dm along a small rotation of one moment and a change of its magnitude
    for i in rng.choice(len(s["pos"]), 2, replace=False):
        axis = np.cross(s["sdir"][i], rng.normal(size=3))
        axis /= np.linalg.norm(axis)
        th = 1e-4
        dirs = []
        for sg in (1, -1):
            d = s["sdir"].copy()
            d[i] = Rotation.from_rotvec(sg * th * axis).apply(s["sdir"][i])
            dirs.append(evaluate(s, sdir=d)[0].sum())
        fd = (dirs[0] - dirs[1]) / (2 * th)
        an = float(np.dot(np.cross(axis, s["sdir"][i] * s["smag"][i]), g0[i]))
        say("  moment %d rotation: dE/dtheta %.10f  central difference %.10f" % (i, an, fd))
        assert abs(fd - an) < 1e-6 * max(1.0, abs(an))
        mags = []
        for sg in (1, -1):
            m = s["smag"].copy()
            m[i] += sg * 1e-5
            mags.append(evaluate(s, smag=m)[0].sum())
        fd = (mags[0] - mags[1]) / 2e-5
        an = float(np.dot(s["sdir"][i], g0[i]))
        say("  moment %d magnitude: dE/dmu %.10f  central difference %.10f" % (i, an, fd))
        assert abs(fd - an) < 1e-6 * max(1.0, abs(an))


def test_rotation_and_time_reversal_invariance(setup):
    pot, s = setup
    say("full size: rotation and time-reversal invariance")
    capi.init_file(pot)
    capi.set_precision(capi.FP64)
    e0, f0, g0 = evaluate(s)
    R = Rotation.from_rotvec([0.3, -0.5, 0.7])
    Rm = R.as_matrix()
    cell_r = s["cell"] @ Rm.T  # rows = rotated lattice vectors
    er, fr, gr = evaluate(s, pos=R.apply(s["pos"]), sdir=R.apply(s["sdir"]), cell=cell_r)
    assert np.abs(er - e0).max() < 1e-10 * np.abs(e0).max()
    assert np.abs(fr - R.apply(f0)).max() < 1e-10 * np.abs(f0).max()
    assert np.abs(gr - R.apply(g0)).max() < 1e-10 * np.abs(g0).max()
    et, ft, gt = evaluate(s, sdir=-s["sdir"])
    assert np.abs(et - e0).max() < 1e-10 * np.abs(e0).max()
    assert np.abs(ft - f0).max() < 1e-10 * np.abs(f0).max()
    assert np.abs(gt + g0).max() < 1e-10 * np.abs(g0).max()


def test_mixed_within_targets_at_full_size(setup):
    pot, s = setup
    say("full size: mixed vs FP64")
    capi.init_file(pot)
    capi.set_precision(capi.FP64)
    e64, f64, g64 = evaluate(s)
    capi.set_precision(capi.MIXED)
    e32, f32, g32 = evaluate(s)
    capi.set_precision(capi.FP64)
    df, dg = np.abs(f32 - f64).max(), np.abs(g32 - g64).max()
    say("  max |dF| %.3e eV/A  max |d dE/dm| %.3e eV/muB  max |dE_i| %.3e eV" % (df, dg, np.abs(e32 - e64).max()))
    assert df < 1e-4 and dg < 1e-4


def test_virial_matches_strain_differences(setup):
    """The virial W_ab = sum_p u_a F_b (u = x_i - x_j, F = dE_i/du; the tally of
    lammps/pair_spin_ace.cpp:635-636) from the GPU pair outputs equals the derivative of the
    total energy under a homogeneous deformation u -> (1 + eps) u of every pair vector: central
    differences with eps = +-1e-5 for all nine components (FP64, <= 1e-6 of max |W|).  The mixed
    path's virial is within 1e-5 of the FP64 one (relative to max |W|)."""
    pot, s = setup
    say("full size: virial vs homogeneous-strain differences (16000 atoms)")
    capi.init_file(pot)
    capi.set_precision(capi.FP64)
    a = nbr.neilist(s["pos"], s["sdir"] * s["smag"][:, None], s["cell"], 5.0)
    out = capi.acelib(a)
    _, _, _, vir = systems.assemble(a, out["eout"], out["flist"], out["tlist"], out["tlist0"])
    u = a["elist"].reshape(-1, 3) * a["dlist"][:, None]
    eps = 1e-5
    scale = np.abs(vir).max()
    for ia in range(3):
        for ib in range(3):
            e = []
            for sg in (1, -1):
                Fd = np.eye(3)
                Fd[ia, ib] += sg * eps
                up = u @ Fd.T
                d = np.linalg.norm(up, axis=1)
                b = dict(a, dlist=d, elist=(up / d[:, None]).ravel())
                e.append(capi.acelib(b)["eout"].sum())
            dE = (e[0] - e[1]) / (2 * eps)
            say("  W[%d,%d] = %.8f   dE/d eps_%d%d = %.8f" % (ib, ia, vir[ib, ia], ia, ib, dE))
            assert abs(dE - vir[ib, ia]) <= 1e-6 * max(1.0, scale)
    capi.set_precision(capi.MIXED)
    o32 = capi.acelib(a)
    capi.set_precision(capi.FP64)
    _, _, _, v32 = systems.assemble(a, o32["eout"], o32["flist"], o32["tlist"], o32["tlist0"])
    err = np.abs(v32 - vir).max() / scale
    say("  mixed virial rel err %.2e (max |W| %.3e eV)" % (err, scale))
    assert err <= 1e-5
