"""Density kernel variants (magace_set_tuning / magace_tune) on the GPU.

Every (variant, neighbour split) candidate the tuner may choose gives the reference's
results: FP64 within 1e-10 of golden vectors written by the reference libace, for the
lmax-4 and lmax-6 headers and the shape of the example potential.  magace_tune times every
candidate on a device-resident system, keeps the fastest and reports its time.
"""
import sys

import numpy as np
import pytest

from helpers import PKG, load_case, relerr, write_potential

capi, systems = PKG.capi, PKG.systems
torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no GPU visible", allow_module_level=True)

DEV = torch.device("cuda:0")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name", ["h500_128", "h2000_16", "hfe_16", "h6w_16", "h4w_16"])
def test_every_variant_matches_reference(name, tmp_path):
    c = load_case(name, tmp_path)
    capi.init_file(c["pot"])
    for prec in (capi.FP64, capi.MIXED):
        capi.set_precision(prec)
        for cand in capi.TUNE_CANDIDATES:
            setting = dict(fp64=cand) if prec == capi.FP64 else dict(mixed=cand)
            capi.set_tuning(**setting)
            got = capi.acelib(c["a"])
            for k in ("eout", "flist", "tlist", "tlist0"):
                e = relerr(got[k], c["ref"][k])
                assert e <= (1e-10 if prec == capi.FP64 else 1e-4), (name, prec, cand, k, e)
        say("  %s precision %d: all %d candidates within the bar" % (name, prec, len(capi.TUNE_CANDIDATES)))
    capi.set_precision(capi.FP64)


def test_tune_keeps_the_fastest(tmp_path):
    pot = write_potential(str(tmp_path / "p.in"), "H500", 2027)
    capi.init_file(pot)
    capi.set_precision(capi.MIXED)
    s = systems.make_system(8, 3, "uniform")
    a = systems.pack_acelib(s["pos"], s["cell"], s["sdir"], s["smag"], 5.0)
    t = lambda x, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(x), dtype=dt, device=DEV)
    N, P = a["natom"], len(a["dlist"])
    ins = [t(a["offsets"], torch.int32)] + [t(a[k]) for k in ("dlist", "elist", "dlist_momi", "elist_momi",
                                                                "dlist_mom", "elist_mom")]
    outs = [torch.zeros(n, dtype=torch.float64, device=DEV) for n in (N, 3 * P, 3 * P, 3 * N)]
    best, ms = capi.tune(N, P, *[x.data_ptr() for x in ins], *[o.data_ptr() for o in outs], reps=3,
                         stream=torch.cuda.current_stream().cuda_stream)
    say("  tune (%d atoms, mixed): best %s, ms %s" % (N, best, ms))
    assert best in capi.TUNE_CANDIDATES
    assert ms[best] is not None and ms[best] > 0
    assert all(v is None or v >= ms[best] for v in ms.values())
    assert capi.get_tuning()[1] == best
    capi.set_precision(capi.FP64)
