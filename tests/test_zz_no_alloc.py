"""magace_compute_device_noalloc: the evaluation with its per-atom intermediates in the
library's fixed device scratch instead of buffers grown with hipMalloc.

Its outputs equal magace_compute_device's on the same device inputs (FP64 <= 1e-10 norm-wise
relative, mixed <= 1e-6) on two calls in a row, the second with a smaller input, in both
precisions; a system too large for the scratch is rejected with MAGACE_ERR_INPUT before any
launch, and the library stays usable.
"""
import sys

import numpy as np
import pytest

from helpers import PKG, relerr, write_potential

capi, systems = PKG.capi, PKG.systems
torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

if not torch.cuda.is_available():
    pytest.skip("no GPU visible", allow_module_level=True)

DEV = torch.device("cuda:0")
KEYS = ("eout", "flist", "tlist", "tlist0")


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def device_inputs(a):
    t = lambda x, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(x), dtype=dt, device=DEV)
    N, P = a["natom"], len(a["dlist"])
    ins = [t(a["offsets"], torch.int32)] + [t(a[k]) for k in ("dlist", "elist", "dlist_momi", "elist_momi",
                                                                "dlist_mom", "elist_mom")]
    return N, P, ins


def run(fn, N, P, ins):
    outs = [torch.full((n,), np.nan, dtype=torch.float64, device=DEV) for n in (N, 3 * P, 3 * P, 3 * N)]
    stream = torch.cuda.current_stream().cuda_stream
    fn(N, P, *[x.data_ptr() for x in ins], *[o.data_ptr() for o in outs], stream=stream)
    capi.check_errors(stream)
    return {k: o.cpu().numpy() for k, o in zip(KEYS, outs)}


@pytest.mark.parametrize("prec", [capi.FP64, capi.MIXED])
def test_noalloc_matches_compute_device(prec, tmp_path):
    pot = write_potential(str(tmp_path / "p.in"), "H500", 2027)
    capi.init_file(pot)
    capi.set_precision(prec)
    tol = 1e-10 if prec == capi.FP64 else 1e-6
    # 2176 atoms (8x8x17 cells: the evaluation runs as two halves on two streams), then 128
    rng = np.random.default_rng(2027)
    pos, cell = systems.bcc_box(8, 8, 17)
    sdir, smag = systems.spins_uniform(len(pos), rng)
    first = dict(pos=systems.jitter(pos, rng), cell=cell, sdir=sdir, smag=smag)
    for s in (first, systems.make_system(4, 2026, "uniform")):
        a = systems.pack_acelib(s["pos"], s["cell"], s["sdir"], s["smag"], 5.0)
        N, P, ins = device_inputs(a)
        say("noalloc: precision %d, %d atoms, %d pairs" % (prec, N, P))
        want = run(capi.compute_device, N, P, ins)
        got = run(capi.compute_device_noalloc, N, P, ins)
        for k in KEYS:
            assert np.isfinite(got[k]).all(), k
            e = relerr(got[k], want[k])
            say("  %-6s rel %.2e" % (k, e))
            assert e <= tol, (k, e)
    capi.set_precision(capi.FP64)


def test_noalloc_rejects_systems_beyond_the_scratch(tmp_path):
    pot = write_potential(str(tmp_path / "p.in"), "H2000", 2028)
    capi.init_file(pot)
    capi.set_precision(capi.FP64)
    # 2000 atoms, beyond the FP64 limit of this header (about 1440 atoms): rejected on the
    # host, nothing launched
    s = systems.make_system(10, 2028, "cone")
    a = systems.pack_acelib(s["pos"], s["cell"], s["sdir"], s["smag"], 5.0)
    N, P, ins = device_inputs(a)
    say("noalloc: %d atoms beyond the scratch" % N)
    with pytest.raises(capi.MagaceError, match="exceed the fixed scratch"):
        run(capi.compute_device_noalloc, N, P, ins)
    # still usable: a small system through both entry points
    s = systems.make_system(2, 2028, "cone")
    a = systems.pack_acelib(s["pos"], s["cell"], s["sdir"], s["smag"], 5.0)
    N, P, ins = device_inputs(a)
    want = run(capi.compute_device, N, P, ins)
    got = run(capi.compute_device_noalloc, N, P, ins)
    for k in KEYS:
        assert relerr(got[k], want[k]) <= 1e-10, k
