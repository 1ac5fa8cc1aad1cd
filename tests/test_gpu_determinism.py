"""Bitwise repeatability of the evaluation on the GPU.

The adjoint's partial sums meet in a fixed order (no atomics) and each atom's walk over the
gradient records depends only on its index in the whole system, so an evaluation gives the
same bits on every launch, whether the atoms run in one launch, as two halves on two
streams, or in the chunks of acelib_ on host arrays (FP64 and mixed precision), also with
body orders 5 and 6, whose products of A entries send their gradient back in a fixed order.
"""
import os
import sys

import numpy as np
import pytest

from helpers import PKG, write_potential

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


def system_2176(seed=2027):
    """8x8x17 bcc cells (2176 atoms): above the two-halves threshold of 2048 atoms."""
    rng = np.random.default_rng(seed)
    pos, cell = systems.bcc_box(8, 8, 17)
    sdir, smag = systems.spins_uniform(len(pos), rng)
    return systems.pack_acelib(systems.jitter(pos, rng), cell, sdir, smag, 5.0)


def device_eval(a):
    t = lambda x, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(x), dtype=dt, device=DEV)
    N, P = a["natom"], len(a["dlist"])
    ins = [t(a["offsets"], torch.int32)] + [t(a[k]) for k in ("dlist", "elist", "dlist_momi", "elist_momi",
                                                                "dlist_mom", "elist_mom")]
    outs = [torch.zeros(n, dtype=torch.float64, device=DEV) for n in (N, 3 * P, 3 * P, 3 * N)]
    stream = torch.cuda.current_stream().cuda_stream
    capi.compute_device(N, P, *[x.data_ptr() for x in ins], *[o.data_ptr() for o in outs], stream=stream)
    capi.check_errors(stream)
    return {k: o.cpu().numpy() for k, o in zip(KEYS, outs)}


def same(x, y):
    return all(np.array_equal(x[k], y[k]) for k in KEYS)


@pytest.mark.parametrize("header", ["H500", "H2000", "H56"])
@pytest.mark.parametrize("prec", [capi.FP64, capi.MIXED, capi.MIXED_ACC])
def test_bitwise_repeatable_across_launches_and_splits(header, prec, tmp_path):
    pot = write_potential(str(tmp_path / "p.in"), header, 2027)
    capi.init_file(pot)
    capi.set_precision(prec)
    a = system_2176()
    say("determinism: %s precision %d, %d atoms" % (header, prec, a["natom"]))
    old = os.environ.get("MAGACE_SPLIT_STREAMS")
    old_chunk = os.environ.get("MAGACE_ACELIB_CHUNK_PAIRS")
    try:
        os.environ["MAGACE_SPLIT_STREAMS"] = "1"
        r1 = device_eval(a)
        r2 = device_eval(a)
        assert same(r1, r2), "two launches differ"
        os.environ["MAGACE_SPLIT_STREAMS"] = "0"
        r3 = device_eval(a)
        assert same(r1, r3), "two halves differ from one sequence"
        # acelib_ on host arrays: chunks of whole atoms (here about 20 chunks)
        os.environ["MAGACE_ACELIB_CHUNK_PAIRS"] = "6000"
        h = capi.acelib(a)
        assert same(r1, h), "chunked host path differs"
    finally:
        for k, v in (("MAGACE_SPLIT_STREAMS", old), ("MAGACE_ACELIB_CHUNK_PAIRS", old_chunk)):
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
        capi.set_precision(capi.FP64)
