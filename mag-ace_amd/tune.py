"""Sweep of the density kernel's variants per potential header; writes the tuning table.

For each header below and both precisions, magace_tune times every (variant, neighbour split)
candidate on the 16000-atom bcc Fe workload (20x20x20 cells, jitter N(0, 0.05 A), cone spins,
r_c 5.0 A, seed 2028), twice; the per-candidate minimum decides.  Results go to

  mag-ace_amd/tuning.json          the record: per header and precision, ms per evaluation of
                                   every candidate, the choice and the built-in rule's time;
  mag-ace_amd/csrc/tuning_table.inc the table compiled into libmagace.so (rebuild after).

Needs a GPU.  Usage: python mag-ace_amd/tune.py [--out DIR] [--reps N]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

# the benchmark headers, the shape of the reference's example potential and the two headers of
# the medium shape class (names of mag-ace_amd/potgen.py)
HEADER_NAMES = ("H2000", "H500", "HFE", "H6W", "H4W")
KEYS = ("nb_a", "nb_m", "nr2_m", "l2m", "nr3_a", "nr3_m", "l3a", "l3m", "nr4_a", "nr4_m", "l4a", "l4m")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=HERE, help="directory for tuning.json and csrc/tuning_table.inc")
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--headers", nargs="*", default=list(HEADER_NAMES))
    args = ap.parse_args()
    import torch
    import __graft_entry__
    pkg = __graft_entry__.load_package()
    capi, potgen, systems = pkg.capi, pkg.potgen, pkg.systems
    dev = torch.device("cuda:0")
    s = systems.make_system(20, 2028, spins="cone")
    a = systems.pack_acelib(s["pos"], s["cell"], s["sdir"], s["smag"], 5.0)
    N, P = a["natom"], len(a["dlist"])
    T = lambda x, dt=torch.float64: torch.as_tensor(np.ascontiguousarray(x), dtype=dt, device=dev)
    ins = [T(a["offsets"], torch.int32)] + [T(a[k]) for k in ("dlist", "elist", "dlist_momi", "elist_momi",
                                                              "dlist_mom", "elist_mom")]
    outs = [torch.zeros(n, dtype=torch.float64, device=dev) for n in (N, 3 * P, 3 * P, 3 * N)]
    ptrs = [x.data_ptr() for x in ins] + [o.data_ptr() for o in outs]
    stream = torch.cuda.current_stream().cuda_stream
    record = {"workload": "bcc Fe 20x20x20 (%d atoms, %d pairs), cone spins, r_c 5.0 A, seed 2028; "
                          "ms per evaluation (magace_compute_device), minimum of two sweeps" % (N, P),
              "gpu": torch.cuda.get_device_name(0), "date": time.strftime("%Y-%m-%d"), "headers": {}}
    lines = []
    td = tempfile.mkdtemp(prefix="magace_tune_")
    for name in args.headers:
        h = getattr(potgen, name)
        pot = os.path.join(td, name + ".pbe.in")
        potgen.write(pot, h, capi.enumerate_basis(h), 2028)
        capi.init_file(pot)
        entry = {"header": [int(h[k]) for k in KEYS]}
        choice = {}
        for prec, mode in (("fp64", capi.FP64), ("mixed", capi.MIXED)):
            capi.set_precision(mode)
            # the built-in rule first, as the reference point
            capi.set_tuning()
            for _ in range(2):
                capi.compute_device(N, P, *ptrs, stream=stream)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(args.reps):
                capi.compute_device(N, P, *ptrs, stream=stream)
            torch.cuda.synchronize()
            rule_ms = 1e3 * (time.perf_counter() - t0) / args.reps
            sweeps = [capi.tune(N, P, *ptrs, reps=args.reps, stream=stream)[1] for _ in range(2)]
            ms = {c: min((sw[c] for sw in sweeps if sw[c] is not None), default=None) for c in capi.TUNE_CANDIDATES}
            best = min((c for c in ms if ms[c] is not None), key=lambda c: ms[c])
            choice[prec] = best
            entry[prec] = {"candidates": {"%d/%d" % c: (None if t is None else round(t, 4)) for c, t in ms.items()},
                           "choice": list(best), "choice_ms": round(ms[best], 4), "rule_ms": round(rule_ms, 4)}
            print("%-6s %-5s rule %.4f ms, best %s %.4f ms" % (name, prec, rule_ms, best, ms[best]), flush=True)
        capi.set_precision(capi.FP64)
        record["headers"][name] = entry
        lines.append("{{%s}, {%d, %d}, {%d, %d}},  // %s" % (", ".join(str(v) for v in entry["header"]),
                                                           *choice["fp64"], *choice["mixed"], name))
    os.makedirs(os.path.join(args.out, "csrc"), exist_ok=True)
    with open(os.path.join(args.out, "tuning.json"), "w") as f:
        json.dump(record, f, indent=1)
    with open(os.path.join(args.out, "csrc", "tuning_table.inc"), "w") as f:
        f.write("// Written by mag-ace_amd/tune.py from measurements on MI355X (record: mag-ace_amd/tuning.json).\n"
                "// One line per potential header: {header}, {FP64 variant, split}, {mixed variant, split}.\n")
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
