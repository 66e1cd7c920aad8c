"""Algorithmic FP64 operations per cell-update of the reference's path
(TEST INFRASTRUCTURE ONLY).

Builds the oracle with its operation-counting real (oracle/opcount.cpp:
add/sub, mul, div, sqrt = 1 each; no FMA, since the reference CPU build does
not contract), runs the seeded smooth periodic PPM workload untiled (one
box: the GPU build's TilingIfNotGPU), and prints the operations of one full
level step (computeNewDt + CAMR_advance) per cell-update, by stage.

    python oracle/count_ops.py --n 256
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import tempfile
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
STAGES = ["fill_clean_prim", "trace", "riemann_normal", "transverse_1", "divu",
          "transverse_2", "consup", "clean_estdt", "flatten"]
OPS = ["add", "mul", "div", "sqrt"]


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--n", type=int, default=64)
    ap.add_argument("--ppm", type=int, default=1)
    ap.add_argument("--json", action="store_true")
    args = ap.parse_args()
    out = tempfile.mkdtemp(prefix="oracle_opcount_")
    subprocess.run(["make", "-s", "-C", HERE, "opcount", f"OUT={out}"], check=True)
    os.environ["CAMR_ORACLE_LIB"] = os.path.join(out, "liboracle_opcount.so")
    sys.path.insert(0, ROOT)
    from camr_amd import workloads
    from oracle import oracle as O
    lib = O.lib()
    lib.ora_opcount_get.argtypes = [C.POINTER(C.c_ulonglong)]
    n = args.n
    prob = workloads.smooth_periodic(n=n, ppm_type=args.ppm, cfl=0.4)
    p = O.params(ppm_type=args.ppm, cfl=0.4, bclo=prob.bclo, bchi=prob.bchi,
                 domlo=prob.domlo, domhi=prob.domhi, is_periodic=prob.is_periodic,
                 dx=prob.dx, tile=(1 << 20,) * 3)
    L = O.Level(p, prob.boxes)
    L.set_box(0, workloads.smooth_init(prob, 0))
    L.init_post()
    dt = L.initial_dt()
    assert L.advance(dt) == 0
    lib.ora_opcount_reset()
    t0 = time.perf_counter()
    dt = L.new_dt(dt)
    assert L.advance(dt) == 0
    el = time.perf_counter() - t0
    raw = (C.c_ulonglong * (len(STAGES) * len(OPS)))()
    lib.ora_opcount_get(raw)
    cells = n ** 3
    per = {s: {o: raw[i * len(OPS) + k] / cells for k, o in enumerate(OPS)}
           for i, s in enumerate(STAGES)}
    if args.ppm == 1:
        # trace_ppm recomputes the flattening coefficient of every cell in
        # each of its three direction passes (PPM.cpp:120-126); the
        # algorithm needs it once
        per["flatten"] = {o: v / 3 for o, v in per["flatten"].items()}
    total = sum(sum(v.values()) for v in per.values())
    res = {"n": n, "ppm_type": args.ppm, "cells": cells, "seconds": el,
           "per_cell_update": per, "total_per_cell_update": total}
    if args.json:
        print(json.dumps(res))
    else:
        print(f"{n}^3 PPM={args.ppm}: {total:.1f} FP64 operations per cell-update "
              f"({el:.1f} s single-threaded)")
        for s in STAGES:
            v = per[s]
            print(f"  {s:16s} " + " ".join(f"{o} {v[o]:8.1f}" for o in OPS) +
                  f"  total {sum(v.values()):8.1f}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
