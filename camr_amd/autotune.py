"""Autotuner for the batched level engine's launch shapes.

What is tuned (``Level.tuning`` and ``Level.tiles``):

* the block of each of the 13 stage ranges of the fused engine
  (``CAMR_NSTAGE``: ghost fill / primitives on grow(bx,4), flattening /
  traces on grow(bx,2), first-flux faces (x, y, z), transdo faces (x, y, z),
  nodes, final faces (x, y, z) and the valid box): either a run of 64, 128
  or 256 consecutive points in x-fastest order, or a 3-D tile of points
  (e.g. 16 x 4 x 4) that keeps the y and z stencil neighbours of a block's
  points inside the block, so that they are re-read from its own caches
  rather than from HBM;
* the XCD chunk of each stage range: the hardware hands consecutive blocks
  to the eight XCDs in turn, so a block's neighbouring rows sit in other
  L2 caches; with chunk K every XCD takes K consecutive blocks of each
  window of 8K;
* the order of the boxes in the batched launches (level order, or largest
  box first so that the longest boxes start first and the tail is short);
* the transverse stages as the staged transdo / transdd launches or as the
  LDS-staged CTU face kernels (``Level.ctu``), a whole-step choice.

How: every stage kernel is timed with the level's hipEvent profiler over a
few steps for each candidate block (one run per candidate covers every
stage, since a stage's cost depends only on its own launch shape); the
fastest candidate per stage range is kept (the three directions of a face
stage share one profiler name, so they share one candidate).  The combined
choice is then checked against the default on whole steps and kept only if
faster.  Results are cached in a JSON file keyed by the device name and the
histogram of box sizes, so a level with the same box population reuses
them without timing anything.
"""
from __future__ import annotations

import hashlib
import json
import os
import time

import numpy as np

from . import _lib

BLOCKS = (64, 128, 256)
LINEAR = (0, 0, 0)
# 3-D tiles (x, y, z) of 256 or 128 points; x at least 8 for whole 64-byte
# row segments
TILES = ((64, 4, 1), (32, 8, 1), (32, 4, 2), (16, 4, 4), (64, 2, 2), (16, 8, 2),
         (8, 8, 4), (32, 2, 4), (16, 4, 2), (32, 2, 2))
XCD_CHUNKS = (0, 16)
# candidates: (block size, tile, XCD chunk); linear blocks carry LINEAR
CANDIDATES = tuple((bs, LINEAR, k) for bs in BLOCKS for k in XCD_CHUNKS) + tuple(
    (t[0] * t[1] * t[2], t, 0) for t in TILES[:4])
# profiler region name -> stage-range indices (CAMR_NSTAGE order)
STAGES = {
    "fill_clean_prim": (0,), "prim": (0,),
    "flatten": (1,), "trace_ppm": (1,), "trace_plm": (1,),
    "riemann_normal": (2, 3, 4),
    "transverse_1": (5, 6, 7),
    "divu": (8,),
    "transverse_2": (9, 10, 11),
    "consup": (12,), "clean_estdt": (12,), "consup_clean_estdt": (12,),
}
# groups of stage ranges that must share a block size (one profiler name)
GROUPS = [(0,), (1,), (2, 3, 4), (5, 6, 7), (8,), (9, 10, 11), (12,)]


def cache_path() -> str:
    return os.environ.get("CAMR_TUNING_CACHE",
                          os.path.join(os.path.expanduser("~"), ".cache", "camr_amd",
                                       "tuning.json"))


def key_of(level, device: str) -> str:
    """Device + ppm_type + histogram of box shapes."""
    shapes = sorted(tuple(b[3 + d] - b[d] + 1 for d in range(3)) for b in level.prob.boxes)
    h = hashlib.sha1(repr(shapes).encode()).hexdigest()[:16]
    return f"{device}|ppm{level.prob.value('ppm_type')}|{len(shapes)}boxes|{h}"


def _device_name() -> str:
    try:
        import torch
        return torch.cuda.get_device_name(0)
    except Exception:  # no torch / no device: key on the library only
        return "gfx950"


def _stage_times(level, steps: int) -> dict:
    """ms per step of every profiler region over `steps` steps."""
    level.profile(True)
    level.run(steps)
    rep = level.profile_report()
    level.profile(False)
    return {k: v["ms"] / steps for k, v in rep.items() if k != "advance"}


def _step_time(level, steps: int) -> float:
    level.sync()
    t0 = time.perf_counter()
    level.run(steps)
    level.sync()
    return (time.perf_counter() - t0) / steps


CTU_PLANES = 8


def apply(level, blocks, order, tiles, xcd=None, ctu=None):
    """Set per-stage blocks (linear block sizes or tiles), XCD chunks, the
    box order and the CTU switch; unnamed settings keep the library's
    defaults."""
    level.tuning = ([b if t == LINEAR else 256 for b, t in zip(blocks, tiles)], order)
    level.tiles = tiles
    if xcd is not None:
        level.xcd = xcd
    if ctu is not None:
        level.ctu = tuple(ctu)


def _cand_name(c) -> str:
    bs, tile, xk = c
    shape = f"{bs} linear" if tile == LINEAR else "x".join(map(str, tile))
    return f"{shape}, xcd {xk}"


def report_markdown(table, results, t_def) -> str:
    """The tuner's measurements as markdown: per stage group the time of
    every candidate (ms per step, level order), then the whole-step checks."""
    out = ["| stage ranges | " + " | ".join(_cand_name(c) for c in CANDIDATES) + " |",
           "|---|" + "---:|" * len(CANDIDATES)]
    for order, grp, cost in table:
        if order != 0:
            continue
        best = min(cost.values())
        cells = [(f"**{cost[c]:.3f}**" if cost[c] == best else f"{cost[c]:.3f}") for c in CANDIDATES]
        out.append(f"| {','.join(map(str, grp))} | " + " | ".join(cells) + " |")
    out += ["", "| whole step | ms |", "|---|---:|", f"| defaults | {t_def * 1e3:.3f} |"]
    for t, blocks, order, tiles, xcd, ctu in results:
        out.append(f"| per-stage picks, box order {order}, CTU {'on' if ctu[0] else 'off'} | "
                   f"{t * 1e3:.3f} |")
    return "\n".join(out)


def tune(level, steps: int = 3, use_cache: bool = True, verbose: bool = False):
    """Choose and apply the launch shapes of `level` (which must hold a
    valid state; it is advanced by the timing runs).  Returns the tuning
    (blocks, order, tiles, xcd, ctu) now set on the level: per stage range
    its block size, its tile ((0, 0, 0) for linear blocks) and its XCD
    chunk, the box order and the CTU switch (on, planes)."""
    device = _device_name()
    key = key_of(level, device)
    path = cache_path()
    cache = {}
    if use_cache and os.path.exists(path):
        try:
            with open(path) as f:
                cache = json.load(f)
        except (OSError, ValueError):
            cache = {}
    if key in cache and "ctu" in cache[key]:
        t = cache[key]
        tiles = [tuple(x) for x in t["tiles"]]
        ctu = tuple(t["ctu"])
        apply(level, t["blocks"], t["order"], tiles, t["xcd"], ctu)
        return list(t["blocks"]), t["order"], tiles, list(t["xcd"]), ctu

    default = (list(level.tuning[0]), level.tuning[1], list(level.tiles), list(level.xcd),
               tuple(level.ctu))
    level.run(1)  # warm-up (graph capture, first launches)
    # per-stage times for each uniform candidate, both box orders
    times = {}
    for order in (0, 1):
        for cand in CANDIDATES:
            bs, tile, xk = cand
            apply(level, [bs] * _lib.NSTAGE, order, [tile] * _lib.NSTAGE, [xk] * _lib.NSTAGE,
                  (0, CTU_PLANES))
            times[(order, cand)] = _stage_times(level, steps)
    results = []
    table = []  # (order, group, {candidate: ms per step})
    for order in (0, 1):
        blocks = [256] * _lib.NSTAGE
        tiles = [LINEAR] * _lib.NSTAGE
        xcd = [0] * _lib.NSTAGE
        for grp in GROUPS:
            names = [n for n, st in STAGES.items() if st == grp or set(st) & set(grp)]
            cost = {c: sum(times[(order, c)].get(n, 0.0) for n in names) for c in CANDIDATES}
            table.append((order, grp, cost))
            pick = min(cost, key=cost.get)
            for g in grp:
                blocks[g], tiles[g], xcd[g] = pick
        for ctu in ((0, CTU_PLANES), (1, CTU_PLANES)):
            apply(level, blocks, order, tiles, xcd, ctu)
            results.append((_step_time(level, steps), blocks, order, tiles, xcd, ctu))
    apply(level, *default)
    t_def = _step_time(level, steps)
    results.sort(key=lambda r: r[0])
    t_best, blocks, order, tiles, xcd, ctu = results[0]
    if t_best >= t_def:
        blocks, order, tiles, xcd, ctu = default
        t_best = t_def
    apply(level, blocks, order, tiles, xcd, ctu)
    if verbose:
        print(f"autotune {key}: default {t_def * 1e3:.3f} ms/step, chosen "
              f"{t_best * 1e3:.3f} ms/step, blocks {blocks}, tiles {tiles}, xcd {xcd}, "
              f"order {order}, ctu {ctu}", flush=True)
        print(report_markdown(table, results, t_def), flush=True)
    cache[key] = {"blocks": list(map(int, blocks)), "order": int(order),
                  "tiles": [list(map(int, t)) for t in tiles], "xcd": list(map(int, xcd)),
                  "ctu": list(map(int, ctu)),
                  "ms_per_step": t_best * 1e3, "default_ms_per_step": t_def * 1e3}
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    except OSError:
        pass
    return list(blocks), order, [tuple(t) for t in tiles], list(xcd), tuple(ctu)


def _main() -> int:
    """Tune the C1 level (256^3 periodic PPM, one box) and the two levels of
    the C2 hierarchy without the cache, printing every measurement."""
    import argparse
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--which", choices=["c1", "c2", "all"], default="all")
    args = ap.parse_args()
    from . import workloads
    from .hydro import Level
    if args.which in ("c1", "all"):
        prob = workloads.smooth_periodic(n=256, ppm_type=1, cfl=0.4)
        L = Level(prob)
        L.set_box(0, workloads.smooth_init(prob, 0))
        L.init_post()
        print("## C1: 256^3 periodic PPM, one box\n", flush=True)
        tune(L, steps=args.steps, use_cache=False, verbose=True)
        L.close()
    if args.which in ("c2", "all"):
        lev0, lev1 = workloads.amr_levels(128)
        C = Level(lev0)
        for b in range(len(lev0.boxes)):
            C.set_box(b, workloads.smooth_init(lev0, b))
        C.init_post()
        print("\n## C2 level 0: 128^3 in 64 boxes of 32^3\n", flush=True)
        tune(C, steps=args.steps, use_cache=False, verbose=True)
        ext1 = workloads.exterior(lev1)
        F = Level(lev1)
        F.set_exterior(ext1)
        for b in range(len(lev1.boxes)):
            lo, hi = lev1.boxes[b][:3], lev1.boxes[b][3:]
            F.set_box(b, ext1[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1].copy())
        F.init_post()
        print(f"\n## C2 level 1: {len(lev1.boxes)} boxes of 8^3 to 64^3\n", flush=True)
        tune(F, steps=args.steps, use_cache=False, verbose=True)
        F.close()
        C.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(_main())
