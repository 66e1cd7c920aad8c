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
    "consup": (12,), "clean_estdt": (12,),
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
    for order in (0, 1):
        blocks = [256] * _lib.NSTAGE
        tiles = [LINEAR] * _lib.NSTAGE
        xcd = [0] * _lib.NSTAGE
        for grp in GROUPS:
            names = [n for n, st in STAGES.items() if st == grp or set(st) & set(grp)]
            cost = {c: sum(times[(order, c)].get(n, 0.0) for n in names) for c in CANDIDATES}
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
