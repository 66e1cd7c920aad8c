"""Autotuner for the batched level engine's launch shapes.

What is tuned (``Level.tuning``):

* the block size (64, 128 or 256 threads: one to four wavefronts) of each of
  the 13 stage ranges of the fused engine (``CAMR_NSTAGE``): ghost fill /
  primitives on grow(bx,4), flattening / traces on grow(bx,2), first-flux
  faces (x, y, z), transdo faces (x, y, z), nodes, final faces (x, y, z)
  and the valid box;
* the order of the boxes in the batched launches (level order, or largest
  box first so that the longest boxes start first and the tail is short).

How: every stage kernel is timed with the level's hipEvent profiler over a
few steps for each candidate block size (one run per size covers every
stage, since a stage's cost depends only on its own launch shape); the
fastest size per stage range is kept (the three directions of a face stage
share one profiler name, so they share one size).  The combined choice is
then checked against the default on whole steps and kept only if faster.
Results are cached in a JSON file keyed by the device name and the
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


def tune(level, steps: int = 3, use_cache: bool = True, verbose: bool = False):
    """Choose and apply the launch shapes of `level` (which must hold a
    valid state; it is advanced by the timing runs).  Returns the tuning
    (blocks, order) now set on the level."""
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
    if key in cache:
        t = cache[key]
        level.tuning = (t["blocks"], t["order"])
        return level.tuning

    default = level.tuning
    level.run(1)  # warm-up (graph capture, first launches)
    # per-stage times for each uniform block size, both box orders
    best_blocks = list(default[0])
    times = {}
    for order in (0, 1):
        for bs in BLOCKS:
            level.tuning = ([bs] * _lib.NSTAGE, order)
            times[(order, bs)] = _stage_times(level, steps)
    results = []
    for order in (0, 1):
        blocks = [256] * _lib.NSTAGE
        for grp in GROUPS:
            names = [n for n, st in STAGES.items() if st == grp or set(st) & set(grp)]
            cost = {bs: sum(times[(order, bs)].get(n, 0.0) for n in names) for bs in BLOCKS}
            pick = min(cost, key=cost.get)
            for g in grp:
                blocks[g] = pick
        level.tuning = (blocks, order)
        results.append((_step_time(level, steps), blocks, order))
    level.tuning = default
    t_def = _step_time(level, steps)
    results.sort(key=lambda r: r[0])
    t_best, blocks, order = results[0]
    if t_best >= t_def:
        blocks, order = list(default[0]), default[1]
        t_best = t_def
    level.tuning = (blocks, order)
    if verbose:
        print(f"autotune {key}: default {t_def * 1e3:.3f} ms/step, chosen "
              f"{t_best * 1e3:.3f} ms/step, blocks {blocks}, order {order}", flush=True)
    cache[key] = {"blocks": list(map(int, blocks)), "order": int(order),
                  "ms_per_step": t_best * 1e3, "default_ms_per_step": t_def * 1e3}
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        with open(path, "w") as f:
            json.dump(cache, f, indent=1, sort_keys=True)
    except OSError:
        pass
    return level.tuning
