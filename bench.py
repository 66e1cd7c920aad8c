#!/usr/bin/env python3
"""Benchmark: FP64 cell-updates/s of CAMR's single-level Godunov PPM advance
on a 256^3 periodic grid (one MI355X).

A step is one full level step as CAMR takes it: computeNewDt (CFL min
reduction + host read-back), then CAMR_advance (clean_state, ghost fill,
ctoprim, PPM CTU hydro, flux divergence and update, clean_state).  Inputs
are resident in HBM before the timed region.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0      # MI355X HBM3E, GB/s
FP64_PEAK_TFS = 78.6       # MI355X FP64 vector, TFLOP/s
NVAR = 10

# Kernel names of the profiler regions and the per-launch algorithmic HBM
# bytes of each, per unit (face or cell) it processes: every array element
# the kernel must read or write, counted once (neighbour re-reads excluded).
#   transverse_2 (one launch per direction d): per d-face: normal states qm, qp (2 x 9), the two
#     corrected transverse fluxes + Godunov (u, p) of its cell (2 x 10),
#     sound speed / floor (2), conserved state of the cell (10, artificial
#     viscosity), one divu node; writes the NVAR flux and Godunov (u, p)
#     = 18 + 20 + 2 + 10 + 1 + 12 = 63 doubles.
#   transverse_1: per d-face: qm, qp (18), two transverse fluxes +
#     Godunov of its cell (2 x 10), sound speed / floor (2); writes two
#     compact fluxes + Godunov (2 x 10) = 60 doubles.
#   trace_ppm: per cell: compact primitives (9 + sound speed) and flat (1);
#     writes the two face states (2 x 9) = 29 doubles.
#   riemann_normal: per face: qm, qp (18), sound speed / floor of both
#     cells (2); writes the compact flux + Godunov (10) = 30 doubles.
#   transverse_ctu (the CTU face kernels, transdo pairs + transdd + final
#     solve of one direction): per d-face: qm, qp of all three directions
#     (54), the first fluxes + Godunov of the two transverse directions
#     (20), sound speed / floor (2), conserved state (10), one divu node;
#     writes the final flux and Godunov (12) = 99 doubles.
ALGO_DOUBLES = {"transverse_2": 63, "transverse_1": 60, "trace_ppm": 29,
                "riemann_normal": 30, "transverse_ctu": 99}
# mol_faces (one launch per direction d, method of lines, flux adjustment
# fused): per d-face, the 9 primitives and 3 aux values (c, csmall, gamma)
# of one cell, its 10 conserved values (artificial viscosity) and one divu
# node read; the NVAR fluxes and Godunov (u, p) written.
NS_MOL_READ, NF_MOL_WRITE = 23, 12

# HBM bytes per launch measured with rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE
# for the fused engine at 256^3 PPM with the default launch shapes
# (profiles/hbm_traffic_fused_256cube_ppm.md, profiles/ctu_faces_256cube_ppm.md;
# FETCH_SIZE scaled by 1.9692, WRITE_SIZE by 0.9961, the factors an
# 8-byte-lane copy of 1 GiB reads back).  Reported only for that engine and size.
MEASURED_TRAFFIC = {("fused", 256, "transverse_2"): 10447411888,
                    ("fused", 256, "transverse_ctu"): 28527364887}

# Executed FP64 work of one PPM cell-update (ADD, MUL, TRANS = 1, FMA = 2 per
# lane; FMAs only from the division / square-root expansions, contraction is
# off), summed over the step's kernels from rocprofv3 SQ_INSTS_VALU_*_F64
# counters at 256^3 (profiles/valu_fused_256cube_ppm.md).
FP64_EXECUTED_PER_CELL_UPDATE = 10069

# Algorithmic FP64 operations of the reference's path per cell-update, by
# stage (add/sub, mul, div, sqrt = 1 each; no FMA): the oracle run with an
# operation-counting real on the 256^3 PPM workload, untiled
# (`python oracle/count_ops.py --n 256`, profiles/fp64_opcount_256cube_ppm.md).
# Keys are the profiler regions of the fused engine; a face stage's count
# covers its three launches (one per direction).
ALGO_FLOP = {"fill_clean_prim": 169.0, "flatten": 59.3, "trace_ppm": 3501.9,
             "riemann_normal": 300.6, "transverse_1": 1285.1, "divu": 30.4,
             "transverse_2": 989.9, "consup_clean_estdt": 111.0 + 129.0}
ALGO_FLOP_PER_CELL_UPDATE = sum(ALGO_FLOP.values())
# the CTU face kernels do the work of both transverse stages
ALGO_FLOP["transverse_ctu"] = ALGO_FLOP["transverse_1"] + ALGO_FLOP["transverse_2"]
LAUNCHES_PER_STEP = {"trace_ppm": 3, "riemann_normal": 3, "transverse_1": 3,
                     "transverse_2": 3, "transverse_ctu": 3}


def unit_count(name: str, n: int) -> int:
    """Faces or cells one launch of kernel `name` covers for an n^3 box."""
    if name.startswith("transverse_2") or name.startswith("transverse_ctu"):
        return (n + 1) * n * n
    if name.startswith("transverse_1"):
        return (n + 3) * (n + 2) * (n + 2)
    if name.startswith("trace_ppm"):
        return (n + 4) ** 3
    if name.startswith("riemann_normal"):
        return (n + 3) * (n + 4) * (n + 4)
    return n ** 3


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n: int, cfl: float, U0, warmup: int = 1, nsteps: int = 5):
    """The same workload with the CPU oracle (the reference's stages
    restated in C, OpenMP over hydro tiles) on this host's cores: `warmup`
    untimed steps, then the median of `nsteps` timed steps, each step
    computeInitialDt / computeNewDt + CAMR_advance.  Returns the baseline and
    the oracle's state and dt sequence after all its steps."""
    import numpy as np
    from oracle import oracle as O
    O.build()
    from camr_amd import workloads
    prob = workloads.smooth_periodic(n=n, ppm_type=1, cfl=cfl)
    p = O.params(ppm_type=1, cfl=cfl, bclo=prob.bclo, bchi=prob.bchi,
                 domlo=prob.domlo, domhi=prob.domhi, is_periodic=prob.is_periodic,
                 dx=prob.dx)
    L = O.Level(p, prob.boxes)
    L.set_box(0, U0)
    L.init_post()
    dts, times = [], []
    dt = None
    for s in range(warmup + nsteps):
        t0 = time.perf_counter()
        dt = L.initial_dt() if s == 0 else L.new_dt(dt)
        err = L.advance(dt)
        el = time.perf_counter() - t0
        assert err == 0
        dts.append(dt)
        if s >= warmup:
            times.append(el)
    state = L.get_box(0)
    L.close()
    med = float(np.median(times))
    threads = O.lib().ora_get_threads()
    return ({"value": n ** 3 / med, "unit": "cell-updates/s", "cores": threads,
             "kind": "port", "cpu_model": cpu_model(),
             "sample": f"the same {n}^3 PPM workload from the same initial state: "
                       f"{warmup} warm-up step, then the median of {nsteps} timed steps "
                       f"(computeNewDt + CAMR_advance), {med:.2f} s per step on "
                       f"{threads} OpenMP threads, hydro tiles (1024,16,16)"},
            state, np.array(dts))


def _amr_level(prob, mode, coarse, init_field):
    """Level in one of the launch modes: per-box, batched, batched + graph;
    with `coarse`, its ghost cells outside its boxes come from that level by
    FillPatch's coarse-fine interpolation."""
    from camr_amd.hydro import Level
    L = Level(prob, batched=(mode != "per_box"), graph=(mode == "graph"))
    if coarse is not None:
        L.exterior_from_coarse(coarse, 2)
    for b in range(len(prob.boxes)):
        L.set_box(b, init_field(b))
    L.init_post()
    return L


def bench_amr(args) -> int:
    """Batched AMR levels: per-box launches vs one launch per stage over all
    boxes vs captured step graphs, on the same two-level hierarchy."""
    import numpy as np
    import torch
    from camr_amd import autotune, workloads
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    lev0, lev1 = workloads.amr_levels(args.amr_n0)
    ext1 = workloads.exterior(lev1)
    # the coarse level's initial state: the source of level 1's ghost data
    C0 = _amr_level(lev0, "batched", None, lambda b: workloads.smooth_init(lev0, b))
    levels = []
    for name, prob, ext in (("level0", lev0, None), ("level1", lev1, ext1)):
        n = prob.domhi[0] + 1
        if ext is not None:
            def init(b, prob=prob, ext=ext):
                lo, hi = prob.boxes[b][:3], prob.boxes[b][3:]
                return ext[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1]
        else:
            def init(b, prob=prob):
                return workloads.smooth_init(prob, b)
        cells = sum((b[3] - b[0] + 1) * (b[4] - b[1] + 1) * (b[5] - b[2] + 1) for b in prob.boxes)
        res = {"boxes": len(prob.boxes), "cells": cells, "index_space": n}
        finals = {}
        for mode in ("per_box", "batched", "graph"):
            src = C0 if ext is not None else None
            L = _amr_level(prob, mode, src, init)
            if mode != "per_box" and not args.no_tune:
                # launch shapes from the autotuner (cached per device and box
                # population); the tuning runs advance a scratch copy
                T = _amr_level(prob, mode, src, init)
                res["tuning"] = autotune.tune(T, steps=2)
                T.close()
                autotune.apply(L, *res["tuning"])
            L.run(args.warmup)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            L.run(args.steps)
            torch.cuda.synchronize()
            el = time.perf_counter() - t0
            res[mode] = {"ms_per_step": el / args.steps * 1e3,
                         "cell_updates_per_s": cells * args.steps / el}
            finals[mode] = [L.get_box(b) for b in range(len(prob.boxes))]
            L.close()
        res["identical"] = all(np.array_equal(a, b) for m in ("batched", "graph")
                               for a, b in zip(finals["per_box"], finals[m]))
        res["speedup_batched"] = res["per_box"]["ms_per_step"] / res["batched"]["ms_per_step"]
        res["speedup_graph"] = res["per_box"]["ms_per_step"] / res["graph"]["ms_per_step"]
        levels.append((name, res))
    out = {"metric": "FP64 cell-updates/s of the batched AMR level step (PPM)",
           "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "higher_is_better": True, "dtype": "f64", "data": "synthetic",
           "config": {"workload": f"level 0 {args.amr_n0}^3 periodic in 32^3 boxes; level 1 "
                                  f"(ratio 2) on the sphere r=0.3, boxes 8..64 clustered "
                                  f"(blocking factor 8, efficiency 0.8); level-1 ghosts "
                                  f"outside its boxes interpolated from level 0 (FillPatch)"},
           "levels": dict(levels)}
    tot_cells = sum(r["cells"] for _, r in levels)
    t_graph = sum(r["graph"]["ms_per_step"] for _, r in levels)
    t_box = sum(r["per_box"]["ms_per_step"] for _, r in levels)
    out["value"] = tot_cells / (t_graph * 1e-3)
    out["speedup_vs_per_box"] = t_box / t_graph
    print(json.dumps(out), flush=True)
    return 0


def bench_sod(args) -> int:
    """The reference's CPU-runnable case: Sod 256x4x4, PLM, 100 steps (dt
    from the CFL estimate every step), on the GPU as per-step launches and
    as captured step graphs, beside the oracle on the host cores; the three
    final states must be bit-identical."""
    import numpy as np
    import torch
    from camr_amd import workloads
    from camr_amd.hydro import Level
    from oracle import oracle as O
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    prob = workloads.sod_tube()
    cells = 256 * 4 * 4
    nsteps = 100
    res, finals = {}, {}
    for mode in ("launches", "graph"):
        L = Level(prob, graph=(mode == "graph"))
        L.set_box(0, workloads.sod_init(prob, 0))
        L.init_post()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L.run(nsteps)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        res[mode] = {"ms_per_step": el / nsteps * 1e3, "cell_updates_per_s": cells * nsteps / el}
        finals[mode] = L.get_box(0)
        L.close()
    O.build()
    p = O.params(ppm_type=0, cfl=0.5, bclo=prob.bclo, bchi=prob.bchi, domlo=prob.domlo,
                 domhi=prob.domhi, is_periodic=prob.is_periodic, dx=prob.dx)
    OL = O.Level(p, prob.boxes)
    OL.set_box(0, workloads.sod_init(prob, 0))
    OL.init_post()
    t0 = time.perf_counter()
    err, _ = OL.run(nsteps)
    el = time.perf_counter() - t0
    assert err == 0
    ref = OL.get_box(0)
    OL.close()
    out = {"metric": "FP64 cell-updates/s, Sod 256x4x4 PLM, 100 steps", "unit": "cell-updates/s",
           "value": res["graph"]["cell_updates_per_s"], "higher_is_better": True, "n_gpus": 1,
           "steps": nsteps, "warmup": 0, "dtype": "f64", "data": "synthetic: Sod states",
           "vs_baseline": None,
           "config": {"workload": "Sod shock tube 256x4x4, PLM (plm_iorder 4), CFL 0.5, "
                                  "outflow in x, periodic in y/z, dt from the CFL estimate"},
           "modes": res,
           "identical_to_oracle": all(np.array_equal(f, ref) for f in finals.values()),
           "cpu_baseline": {"value": cells * nsteps / el, "unit": "cell-updates/s",
                            "cores": O.lib().ora_get_threads(), "kind": "port",
                            "sample": f"the same 100 steps, {el * 1e3:.1f} ms wall"}}
    print(json.dumps(out), flush=True)
    return 0


def bench_eb(args) -> int:
    """Sod past a sphere (embedded boundary, FluxRedist): (2n) x n x n cells
    in 32^3 boxes, PLM, CFL 0.3.  The level takes the regular path in boxes
    without cut cells and hydro_umdrv_eb in the others.  The first two
    steps are checked bit for bit against the oracle (timed on the host
    cores as the CPU baseline), then `warmup` + `steps` steps are timed."""
    import numpy as np
    import torch
    from camr_amd import _lib, workloads
    from camr_amd.hydro import Level
    from oracle import oracle as O
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    n = args.eb_n
    prob = workloads.sod_sphere(n=n, max_grid=32)
    t0 = time.perf_counter()
    geom = workloads.sod_sphere_geometry(prob)
    t_geom = time.perf_counter() - t0
    cells = 2 * n * n * n
    init = [workloads.sod_sphere_init(prob, b, geom) for b in range(len(prob.boxes))]
    L = Level(prob, engine=_lib.ENGINE_STAGED, graph=True, worker_streams=args.workers)
    L.set_eb(geom, merge=not args.no_eb_merge)
    kinds = list(L.kinds)
    nvirt, nmerged = L.merge_info()
    for b, U in enumerate(init):
        L.set_box(b, U)
    L.init_post()
    gdts = L.run(2)
    identical = None
    cpu = None
    if not args.no_cpu_baseline:
        O.build()
        p = O.params(ppm_type=0, cfl=prob.value("cfl"), bclo=prob.bclo, bchi=prob.bchi,
                     domlo=prob.domlo, domhi=prob.domhi, is_periodic=prob.is_periodic,
                     dx=prob.dx, tile=(1 << 20,) * 3)
        OL = O.Level(p, prob.boxes)
        OL.set_eb(geom)
        for b, U in enumerate(init):
            OL.set_box(b, U)
        OL.init_post()
        t0 = time.perf_counter()
        err, odts = OL.run(2)
        el = time.perf_counter() - t0
        assert err == 0
        identical = bool(np.array_equal(np.asarray(gdts), odts) and
                         all(np.array_equal(L.get_box(b), OL.get_box(b))
                             for b in range(len(prob.boxes))))
        OL.close()
        cpu = {"value": cells * 2 / el, "unit": "cell-updates/s",
               "cores": O.lib().ora_get_threads(), "kind": "port",
               "sample": f"the first 2 steps of the same workload, {el:.2f} s wall"}
    L.run(args.warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L.run(args.steps)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    value = cells * args.steps / el
    force = [float(x) for x in L.eb_force()]
    out = {"metric": f"FP64 cell-updates/s, Sod past a sphere {2 * n}x{n}x{n} (EB, FluxRedist)",
           "unit": "cell-updates/s", "value": value, "higher_is_better": True, "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic: Sod states around a sphere",
           "config": {"workload": f"{2 * n}x{n}x{n} on [0,2]x[0,1]^2, sphere r=0.15 at "
                                  "(0.6,0.5,0.5), Sod split at x=0.3, PLM, CFL 0.3, outflow x, "
                                  "slip walls y/z, 32^3 boxes, one captured graph per step",
                      "boxes": {"regular": kinds.count(0), "cut_cell": kinds.count(1),
                                "covered": kinds.count(2)},
                      "regular_batch": {"virtual_boxes": nvirt, "boxes_merged": nmerged,
                                        "boxes_alone": kinds.count(0) - nmerged},
                      "geometry_setup_s": round(t_geom, 2)},
           "identical_to_oracle_2_steps": identical,
           "force_on_sphere_after_run": force,
           "roofline": {"bound": "hbm", "achieved": 160.0 * value / 1e9, "peak": 8000.0,
                        "unit": "GB/s", "frac": 160.0 * value / 8e12, "traffic": None,
                        "kernel": "whole step, 160 B per cell-update"},
           "cpu_baseline": cpu}
    if not args.no_adjoint and args.eb_adj_steps > 0:
        L.close()
        out["adjoint"] = eb_adjoint_measure(prob, geom, init, args)
    print(json.dumps(out), flush=True)
    return 0


def eb_adjoint_measure(prob, geom, init, args) -> dict:
    """Adjoint of the x-force on the sphere after `eb_adj_steps` steps from
    the Sod state (checkpoints every 5, dt frozen from the primal run):
    backward over forward wall time; then the gradient of the same
    functional after `eb_fd_steps` steps from a smooth state on the same grid
    (``workloads.sod_sphere_smooth_init``) against central differences along
    seeded unit directions over the hydro components of the non-covered
    cells.  The Sod state's discontinuity puts limiters on their switches,
    where differences measure secants; its one-sided slope gaps are
    reported beside."""
    import numpy as np
    import torch
    from camr_amd import _lib, problems, workloads
    from camr_amd.hydro import Level
    nb = len(prob.boxes)

    def level(states):
        L = Level(prob, engine=_lib.ENGINE_STAGED, graph=True, worker_streams=args.workers)
        L.set_eb(geom)
        for b, U in enumerate(states):
            L.set_box(b, U)
        L.init_post()
        return L

    res = {}
    nsteps, every = args.eb_adj_steps, 5
    L = level(init)
    S0 = [L.get_box(b) for b in range(nb)]
    dts = L.run(nsteps)
    force_T = [float(x) for x in L.eb_force()]

    def run_from(states, dts_):
        for b in range(nb):
            L.set_box(b, states[b])
        for dt in dts_:
            L.advance(dt)

    def grad_of(states, dts_):
        for b in range(nb):
            L.set_box(b, states[b])
        L.adjoint_forward(dts_, every)
        L.seed_force_adjoint((1.0, 0.0, 0.0))
        L.adjoint_backward()
        return [L.get_adjoint(b) for b in range(nb)]

    grad_of(S0, dts)  # warm-up: allocates the checkpoints and the adjoint arrays
    times = {"forward": [], "forward_checkpointed": [], "backward": []}
    for _ in range(max(1, args.adj_reps)):
        torch.cuda.synchronize()
        for b in range(nb):
            L.set_box(b, S0[b])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for dt in dts:
            L.advance(dt)
        torch.cuda.synchronize()
        times["forward"].append(time.perf_counter() - t0)
        for b in range(nb):
            L.set_box(b, S0[b])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L.adjoint_forward(dts, every)
        torch.cuda.synchronize()
        times["forward_checkpointed"].append(time.perf_counter() - t0)
        L.seed_force_adjoint((1.0, 0.0, 0.0))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L.adjoint_backward()
        torch.cuda.synchronize()
        times["backward"].append(time.perf_counter() - t0)
    ms = {k: 1e3 * min(v) for k, v in times.items()}
    grad = [L.get_adjoint(b) for b in range(nb)]
    res.update({"functional": "x-force on the sphere at the final state",
                "steps": nsteps, "checkpoint_every": every, "force_at_T": force_T,
                "forward_ms": ms["forward"], "forward_checkpointed_ms": ms["forward_checkpointed"],
                "backward_ms": ms["backward"], "ratio": ms["backward"] / ms["forward"],
                "grad_norm_sod": float(np.sqrt(sum(float((g * g).sum()) for g in grad)))})
    # finite differences, smooth state
    cov = [geom.covered_mask(prob.boxes[b][:3], prob.boxes[b][3:]) for b in range(nb)]
    L.close()
    if args.eb_fd_dirs <= 0:
        return res
    L = level([workloads.sod_sphere_smooth_init(prob, b, geom) for b in range(nb)])
    S1 = [L.get_box(b) for b in range(nb)]
    fdts = L.run(args.eb_fd_steps)
    g = grad_of(S1, fdts)
    gnorm = float(np.sqrt(sum(float((x * x).sum()) for x in g)))
    rng = np.random.default_rng(problems.SEED_DIRECTIONS)
    eps, rows = 1e-6, []
    run_from(S1, fdts)
    l0 = float(L.eb_force()[0])
    for _ in range(args.eb_fd_dirs):
        v = []
        for s, c in zip(S1, cov):
            x = np.zeros_like(s)
            x[:6] = rng.standard_normal(x[:6].shape)
            x[:, c] = 0.0
            v.append(x)
        nv = np.sqrt(sum(float((x * x).sum()) for x in v))
        v = [x / nv for x in v]
        run_from([s + eps * x for s, x in zip(S1, v)], fdts)
        lp = float(L.eb_force()[0])
        run_from([s - eps * x for s, x in zip(S1, v)], fdts)
        lm = float(L.eb_force()[0])
        fd = (lp - lm) / (2 * eps)
        ad = sum(float((a * x).sum()) for a, x in zip(g, v))
        rows.append((fd, ad, abs((lp - l0) / eps - (l0 - lm) / eps)))
    res["fd_check"] = {
        "state": "smooth state on the same grid", "steps": args.eb_fd_steps,
        "directions": args.eb_fd_dirs, "eps": eps, "grad_norm": gnorm,
        "max_abs_err_over_grad_norm": max(abs(f - a) / gnorm for f, a, _ in rows),
        "max_abs_err_over_directional": max(abs(f - a) / abs(f) for f, a, _ in rows),
        "max_one_sided_slope_gap_over_directional": max(k / abs(f) for f, _, k in rows),
        "directions_fd_adjoint_slope_gap": [[f, a, k] for f, a, k in rows]}
    L.close()
    return res


def bench_mol(args) -> int:
    """The method-of-lines integrator (do_mol = 1) on the n^3 periodic
    smooth field, CFL 0.3 (the MOL limit, CAMR.cpp:175-177), one box on
    the staged engine.  The first step is checked bit for bit against the
    oracle (timed on the host cores as the CPU baseline); the dominant
    kernel is timed with HIP events on the level's stream during the timed
    steps."""
    import numpy as np
    import torch
    from camr_amd import _lib, workloads
    from camr_amd.hydro import Level
    from oracle import oracle as O
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    n, cfl = args.n, 0.3
    prob = workloads.smooth_periodic(n=n, ppm_type=0, cfl=cfl)
    prob.params["do_mol"] = 1
    U0 = workloads.smooth_init(prob, 0)
    L = Level(prob, engine=_lib.ENGINE_STAGED)
    L.set_box(0, U0)
    L.init_post()
    dt = L.initial_dt()
    L.advance(dt)
    L.sync()
    identical, cpu = None, None
    if not args.no_cpu_baseline:
        O.build()
        p = O.params(do_mol=1, cfl=cfl, bclo=prob.bclo, bchi=prob.bchi, domlo=prob.domlo,
                     domhi=prob.domhi, is_periodic=prob.is_periodic, dx=prob.dx)
        OL = O.Level(p, prob.boxes)
        OL.set_box(0, U0)
        OL.init_post()
        odt = OL.initial_dt()
        t0 = time.perf_counter()
        err = OL.advance(odt)
        el = time.perf_counter() - t0
        assert err == 0
        identical = bool(odt == dt and np.array_equal(OL.get_box(0), L.get_box(0)))
        OL.close()
        cpu = {"value": n ** 3 / el, "unit": "cell-updates/s",
               "cores": O.lib().ora_get_threads(), "kind": "port",
               "sample": f"the first step of the same {n}^3 MOL workload, {el:.1f} s wall"}
    for _ in range(max(args.warmup - 1, 0)):
        dt = L.new_dt(dt)
        L.advance(dt)
    L.sync()
    L.profile(True)
    dt = L.new_dt(dt)
    L.advance(dt)
    L.sync()
    rep = L.profile_report()
    kern = max((k for k in rep if k != "advance"), key=lambda k: rep[k]["ms"])
    L.profile(True, kern)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dt = L.new_dt(dt)
        L.advance(dt)
    L.sync()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    rep = L.profile_report()[kern]
    L.profile(False)
    cups = n ** 3 * args.steps / el
    avg_ms = rep["ms"] / rep["count"]
    # algorithmic bytes per launch of the dominant kernel
    if kern == "mol_faces":
        # per face: 9 primitives + 3 aux values of one cell read, 9 fluxes
        # + 2 Godunov values written
        algo = (NS_MOL_READ + NF_MOL_WRITE) * 8 * (n + 1) * n * n
    else:
        algo = ALGO_DOUBLES.get(kern, 2 * NVAR) * 8 * n ** 3
    achieved = algo / (avg_ms * 1e-3) / 1e9
    out = {"metric": f"FP64 cell-updates/s at {n}^3, method of lines (do_mol = 1)",
           "value": cups, "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic: seeded smooth periodic field (as the default workload)",
           "config": {"workload": f"{n}^3 periodic single-level MOL advance (two-stage RK, "
                                  f"characteristic MC slopes), CFL {cfl}, one box, staged engine",
                      "cells": n ** 3, "cfl": cfl},
           "identical_to_oracle_1_step": identical,
           "roofline": {"bound": "hbm", "kernel": kern, "achieved": achieved,
                        "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                        "traffic": None, "avg_launch_ms": avg_ms, "launches": rep["count"]},
           "cpu_baseline": cpu}
    print(json.dumps(out), flush=True)
    return 0


def bench_amr_sync(args) -> int:
    """Two-level subcycled step with flux registers (camr_amr_step): a
    periodic n0^3 coarse level and a fine block over the middle half of the
    domain in each direction at ratio 2, both in 32^3 boxes, PPM at CFL
    0.4, on the fused engine (--engine staged for the per-stage kernels).
    One coarse step = the coarse advance (CrseAdd), two fine steps
    with time-interpolated coarse-fine ghosts (FineAdd), reflux,
    average_down and the coarse clean.  The first step is checked bit for
    bit against the oracle (timed on the host cores as the CPU baseline)."""
    import numpy as np
    import torch
    from camr_amd import _lib, workloads
    from camr_amd.hydro import FluxRegister, Level, Problem, amr_step
    from oracle import oracle as O
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    n0, r, mg, cfl = args.amr_n0, 2, 32, 0.4
    coarse = workloads.smooth_periodic(n=n0, ppm_type=1, cfl=cfl, max_grid=mg)
    nf = n0 * r
    lo, hi = nf // 4, 3 * nf // 4
    fboxes = [(i, j, k, i + mg - 1, j + mg - 1, k + mg - 1) for k in range(lo, hi, mg)
              for j in range(lo, hi, mg) for i in range(lo, hi, mg)]
    fine = Problem(domlo=(0, 0, 0), domhi=(nf - 1,) * 3, dx=(1.0 / nf,) * 3, bclo=coarse.bclo,
                   bchi=coarse.bchi, is_periodic=coarse.is_periodic, boxes=fboxes,
                   params=dict(ppm_type=1, cfl=cfl))
    cells_c, cells_f = n0 ** 3, len(fboxes) * mg ** 3
    ups = cells_c + r * cells_f  # cell-updates per coarse step

    def gpu_pair():
        eng = _lib.ENGINE_FUSED if args.engine == "fused" else _lib.ENGINE_STAGED
        C = Level(coarse, engine=eng, worker_streams=args.workers)
        F = Level(fine, engine=eng, worker_streams=args.workers)
        for b in range(len(coarse.boxes)):
            C.set_box(b, workloads.smooth_init(coarse, b))
        for b in range(len(fboxes)):
            F.set_box(b, workloads.smooth_init(fine, b))
        R = FluxRegister(C, F, r)
        C.average_down(F, r)
        C.init_post()
        F.init_post()
        return C, F, R

    C, F, R = gpu_pair()
    dt = min(C.initial_dt(), r * F.initial_dt())
    amr_step(C, F, R, dt)
    identical, cpu = None, None
    if not args.no_cpu_baseline:
        O.build()

        def oparams(prob):
            return O.params(ppm_type=1, cfl=cfl, bclo=prob.bclo, bchi=prob.bchi,
                            domlo=prob.domlo, domhi=prob.domhi, is_periodic=prob.is_periodic,
                            dx=prob.dx)
        OC, OF = O.Level(oparams(coarse), coarse.boxes), O.Level(oparams(fine), fboxes)
        for b in range(len(coarse.boxes)):
            OC.set_box(b, workloads.smooth_init(coarse, b))
        for b in range(len(fboxes)):
            OF.set_box(b, workloads.smooth_init(fine, b))
        OR = O.FluxRegister(OC, OF, r)
        O.average_down(OC, OF, r)
        OC.init_post()
        OF.init_post()
        odt = min(OC.initial_dt(), r * OF.initial_dt())
        t0 = time.perf_counter()
        assert O.amr_step(OC, OF, OR, odt) == 0
        el = time.perf_counter() - t0
        identical = bool(odt == dt and
                         all(np.array_equal(C.get_box(b), OC.get_box(b)) for b in range(len(coarse.boxes))) and
                         all(np.array_equal(F.get_box(b), OF.get_box(b)) for b in range(len(fboxes))))
        OR.close()
        cpu = {"value": ups / el, "unit": "cell-updates/s", "cores": O.lib().ora_get_threads(),
               "kind": "port",
               "sample": f"the first composite step of the same hierarchy, {el:.1f} s wall"}
    for _ in range(args.warmup):
        dt = min(C.new_dt(dt), r * F.new_dt(dt / r))
        amr_step(C, F, R, dt)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dt = min(C.new_dt(dt), r * F.new_dt(dt / r))
        amr_step(C, F, R, dt)
    C.sync()
    F.sync()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    value = ups * args.steps / el
    out = {"metric": "FP64 cell-updates/s of the two-level subcycled step with reflux (PPM)",
           "unit": "cell-updates/s", "value": value, "higher_is_better": True, "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": el / args.steps * 1e3,
           "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic: seeded smooth periodic field on both levels",
           "config": {"workload": f"{n0}^3 periodic coarse level + fine block "
                                  f"[{lo},{hi - 1}]^3 at ratio 2, 32^3 boxes, PPM, CFL {cfl}; "
                                  "one step = coarse step + 2 fine steps + reflux + average_down",
                      "engine": args.engine, "coarse_cells": cells_c, "fine_cells": cells_f,
                      "cell_updates_per_step": ups, "boxes": [len(coarse.boxes), len(fboxes)]},
           "identical_to_oracle_1_step": identical,
           "roofline": {"bound": "hbm", "achieved": 160.0 * value / 1e9, "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": 160.0 * value / 8e12, "traffic": None,
                        "kernel": "whole composite step, 160 B per cell-update"},
           "cpu_baseline": cpu}
    R.close()
    print(json.dumps(out), flush=True)
    return 0


def adjoint_measure(n: int, nsteps: int, every: int, reps: int, ndirs: int,
                    field: str = "varp") -> dict:
    """Adjoint of `nsteps` PPM level steps at n^3 (checkpoints every
    `every`): wall time of the backward pass (recomputation from the
    checkpoints included) over that of the forward steps, best of `reps`;
    with `ndirs` > 0 the gradient is checked against central differences
    along seeded unit directions.  field "varp": the seeded smooth field
    with a smooth pressure variation; "c1": the same field at p = 1."""
    import numpy as np
    import torch
    from camr_amd import problems, workloads
    from camr_amd.hydro import Level
    prob = workloads.smooth_periodic(n=n, ppm_type=1, cfl=0.4)

    init = workloads.smooth_init_varp if field == "varp" else workloads.smooth_init

    def level_from(seed):
        L = Level(prob)
        L.set_box(0, init(prob, 0, seed))
        L.init_post()
        return L

    L = level_from(problems.SEED_FIELD)
    S0 = L.get_box(0)
    dts = L.run(nsteps)  # the primal dt sequence, frozen from here on
    T = level_from(problems.SEED_TARGET)
    for dt in dts:
        T.advance(dt)
    rho_t = T.get_box(0)[0]
    T.close()

    def forward(U):
        L.set_box(0, U)
        for dt in dts:
            L.advance(dt)
        return L.get_box(0)

    def gradient():
        L.set_box(0, S0)
        L.adjoint_forward(dts, every)
        g = np.zeros_like(S0)
        g[0] = L.get_box(0)[0] - rho_t
        L.set_adjoint(0, g)
        L.adjoint_backward()
        return L.get_adjoint(0)

    gradient()  # warm-up: allocates the checkpoints and adjoint scratch
    times = {"forward": [], "forward_checkpointed": [], "backward": []}
    for _ in range(max(1, reps)):
        L.set_box(0, S0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for dt in dts:
            L.advance(dt)
        torch.cuda.synchronize()
        times["forward"].append(time.perf_counter() - t0)
        L.set_box(0, S0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L.adjoint_forward(dts, every)
        torch.cuda.synchronize()
        times["forward_checkpointed"].append(time.perf_counter() - t0)
        g = np.zeros_like(S0)
        g[0] = L.get_box(0)[0] - rho_t
        L.set_adjoint(0, g)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        L.adjoint_backward()
        torch.cuda.synchronize()
        times["backward"].append(time.perf_counter() - t0)
    ms = {k: 1e3 * min(v) for k, v in times.items()}
    out = {"n": n, "steps": nsteps, "checkpoint_every": every,
           "forward_ms": ms["forward"], "forward_checkpointed_ms": ms["forward_checkpointed"],
           "backward_ms": ms["backward"], "ratio": ms["backward"] / ms["forward"],
           "backward_over_checkpointed_forward": ms["backward"] / ms["forward_checkpointed"]}
    if ndirs > 0:
        grad = L.get_adjoint(0)
        gnorm = float(np.sqrt((grad * grad).sum()))
        rng = np.random.default_rng(problems.SEED_DIRECTIONS)
        eps, fd_rows = 1e-6, []
        l0 = 0.5 * float(((forward(S0)[0] - rho_t) ** 2).sum())
        for _ in range(ndirs):
            v = rng.standard_normal(S0.shape)
            v /= np.sqrt((v * v).sum())
            lp = 0.5 * float(((forward(S0 + eps * v)[0] - rho_t) ** 2).sum())
            lm = 0.5 * float(((forward(S0 - eps * v)[0] - rho_t) ** 2).sum())
            fd = (lp - lm) / (2 * eps)
            # one-sided slopes: they differ where the step crosses a limiter
            # or upwind switch within +-eps, which the central difference
            # averages
            fr, fl = (lp - l0) / eps, (l0 - lm) / eps
            ad = float((grad * v).sum())
            fd_rows.append((fd, ad, abs(fr - fl)))
        out["fd_check"] = {
            "directions": ndirs, "eps": eps, "grad_norm": gnorm,
            "max_abs_err_over_grad_norm": max(abs(f - a) / gnorm for f, a, _ in fd_rows),
            "max_abs_err_over_directional": max(abs(f - a) / abs(f) for f, a, _ in fd_rows),
            "max_one_sided_slope_gap_over_directional": max(k / abs(f) for f, _, k in fd_rows),
            "directions_fd_adjoint_slope_gap": [[f, a, k] for f, a, k in fd_rows]}
    L.close()
    return out


def bench_adjoint(args) -> int:
    """Adjoint of `steps` PPM level steps at n^3 (checkpoints every 5):
    backward over forward wall time, and the gradient checked against
    central differences along seeded unit directions."""
    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    r = adjoint_measure(args.adj_n, args.steps, 5, args.adj_reps, args.adj_dirs,
                        args.adj_field)
    out = {"metric": "adjoint/forward cost ratio", "value": r["ratio"], "unit": "x",
           "higher_is_better": False, "n_gpus": 1, "steps": args.steps, "warmup": 1,
           "dtype": "f64", "data": "synthetic", "vs_baseline": None,
           "config": {"workload": f"adjoint of {args.steps} periodic PPM level steps at "
                                  f"{args.adj_n}^3, checkpoints every 5 steps, frozen dt "
                                  "sequence, L = 1/2 sum (rho(T) - rho_target)^2 over all "
                                  "initial conserved components; "
                                  + ("smooth field with a smooth pressure variation"
                                     if args.adj_field == "varp" else
                                     "the C1 field (p = 1)")}}
    out.update({k: v for k, v in r.items() if k != "ratio"})
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--engine", choices=["fused", "staged"], default="fused")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="c1 / eb: leave out the adjoint sub-measurement")
    ap.add_argument("--workload", choices=["c1", "amr", "adjoint", "sod", "eb", "mol", "amr_sync"],
                    default="c1",
                    help="c1: 256^3 PPM level step (default); amr: batched "
                         "AMR levels against per-box launches; adjoint: "
                         "backward/forward cost of the level-step adjoint; sod: "
                         "the 256x4x4 Sod tube, GPU beside the host oracle; eb: "
                         "Sod past a sphere, embedded boundary with FluxRedist; "
                         "mol: the n^3 periodic field with the method of lines; "
                         "amr_sync: two-level subcycled steps with flux registers")
    ap.add_argument("--amr-n0", type=int, default=128)
    ap.add_argument("--workers", type=int, default=0,
                    help="eb / amr_sync: worker streams of the staged engine (0: the "
                         "library's default of 4)")
    ap.add_argument("--adj-n", type=int, default=128)
    ap.add_argument("--eb-n", type=int, default=128,
                    help="eb workload: n of the (2n) x n x n grid")
    ap.add_argument("--no-eb-merge", action="store_true",
                    help="eb workload: run every regular box as itself in the batch (no "
                         "virtual boxes)")
    ap.add_argument("--eb-adj-steps", type=int, default=50,
                    help="eb workload: steps of the x-force adjoint (0: left out)")
    ap.add_argument("--eb-fd-steps", type=int, default=10,
                    help="eb workload: steps of the adjoint's finite-difference check")
    ap.add_argument("--eb-fd-dirs", type=int, default=4,
                    help="eb workload: directions of the adjoint's finite-difference check")
    ap.add_argument("--adj-dirs", type=int, default=16,
                    help="finite-difference directions of the adjoint check")
    ap.add_argument("--adj-reps", type=int, default=3)
    ap.add_argument("--adj-field", choices=["varp", "c1"], default="varp",
                    help="adjoint workload field: smooth pressure variation, or p = 1")
    ap.add_argument("--ctu", type=int, choices=[0, 1], default=None,
                    help="C1: force the CTU face kernels on (1) or off (0)")
    ap.add_argument("--ctu-planes", type=int, default=8)
    ap.add_argument("--no-tune", action="store_true",
                    help="default launch shapes instead of the autotuner's")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="after the timed steps, run one elementwise copy "
                         "(torch.neg) of MIB MiB of float64: a kernel of known "
                         "HBM traffic to scale hardware byte counters with")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one MI355X: runs on several GPUs are not "
              "supported (the level would shard by boxes, with one halo "
              "exchange and one dt all-reduce per step)", file=sys.stderr)
        return 2
    if args.workload == "amr":
        return bench_amr(args)
    if args.workload == "adjoint":
        return bench_adjoint(args)
    if args.workload == "eb":
        return bench_eb(args)
    if args.workload == "sod":
        return bench_sod(args)
    if args.workload == "mol":
        return bench_mol(args)
    if args.workload == "amr_sync":
        return bench_amr_sync(args)

    import numpy as np
    import torch
    from camr_amd import _lib, workloads
    from camr_amd.hydro import Level

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    n, cfl = args.n, 0.4
    prob = workloads.smooth_periodic(n=n, ppm_type=1, cfl=cfl)
    U0 = workloads.smooth_init(prob, 0)
    eng = _lib.ENGINE_FUSED if args.engine == "fused" else _lib.ENGINE_STAGED
    L = Level(prob, engine=eng)
    L.set_box(0, U0)
    L.init_post()
    tuning = None
    if args.engine == "fused" and not args.no_tune:
        # launch shapes from the autotuner (cached per device and box shape);
        # its timing steps advance this level before the warm-up
        from camr_amd import autotune
        tuning = autotune.tune(L, steps=2)
    if args.ctu is not None and args.engine == "fused":
        # transverse stages as the LDS-staged CTU face kernels (1) or the
        # staged transdo / transdd launches (0)
        L.ctu = (args.ctu, args.ctu_planes)

    # warm-up steps follow the reference dt sequence
    dt = L.initial_dt()
    L.advance(dt)
    for _ in range(max(args.warmup - 1, 0)):
        dt = L.new_dt(dt)
        L.advance(dt)
    L.sync()

    # find the dominant kernel over one profiled step
    L.profile(True)
    dt = L.new_dt(dt)
    L.advance(dt)
    L.sync()
    rep = L.profile_report()
    kern = max((k for k in rep if k != "advance"), key=lambda k: rep[k]["ms"])
    L.profile(True, kern)  # time only that kernel during the timed region

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dt = L.new_dt(dt)
        L.advance(dt)
    L.sync()
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    rep = L.profile_report()[kern]
    L.profile(False)
    if args.calib_copy > 0:
        # aligned: 32-byte vector loads per lane; offset by one element:
        # 8-byte loads per lane, the access pattern of the hydro kernels
        src = torch.ones(args.calib_copy << 17, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.neg(src, out=dst)
        torch.neg(src[1:], out=dst[1:])
        torch.cuda.synchronize()
        del src, dst

    cups = n ** 3 * args.steps / el
    avg_ms = rep["ms"] / rep["count"]
    # algorithmic FP64 operations of one launch of the dominant kernel: its
    # stage's count per cell-update over the launches of that stage per step
    flop_launch = ALGO_FLOP.get(kern, 0.0) * n ** 3 / LAUNCHES_PER_STEP.get(kern, 1)
    achieved = flop_launch / (avg_ms * 1e-3) / 1e12
    algo_bytes = ALGO_DOUBLES.get(kern, 2 * NVAR) * 8 * unit_count(kern, n)
    tiles = L.tiles if args.engine == "fused" else None
    ctu = L.ctu if args.engine == "fused" else None

    # parity at full size: the cpu baseline's steps from U0 on the GPU
    n_check = 6
    L.set_box(0, U0)
    L.init_post()
    gdts = []
    dt = None
    for s_ in range(n_check):
        dt = L.initial_dt() if s_ == 0 else L.new_dt(dt)
        L.advance(dt)
        gdts.append(dt)
    L.sync()
    U_gpu = L.get_box(0)
    L.close()
    del L
    torch.cuda.synchronize()

    out = {
        "metric": "FP64 cell-updates/s at 256³ PPM; % HBM peak; adjoint/forward cost ratio",
        "value": cups,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: seeded smooth periodic field (rho = 1 + 0.1 x 8 sines, "
                "k^-2 Gaussian velocity at rms Mach 0.5, p = 1)",
        "config": {"workload": f"{n}^3 periodic single-level Godunov PPM advance "
                               f"(CAMR_advance + computeNewDt), CFL {cfl}",
                   "cells": n ** 3, "ppm_type": 1, "cfl": cfl,
                   "engine": args.engine, "tuning": tuning, "tiles": tiles,
                   "ctu": list(ctu) if ctu else None},
        "roofline": {"bound": "fp64", "kernel": kern, "achieved": achieved,
                     "peak": FP64_PEAK_TFS, "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFS,
                     "traffic": MEASURED_TRAFFIC.get((args.engine, n, kern)),
                     "algorithmic_flop_per_launch": flop_launch,
                     "algorithmic_bytes_per_launch": algo_bytes,
                     "hbm_frac_of_algorithmic_bytes":
                         algo_bytes / (avg_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "avg_launch_ms": avg_ms, "launches": rep["count"]},
        "step_fp64": {"algorithmic_flop_per_cell_update": ALGO_FLOP_PER_CELL_UPDATE,
                      "achieved_TFs": cups * ALGO_FLOP_PER_CELL_UPDATE / 1e12,
                      "peak_TFs": FP64_PEAK_TFS,
                      "frac": cups * ALGO_FLOP_PER_CELL_UPDATE / 1e12 / FP64_PEAK_TFS,
                      "executed_flop_per_cell_update": FP64_EXECUTED_PER_CELL_UPDATE},
        "step_hbm": {"bytes_per_cell_update": {"NC6": 96, "NC10": 2 * NVAR * 8},
                     "frac_NC6": cups * 96 / 1e9 / HBM_PEAK_GBS,
                     "frac_NC10": cups * 2 * NVAR * 8 / 1e9 / HBM_PEAK_GBS},
    }
    if not args.no_adjoint:
        # the metric's third part: the adjoint of 20 PPM steps at 128^3 with
        # checkpoints every 5 steps, backward over forward wall time
        r = adjoint_measure(128, 20, 5, 2, 0)
        out["adjoint"] = {"ratio": r["ratio"], "forward_ms": r["forward_ms"],
                          "backward_ms": r["backward_ms"],
                          "workload": "adjoint of 20 periodic PPM level steps at 128^3, "
                                      "checkpoints every 5, L = 1/2 sum (rho(T) - rho_t)^2"}
    if not args.no_cpu_baseline:
        base, U_ora, odts = cpu_baseline(n, cfl, U0, warmup=1, nsteps=n_check - 1)
        out["cpu_baseline"] = base
        out["identical_to_oracle"] = bool(np.array_equal(U_gpu, U_ora) and
                                          np.array_equal(np.array(gdts), odts))
        out["oracle_check"] = (f"state and dt sequence after {n_check} steps from the same "
                               "initial state, GPU against the CPU oracle, bit for bit")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
