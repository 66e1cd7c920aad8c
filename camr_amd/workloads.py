"""The hydro-advance workloads (domain, BCs, boxes, parameters, initial data).

* ``sod_tube``: Sod shock tube, PLM, 256x4x4 cells, x in [0,1], outflow in x,
  periodic in y/z, CFL 0.5, 100 steps (Exec/Sod/inputs-x with the tube
  resolution and CFL of the CPU configuration).
* ``smooth_periodic``: 256^3 (or n^3) periodic PPM FP64 advance of the seeded
  smooth field of ``problems.SmoothField``, CFL 0.4, 20 steps.
"""
from __future__ import annotations

from . import problems
from .hydro import Problem
from ._lib import INTERIOR, OUTFLOW


def sod_tube(nx=256, ny=4, nz=4, ppm_type=0, cfl=0.5, boxes=None):
    dx = 1.0 / nx
    prob = Problem(domlo=(0, 0, 0), domhi=(nx - 1, ny - 1, nz - 1),
                   dx=(dx, dx, dx), bclo=(OUTFLOW, INTERIOR, INTERIOR),
                   bchi=(OUTFLOW, INTERIOR, INTERIOR), is_periodic=(0, 1, 1),
                   boxes=boxes or [(0, 0, 0, nx - 1, ny - 1, nz - 1)],
                   params=dict(ppm_type=ppm_type, cfl=cfl))
    return prob


def sod_init(prob: Problem, b):
    lo, hi = prob.boxes[b][:3], prob.boxes[b][3:]
    return problems.sod(lo, hi, prob.dx, prob.prob_lo)


def smooth_periodic(n=256, ppm_type=1, cfl=0.4, max_grid=None):
    dx = 1.0 / n
    boxes = []
    mg = max_grid or n
    for k in range(0, n, mg):
        for j in range(0, n, mg):
            for i in range(0, n, mg):
                boxes.append((i, j, k, min(i + mg, n) - 1, min(j + mg, n) - 1,
                              min(k + mg, n) - 1))
    return Problem(domlo=(0, 0, 0), domhi=(n - 1,) * 3, dx=(dx,) * 3,
                   bclo=(INTERIOR,) * 3, bchi=(INTERIOR,) * 3,
                   is_periodic=(1, 1, 1), boxes=boxes,
                   params=dict(ppm_type=ppm_type, cfl=cfl))


_FIELDS = {}


def smooth_field(seed=problems.SEED_FIELD):
    if seed not in _FIELDS:
        _FIELDS[seed] = problems.SmoothField(seed=seed)
    return _FIELDS[seed]


def smooth_init(prob: Problem, b, seed=problems.SEED_FIELD):
    n = prob.domhi[0] - prob.domlo[0] + 1
    lo, hi = prob.boxes[b][:3], prob.boxes[b][3:]
    return smooth_field(seed).state(n, lo, hi)
