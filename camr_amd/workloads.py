"""The hydro-advance workloads (domain, BCs, boxes, parameters, initial data).

* ``sod_tube``: Sod shock tube, PLM, 256x4x4 cells, x in [0,1], outflow in x,
  periodic in y/z, CFL 0.5, 100 steps (Exec/Sod/inputs-x with the tube
  resolution and CFL of the CPU configuration).
* ``smooth_periodic``: 256^3 (or n^3) periodic PPM FP64 advance of the seeded
  smooth field of ``problems.SmoothField``, CFL 0.4, 20 steps.
* ``amr_levels``: the two levels of the batched AMR benchmark: level 0 is
  128^3 periodic cut into 32^3 boxes; level 1 (refinement ratio 2, a 256^3
  index space) covers the cells whose centre lies within 0.3 of the domain
  centre, clustered into boxes aligned to 8 cells with sides 8..64
  (``amr.cluster``, efficiency 0.8: about 300 boxes).  Both use the smooth
  field at their own cell centres; level-1 ghost cells outside level 1 take
  the smooth field as exterior data (standing for the coarse-level
  interpolation).
"""
from __future__ import annotations

from . import amr, problems
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


def smooth_init_varp(prob: Problem, b, seed=problems.SEED_FIELD, amp=0.1):
    """``smooth_init`` with the pressure 1 + amp * (two fixed sine modes of
    the cell centre) instead of the uniform p = 1.

    Gradient checks need it: with a uniform pressure every PPM pressure
    stencil is flat to rounding, so the edge limiter of every cell sits on
    its switching point and the step is non-differentiable at the scale of
    one ulp there; a finite difference then measures a secant across those
    kinks, not the derivative."""
    import numpy as np
    U = smooth_init(prob, b, seed)
    n = prob.domhi[0] - prob.domlo[0] + 1
    lo, hi = prob.boxes[b][:3], prob.boxes[b][3:]
    x, y, z = [(np.arange(lo[d], hi[d] + 1) + 0.5) / n for d in range(3)]
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    p = (1.0 + amp * np.sin(2 * np.pi * (X + 2 * Y) + 0.3)
         + 0.5 * amp * np.sin(2 * np.pi * (3 * Z - Y) + 1.1))
    P = problems
    ke = 0.5 * (U[P.UMX] ** 2 + U[P.UMY] ** 2 + U[P.UMZ] ** 2) / U[P.URHO]
    U[P.UEINT] = p / (P.GAMMA - 1.0)
    U[P.UEDEN] = U[P.UEINT] + ke
    return U


def refined_level(n=256, radius=0.3, blocking_factor=8, max_size=64, efficiency=0.8,
                  ppm_type=1, cfl=0.4):
    """Level covering the sphere of `radius` around the centre of an n^3
    periodic index space (the refined level of ``amr_levels``)."""
    tags = amr.sphere_tags(n, radius)
    boxes = amr.cluster(tags, blocking_factor, max_size, efficiency)
    dx = 1.0 / n
    return Problem(domlo=(0, 0, 0), domhi=(n - 1,) * 3, dx=(dx,) * 3,
                   bclo=(INTERIOR,) * 3, bchi=(INTERIOR,) * 3,
                   is_periodic=(1, 1, 1), boxes=boxes,
                   params=dict(ppm_type=ppm_type, cfl=cfl))


def exterior(prob: Problem, seed=problems.SEED_FIELD):
    """The smooth field over the whole domain of `prob` (exterior data)."""
    n = prob.domhi[0] - prob.domlo[0] + 1
    return smooth_field(seed).state(n, prob.domlo, prob.domhi)


def amr_levels(n0=128):
    """(level 0, level 1) of the batched AMR benchmark."""
    return (smooth_periodic(n=n0, ppm_type=1, cfl=0.4, max_grid=32),
            refined_level(n=2 * n0))


def sod_sphere(n=128, nx=None, max_grid=32, cfl=0.3, interface=0.3,
               center=(0.6, 0.5, 0.5), radius=0.15):
    """Sod shock tube past a sphere: (2n) x n x n cells on [0,2] x [0,1]^2
    (dx = 1/n), Outflow in x, SlipWall in y and z, PLM, the sphere of
    `radius` at `center` with the fluid outside it (Exec/SodPlusSphere/
    inputs:18-19, prob.H:15-49, eb_sphere.cpp:14-44), boxes of `max_grid`."""
    from ._lib import SLIPWALL
    nx = nx or 2 * n
    dx = 1.0 / n
    boxes = []
    for k in range(0, n, max_grid):
        for j in range(0, n, max_grid):
            for i in range(0, nx, max_grid):
                boxes.append((i, j, k, min(i + max_grid, nx) - 1, min(j + max_grid, n) - 1,
                              min(k + max_grid, n) - 1))
    prob = Problem(domlo=(0, 0, 0), domhi=(nx - 1, n - 1, n - 1), dx=(dx,) * 3,
                   bclo=(OUTFLOW, SLIPWALL, SLIPWALL), bchi=(OUTFLOW, SLIPWALL, SLIPWALL),
                   is_periodic=(0, 0, 0), boxes=boxes,
                   params=dict(ppm_type=0, cfl=cfl))
    prob.sphere = (tuple(center), float(radius))
    prob.interface = float(interface)
    return prob


def sod_sphere_geometry(prob):
    from . import eb
    c, r = prob.sphere
    return eb.Geometry(prob, eb.Sphere(c, r))


def sod_sphere_smooth_init(prob, b, geom=None, amp=0.2):
    """A smooth state on the Sod-past-a-sphere grid for gradient checks:
    density, pressure and the three velocities are fixed sine modes of the
    cell centre along oblique wave vectors (a flow of Mach ~0.4 in +x),
    covered cells zero.  The
    Sod states' discontinuity would put limiters on their switching points,
    where finite differences measure secants across kinks."""
    import numpy as np
    lo, hi = prob.boxes[b][:3], prob.boxes[b][3:]
    x, y, z = [prob.prob_lo[d] + (np.arange(lo[d], hi[d] + 1) + 0.5) * prob.dx[d]
               for d in range(3)]
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    tp = 2 * np.pi
    # every field varies along every axis: one constant along a direction
    # would put its slopes there exactly on the limiter's kink
    rho = (1.0 + amp * np.sin(tp * (0.5 * X + 0.3 * Y + 0.2 * Z) + 0.4)
           + 0.5 * amp * np.cos(tp * (0.35 * X + Z - Y) + 0.2))
    u = 0.4 + 0.5 * amp * np.sin(tp * (0.25 * X + Y + Z) + 0.3)
    v = 0.5 * amp * np.sin(tp * (0.5 * X + 0.4 * Y + Z) + 0.7)
    w = 0.5 * amp * np.cos(tp * (0.5 * X - Y + 0.3 * Z) + 0.1)
    p = 1.0 + amp * np.sin(tp * (0.5 * X + Y + 0.5 * Z) + 1.1)
    P = problems
    U = np.zeros((P.NVAR,) + rho.shape)
    U[P.URHO] = rho
    U[P.UMX], U[P.UMY], U[P.UMZ] = rho * u, rho * v, rho * w
    U[P.UEINT] = p / (P.GAMMA - 1.0)
    U[P.UEDEN] = U[P.UEINT] + 0.5 * rho * (u * u + v * v + w * w)
    U[P.UFS] = rho
    if geom is not None:
        cov = geom.covered_mask(lo, hi)
        U[:, cov] = 0.0
    return U


def sod_sphere_init(prob, b, geom=None):
    """Left state (rho, p) = (1, 1) for x < interface, right (0.125, 0.1),
    at rest; covered cells zero (EB_set_covered, CAMR.cpp:419-421)."""
    import numpy as np
    lo, hi = prob.boxes[b][:3], prob.boxes[b][3:]
    U = problems.sod(lo, hi, prob.dx, prob.prob_lo, interface=prob.interface)
    if geom is not None:
        cov = geom.covered_mask(lo, hi)
        U[:, cov] = 0.0
    return U
