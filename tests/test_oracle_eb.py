"""The embedded-boundary path of the oracle (hydro_umdrv_eb with
FluxRedist, Source/Hydro/Hydro_umdrv_eb.cpp:14-191) and the cut-cell
geometry it consumes (camr_amd/eb.py), pinned by what the mathematics fixes:

* geometry: the body volume converges to 4/3 pi r^3; covered cells have
  every face closed and regular cells every face open; connectivity is
  symmetric;
* free stream: a uniform gas at rest around the sphere stays at rest (the
  wall pressure flux balances the cut-face fluxes to rounding);
* conservation: with FluxRedist on one box, mass and total energy summed as
  vfrac * U over the valid cells change only by the domain-boundary fluxes
  (zero here: slip walls and a quiescent outflow);
* a level of the EB build with no cut cell reproduces the plain level bit
  for bit;
* the force on the body (the wall momentum flux of the cut cells times
  dx^2): zero for a uniform pressure (a closed surface), -b V_body for the
  pressure 1 + b x (Archimedes) to first order in dx, no y / z force by
  symmetry.

AMReX's EB2 geometry and its redistribution routine are not available, so
their numbers are parity unpinned; these checks pin the restatement.
"""
import itertools

import numpy as np
import pytest

from camr_amd import eb, problems, workloads

UNTILED = (1 << 20,) * 3


def _params(O, prob, **kw):
    return O.params(
        ppm_type=prob.value("ppm_type"), plm_iorder=prob.value("plm_iorder"),
        use_flattening=prob.value("use_flattening"), use_pslope=prob.value("use_pslope"),
        transverse_reset_density=prob.value("transverse_reset_density"),
        cfl=prob.value("cfl"), bclo=prob.bclo, bchi=prob.bchi, domlo=prob.domlo,
        domhi=prob.domhi, is_periodic=prob.is_periodic, dx=prob.dx, tile=UNTILED, **kw)


def _eb_level(O, prob, geom, init):
    L = O.Level(_params(O, prob), prob.boxes)
    L.set_eb(geom)
    for b in range(len(prob.boxes)):
        L.set_box(b, init(b))
    L.init_post()
    return L


def test_geometry_invariants():
    errs = []
    for n in (16, 32):
        prob = workloads.sod_sphere(n=n, max_grid=16)
        g = workloads.sod_sphere_geometry(prob)
        core = tuple(slice(eb.GEOM_GROW, -eb.GEOM_GROW) for _ in range(3))
        body = 2.0 - g.vfrac[core].sum() * prob.dx[0] ** 3
        errs.append(abs(body - 4.0 / 3.0 * np.pi * 0.15 ** 3) / (4.0 / 3.0 * np.pi * 0.15 ** 3))
        assert np.all((g.vfrac >= 0) & (g.vfrac <= 1))
        apx, apy, apz = g.ap
        for a in g.ap:
            assert np.all((a >= 0) & (a <= 1))
        cov = g.type == eb.COVERED
        reg = g.type == eb.REGULAR
        for t, want in ((cov, 0.0), (reg, 1.0)):
            k, j, i = np.nonzero(t)
            for a, (dk, dj, di) in ((apx, (0, 0, 1)), (apy, (0, 1, 0)), (apz, (1, 0, 0))):
                assert np.all(a[k, j, i] == want) and np.all(a[k + dk, j + dj, i + di] == want)
        # connectivity is symmetric
        f = g.flags
        K, J, I = f.shape
        for kk, jj, ii in itertools.product((-1, 0, 1), repeat=3):
            a = f[1:-1, 1:-1, 1:-1]
            b = f[1 + kk:K - 1 + kk, 1 + jj:J - 1 + jj, 1 + ii:I - 1 + ii]
            ca = (a >> eb.conn_bit(ii, jj, kk)) & 1
            cb = (b >> eb.conn_bit(-ii, -jj, -kk)) & 1
            live = ((a >> eb.TYPE_SHIFT) != eb.COVERED) & ((b >> eb.TYPE_SHIFT) != eb.COVERED)
            assert np.array_equal(ca[live], cb[live])
    # second-order convergence of the body volume
    assert errs[1] < errs[0] / 3.0


def test_eb_free_stream(oracle):
    prob = workloads.sod_sphere(n=16, max_grid=16)
    geom = workloads.sod_sphere_geometry(prob)

    def uniform(b):
        U = workloads.sod_sphere_init(prob, b, geom)
        lo, hi = prob.boxes[b][:3], prob.boxes[b][3:]
        live = ~geom.covered_mask(lo, hi)
        U[problems.URHO][live] = 1.0
        U[problems.UEINT][live] = 2.5
        U[problems.UEDEN][live] = 2.5
        U[problems.UFS][live] = 1.0
        return U
    L = _eb_level(oracle, prob, geom, uniform)
    assert 1 in L.kinds
    err, _ = L.run(3)
    assert err == 0
    for b in range(len(prob.boxes)):
        U = L.get_box(b)
        lo, hi = prob.boxes[b][:3], prob.boxes[b][3:]
        live = ~geom.covered_mask(lo, hi)
        assert np.abs(U[problems.URHO][live] - 1.0).max() <= 1e-12
        assert np.abs(U[problems.UEDEN][live] - 2.5).max() <= 1e-12
        for c in (problems.UMX, problems.UMY, problems.UMZ):
            assert np.abs(U[c][live]).max() <= 1e-12
        assert np.all(U[:, ~live] == 0.0)


def test_eb_conservation_one_box(oracle):
    # the interface starts 0.05 from the sphere: the shock reaches the cut
    # cells within two steps, long before any wave reaches the outflow faces
    prob = workloads.sod_sphere(n=16, max_grid=32, interface=0.4)
    assert len(prob.boxes) == 1
    geom = workloads.sod_sphere_geometry(prob)
    L = _eb_level(oracle, prob, geom, lambda b: workloads.sod_sphere_init(prob, b, geom))
    assert L.kinds == [1]
    lo, hi = prob.boxes[0][:3], prob.boxes[0][3:]
    core = geom._cells_of(lo, hi)
    vf = geom.vfrac[core]

    def totals():
        U = L.get_box(0)
        return np.array([(vf * U[c]).sum() for c in (problems.URHO, problems.UEDEN)])
    t0 = totals()
    err, dts = L.run(5)
    assert err == 0
    t1 = totals()
    assert np.all(np.abs(t1 - t0) <= 1e-12 * np.abs(t0)), (t0, t1)
    # the shock has reached the sphere: the state around it has changed
    U = L.get_box(0)
    cut = geom.type[core] == eb.SINGLE
    assert np.abs(U[problems.URHO][cut] - workloads.sod_sphere_init(prob, 0, geom)[0][cut]).max() > 1e-3


def test_eb_level_without_cut_cells_matches_plain(oracle):
    prob = workloads.smooth_periodic(n=16, ppm_type=0, cfl=0.4, max_grid=8)
    geom = eb.Geometry(prob, eb.Sphere((10.0, 10.0, 10.0), 0.1))
    assert np.all(geom.type == eb.REGULAR)
    init = lambda b: workloads.smooth_init_varp(prob, b)
    A = _eb_level(oracle, prob, geom, init)
    assert A.kinds == [0] * len(prob.boxes)
    B = oracle.Level(_params(oracle, prob), prob.boxes)
    for b in range(len(prob.boxes)):
        B.set_box(b, init(b))
    B.init_post()
    ea, da = A.run(3)
    eb_, db = B.run(3)
    assert ea == 0 and eb_ == 0
    assert np.array_equal(da, db)
    for b in range(len(prob.boxes)):
        assert np.array_equal(A.get_box(b), B.get_box(b))


def _pressure_field_level(oracle, n, pfun):
    """Oracle EB level of Sod past a sphere at rest with pressure pfun(X)."""
    import numpy as np
    prob = workloads.sod_sphere(n=n, max_grid=16)
    geom = workloads.sod_sphere_geometry(prob)
    L = oracle.Level(_params(oracle, prob), prob.boxes)
    L.set_eb(geom)
    for b, bb in enumerate(prob.boxes):
        lo, hi = bb[:3], bb[3:]
        x = [(np.arange(lo[d], hi[d] + 1) + 0.5) * prob.dx[d] for d in range(3)]
        Z, Y, X = np.meshgrid(x[2], x[1], x[0], indexing="ij")
        U = np.zeros((10,) + X.shape)
        U[0] = 1.0
        U[5] = pfun(X) / 0.4
        U[4] = U[5]
        U[7] = 1.0
        U[:, geom.covered_mask(lo, hi)] = 0.0
        L.set_box(b, np.ascontiguousarray(U))
    L.init_post()
    vbody = (1.0 - geom.vfrac[7:-7, 7:-7, 7:-7]).sum() * prob.dx[0] ** 3
    return L, vbody


def test_force_of_a_uniform_pressure_vanishes(oracle):
    """The body's surface is closed: the aperture differences of its cut
    cells sum to zero, so a uniform pressure exerts no net force."""
    L, vbody = _pressure_field_level(oracle, 32, lambda X: 1.0 + 0.0 * X)
    f = L.eb_force()
    assert abs(f).max() <= 1e-13 * 3.14159 * 0.15 ** 2


def test_force_of_a_linear_pressure_is_archimedes(oracle):
    """p = 1 + b x: the force is -b V_body (V from the cut-cell volume
    fractions), to first order in dx; y and z vanish by symmetry."""
    errs = []
    for n in (32, 64):
        L, vbody = _pressure_field_level(oracle, n, lambda X: 1.0 + 0.5 * X)
        f = L.eb_force()
        errs.append(abs(f[0] + 0.5 * vbody) / (0.5 * vbody))
        assert abs(f[1]) <= 1e-13 and abs(f[2]) <= 1e-13
    assert errs[0] < 2e-2 and errs[1] < 0.5 * errs[0], errs
