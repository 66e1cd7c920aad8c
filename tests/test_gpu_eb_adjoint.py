"""Adjoint of the embedded-boundary level step (C4's x-force) on the GPU
against central finite differences.

Sod-past-a-sphere grids (Outflow in x, SlipWall in y/z, PLM, FluxRedist)
with a smooth initial state (``workloads.sod_sphere_smooth_init``: the Sod
discontinuity would put limiters on their switching points), n steps with
the dt sequence of the primal run frozen.  Two losses: the x-force on the
sphere at the final state (``Level.eb_force``, seeded by
``seed_force_adjoint``) and a fixed random linear functional of the final
state.  Along seeded random unit directions v over the hydrodynamic
components of the non-covered cells, (L(S0 + eps v) - L(S0 - eps v)) / 2 eps
must match grad . v to 1e-6 of |grad|, as for the regular adjoint.
"""
import numpy as np
import pytest

from camr_amd import _lib, problems, workloads
from camr_amd.hydro import Level

pytestmark = pytest.mark.gpu

NH = 6  # rho, rho u, rho v, rho w, rho E, rho e


def _setup(n, max_grid, nsteps):
    prob = workloads.sod_sphere(n=n, max_grid=max_grid, cfl=0.3)
    geom = workloads.sod_sphere_geometry(prob)
    L = Level(prob, engine=_lib.ENGINE_STAGED)
    L.set_eb(geom)
    for b in range(len(prob.boxes)):
        L.set_box(b, workloads.sod_sphere_smooth_init(prob, b, geom))
    L.init_post()
    S0 = [L.get_box(b) for b in range(len(prob.boxes))]
    dts = L.run(nsteps)
    cov = [geom.covered_mask(prob.boxes[b][:3], prob.boxes[b][3:]) for b in range(len(prob.boxes))]
    return prob, L, S0, dts, cov


def _run(L, prob, U, dts):
    for b in range(len(prob.boxes)):
        L.set_box(b, U[b])
    for dt in dts:
        L.advance(dt)


def _grad(L, prob, S0, dts, every, seed_fn):
    for b in range(len(prob.boxes)):
        L.set_box(b, S0[b])
    L.adjoint_forward(dts, every)
    seed_fn(L)
    L.adjoint_backward()
    return [L.get_adjoint(b) for b in range(len(prob.boxes))]


def _fd(L, prob, S0, dts, cov, grad, loss, ndir, eps=1e-6, tol=1e-6):
    rng = np.random.default_rng(problems.SEED_DIRECTIONS)
    gnorm = np.sqrt(sum(float((g * g).sum()) for g in grad))
    assert gnorm > 0.0
    worst = 0.0
    for _ in range(ndir):
        v = []
        for s, c in zip(S0, cov):
            x = np.zeros_like(s)
            x[:NH] = rng.standard_normal(x[:NH].shape)
            x[:, c] = 0.0
            v.append(x)
        nv = np.sqrt(sum(float((x * x).sum()) for x in v))
        v = [x / nv for x in v]
        _run(L, prob, [s + eps * x for s, x in zip(S0, v)], dts)
        lp = loss(L)
        _run(L, prob, [s - eps * x for s, x in zip(S0, v)], dts)
        lm = loss(L)
        fd = (lp - lm) / (2 * eps)
        ad = sum(float((g * x).sum()) for g, x in zip(grad, v))
        rel = abs(fd - ad) / gnorm
        print(f"direction: fd {fd:.12e} adjoint {ad:.12e} |diff|/|grad| {rel:.2e}", flush=True)
        worst = max(worst, rel)
    assert worst <= tol, worst


@pytest.mark.parametrize("n,max_grid,every", [(16, 16, 2), (16, 8, 3)])
def test_eb_force_adjoint_matches_central_differences(n, max_grid, every):
    print(f"EB adjoint n={n} max_grid={max_grid} every={every}", flush=True)
    nsteps = 4
    prob, L, S0, dts, cov = _setup(n, max_grid, nsteps)
    assert 1 in L.kinds
    grad = _grad(L, prob, S0, dts, every, lambda lv: lv.seed_force_adjoint((1.0, 0.0, 0.0)))
    for g, c in zip(grad, cov):
        assert np.all(np.isfinite(g))
        # rho e is reset from rho E before it is used; T and the species are
        # passive; covered cells do not enter the step
        assert np.all(g[5] == 0.0) and np.all(g[6:] == 0.0)
        assert np.all(g[:, c] == 0.0)
    for b in range(len(prob.boxes)):
        assert np.array_equal(L.get_box(b), S0[b])
    _fd(L, prob, S0, dts, cov, grad, lambda lv: float(lv.eb_force()[0]), ndir=4)


def test_eb_adjoint_of_a_linear_functional():
    nsteps = 3
    prob, L, S0, dts, cov = _setup(16, 8, nsteps)
    rng = np.random.default_rng(problems.SEED_TARGET)
    W = []
    for s, c in zip(S0, cov):
        w = np.zeros_like(s)
        w[:5] = rng.standard_normal(w[:5].shape)
        w[:, c] = 0.0
        W.append(w)

    def seed(lv):
        for b in range(len(prob.boxes)):
            lv.set_adjoint(b, W[b])

    def loss(lv):
        return sum(float((w * lv.get_box(b)).sum()) for b, w in enumerate(W))

    grad = _grad(L, prob, S0, dts, 2, seed)
    for g in grad:
        assert np.all(np.isfinite(g))
    _fd(L, prob, S0, dts, cov, grad, loss, ndir=3)


def test_eb_force_seed_needs_an_adjoint_run():
    prob = workloads.sod_sphere(n=16, max_grid=16)
    geom = workloads.sod_sphere_geometry(prob)
    L = Level(prob, engine=_lib.ENGINE_STAGED)
    L.set_eb(geom)
    with pytest.raises(_lib.CamrError):
        L.seed_force_adjoint((1.0, 0.0, 0.0))


def test_eb_force_autograd_function_matches_direct_adjoint():
    torch = pytest.importorskip("torch")
    from camr_amd.adjoint import EbForce, LevelSteps
    nsteps = 3
    prob, L, S0, dts, cov = _setup(16, 8, nsteps)
    grad = _grad(L, prob, S0, dts, 2, lambda lv: lv.seed_force_adjoint((1.0, 0.0, 0.0)))
    U0 = torch.tensor(np.stack(S0), dtype=torch.float64, requires_grad=True)
    F = EbForce.apply(U0, L, dts, 2)
    F.backward()
    # atomics: two backward passes agree to rounding, not bit for bit
    g = U0.grad.numpy()
    scale = max(float(np.abs(x).max()) for x in grad)
    assert np.abs(g - np.stack(grad)).max() <= 1e-12 * scale
    # the state map of the same level: a linear loss through LevelSteps
    U1 = torch.tensor(np.stack(S0), dtype=torch.float64, requires_grad=True)
    UT = LevelSteps.apply(U1, L, dts, 2)
    UT[:, 0].sum().backward()
    assert np.all(np.isfinite(U1.grad.numpy())) and float(U1.grad.abs().max()) > 0.0


def test_eb_adjoint_over_virtual_boxes_matches_box_by_box():
    """The recompute and the adjoint stages over virtual boxes (blocks of
    regular boxes, their Sborder adjoint added back into the boxes') give
    the gradient of the same level run box by box, to rounding (the order
    of the atomic adds differs)."""
    print("EB adjoint: 64x32x32, 8^3 boxes, merged vs box by box", flush=True)
    nsteps = 3
    prob = workloads.sod_sphere(n=32, max_grid=8, cfl=0.3)
    geom = workloads.sod_sphere_geometry(prob)
    grads, dtss = [], []
    for merge in (True, False):
        L = Level(prob, engine=_lib.ENGINE_STAGED)
        L.set_eb(geom, merge=merge)
        assert (L.merge_info()[0] > 0) == merge
        for b in range(len(prob.boxes)):
            L.set_box(b, workloads.sod_sphere_smooth_init(prob, b, geom))
        L.init_post()
        S0 = [L.get_box(b) for b in range(len(prob.boxes))]
        dts = L.run(nsteps)
        dtss.append(np.asarray(dts))
        grads.append(_grad(L, prob, S0, dts, 2,
                           lambda lv: lv.seed_force_adjoint((1.0, 0.0, 0.0))))
        L.close()
    assert np.array_equal(dtss[0], dtss[1])
    scale = max(float(np.abs(g).max()) for g in grads[1])
    assert scale > 0.0
    worst = max(float(np.abs(a - b).max()) for a, b in zip(grads[0], grads[1]))
    print(f"max |merged - box by box| / max |grad| = {worst / scale:.2e}", flush=True)
    assert worst <= 1e-12 * scale
