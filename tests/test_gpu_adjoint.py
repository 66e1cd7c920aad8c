"""Adjoint of the level step on the GPU against central finite differences.

Loss L = 1/2 sum (rho(T) - rho_target)^2 after n steps with the dt
sequence of the primal run frozen (as the adjoint assumes), rho_target the
same steps from a second seeded field.  The gradient with respect to every
initial conserved component is checked along seeded random unit directions
v: (L(S0 + eps v) - L(S0 - eps v)) / 2 eps with eps = 1e-6 must match
grad . v to 1e-6 of |grad| (the bound of |grad . v| for a unit v).  The
error is taken relative to |grad| rather than to |grad . v|: the central
difference itself carries rounding noise of a few 1e-8 at eps = 1e-6 (it
grows as 1/eps: a few 1e-7 at eps = 1e-7), which exceeds 1e-6 of a small
directional derivative, while larger steps cross limiter switches.

The initial fields vary in pressure (``workloads.smooth_init_varp``): with
a uniform pressure every PPM pressure stencil is flat to rounding, the
edge limiters sit on their switching points, and the step is not
differentiable there at the scale of one ulp.
"""
import numpy as np
import pytest

from camr_amd import _lib, problems, workloads
from camr_amd.hydro import Level

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _state(level, prob):
    return [level.get_box(b) for b in range(len(prob.boxes))]


def _setup(n, max_grid, nsteps, seed=problems.SEED_FIELD):
    prob = workloads.smooth_periodic(n=n, ppm_type=1, cfl=0.4, max_grid=max_grid)
    L = Level(prob)
    for b in range(len(prob.boxes)):
        L.set_box(b, workloads.smooth_init_varp(prob, b, seed))
    L.init_post()
    S0 = _state(L, prob)
    dts = L.run(nsteps)
    return prob, L, S0, dts


def _forward_loss(L, prob, U, dts, target):
    for b in range(len(prob.boxes)):
        L.set_box(b, U[b])
    for dt in dts:
        L.advance(dt)
    S = _state(L, prob)
    return 0.5 * sum(float(((s[0] - t[0]) ** 2).sum()) for s, t in zip(S, target))


def _target(prob, nsteps, dts):
    T = Level(prob)
    for b in range(len(prob.boxes)):
        T.set_box(b, workloads.smooth_init_varp(prob, b, problems.SEED_TARGET))
    T.init_post()
    for dt in dts:
        T.advance(dt)
    return _state(T, prob)


def _gradient(L, prob, S0, dts, target, every):
    for b in range(len(prob.boxes)):
        L.set_box(b, S0[b])
    L.adjoint_forward(dts, every)
    ST = _state(L, prob)
    for b in range(len(prob.boxes)):
        g = np.zeros_like(ST[b])
        g[0] = ST[b][0] - target[b][0]
        L.set_adjoint(b, g)
    L.adjoint_backward()
    return [L.get_adjoint(b) for b in range(len(prob.boxes))]


def _fd_check(L, prob, S0, dts, target, grad, ndir, eps=1e-6, tol=1e-6):
    rng = np.random.default_rng(problems.SEED_DIRECTIONS)
    gnorm = np.sqrt(sum(float((g * g).sum()) for g in grad))
    worst = 0.0
    for _ in range(ndir):
        v = [rng.standard_normal(s.shape) for s in S0]
        nv = np.sqrt(sum(float((x * x).sum()) for x in v))
        v = [x / nv for x in v]
        lp = _forward_loss(L, prob, [s + eps * x for s, x in zip(S0, v)], dts, target)
        lm = _forward_loss(L, prob, [s - eps * x for s, x in zip(S0, v)], dts, target)
        fd = (lp - lm) / (2 * eps)
        ad = sum(float((g * x).sum()) for g, x in zip(grad, v))
        rel = abs(fd - ad) / gnorm
        print(f"direction: fd {fd:.12e} adjoint {ad:.12e} |diff|/|grad| {rel:.2e} "
              f"|diff|/|fd| {abs(fd - ad) / abs(fd):.2e}", flush=True)
        worst = max(worst, rel)
    assert worst <= tol, worst


@pytest.mark.parametrize("max_grid,every", [(None, 2), (12, 3)])
def test_adjoint_matches_central_differences(max_grid, every):
    nsteps = 6
    prob, L, S0, dts = _setup(24, max_grid, nsteps)
    target = _target(prob, nsteps, dts)
    grad = _gradient(L, prob, S0, dts, target, every)
    for g in grad:
        assert np.all(np.isfinite(g))
        # rho e is reset from rho E by clean_state before it is used, and T
        # and the species never enter rho: their gradients vanish
        assert np.all(g[5] == 0.0) and np.all(g[6:] == 0.0)
    assert max(float(np.abs(g[0]).max()) for g in grad) > 0.0
    # the level is back at S0 after the backward pass
    for b in range(len(prob.boxes)):
        assert np.array_equal(L.get_box(b), S0[b])
    _fd_check(L, prob, S0, dts, target, grad, ndir=4)


def test_autograd_function_matches_direct_adjoint():
    from camr_amd.adjoint import LevelSteps
    nsteps = 4
    prob, L, S0, dts = _setup(16, None, nsteps)
    target = _target(prob, nsteps, dts)
    grad = _gradient(L, prob, S0, dts, target, every=2)
    U0 = torch.tensor(S0[0], dtype=torch.float64, requires_grad=True)
    UT = LevelSteps.apply(U0, L, dts, 2)
    loss = 0.5 * ((UT[0] - torch.as_tensor(target[0][0])) ** 2).sum()
    loss.backward()
    # the adjoint kernels accumulate shared contributions with FP64 atomics,
    # so two backward passes agree to rounding, not bit for bit
    g = U0.grad.numpy()
    scale = float(np.abs(grad[0]).max())
    assert np.abs(g - grad[0]).max() <= 1e-12 * scale


def test_adjoint_rejects_settings_out_of_scope():
    prob = workloads.sod_tube(nx=32)  # outflow walls, PLM
    L = Level(prob)
    L.set_box(0, workloads.sod_init(prob, 0))
    L.init_post()
    with pytest.raises(_lib.CamrError) as e:
        L.adjoint_forward([1e-3], 1)
    assert e.value.code == _lib.CAMR_ERR_ADJOINT if hasattr(_lib, "CAMR_ERR_ADJOINT") else e.value.code == 8
