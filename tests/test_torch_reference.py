"""Reference gradients from a float64 PyTorch-autograd restatement of the
level step (tests/torch_step.py).

The restatement is first pinned against the C oracle (same initial state,
same dts: agreement to rounding after two steps); its autograd gradient of
L = 1/2 sum (rho(T) - rho_target)^2 is then compared entry by entry with the
GPU adjoint on a 12^3 periodic box over three steps (checkpoints every 2).
Unlike a finite difference this reference has no truncation or rounding
noise of its own: both sides differentiate the branches the forward takes,
so they agree to rounding wherever the primal runs agree.
"""
import numpy as np
import pytest

from camr_amd import problems, workloads

torch = pytest.importorskip("torch")

N = 12


def _prob():
    return workloads.smooth_periodic(n=N, ppm_type=1, cfl=0.4)


def _torch_forward(S0, dts, prob):
    import torch_step as T
    hc = T.default_consts()
    S = S0
    for dt in dts:
        S = T.level_step(S, float(dt), prob.dx, hc)
    return S


def test_torch_step_matches_oracle(oracle):
    from test_gpu_parity import _oracle_params
    O = oracle
    prob = _prob()
    L = O.Level(_oracle_params(O, prob), prob.boxes)
    L.set_box(0, workloads.smooth_init_varp(prob, 0))
    L.init_post()
    S0 = L.get_box(0)
    err, dts = L.run(2)
    assert err == 0
    ref = L.get_box(0)
    L.close()
    S = _torch_forward(torch.tensor(S0[:6]), dts, prob).numpy()
    for c in range(6):
        rel = np.abs(S[c] - ref[c]).max() / np.abs(ref[c]).max()
        assert rel <= 1e-14, (c, rel)


def _gather(prob, boxes_data):
    """Boxes of a level assembled into the (NVAR, n, n, n) domain array."""
    out = np.zeros((boxes_data[0].shape[0], N, N, N))
    for bb, a in zip(prob.boxes, boxes_data):
        lo, hi = bb[:3], bb[3:]
        out[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = a
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("max_grid,nsteps,every", [(None, 3, 2), (6, 6, 4)])
def test_adjoint_matches_torch_autograd(max_grid, nsteps, every):
    """One 12^3 box over 3 steps, and the same domain in 8 boxes of 6^3
    over 6 steps (checkpoints every 4: a partial last block): the GPU
    adjoint of every box against autograd through the one-box restatement
    (a level's result does not depend on its box decomposition)."""
    from camr_amd.hydro import Level
    prob = workloads.smooth_periodic(n=N, ppm_type=1, cfl=0.4, max_grid=max_grid or N)
    nb = len(prob.boxes)
    assert nb == (8 if max_grid else 1)
    L = Level(prob)
    for b in range(nb):
        L.set_box(b, workloads.smooth_init_varp(prob, b))
    L.init_post()
    S0b = [L.get_box(b) for b in range(nb)]
    dts = L.run(nsteps)
    T = Level(prob)
    for b in range(nb):
        T.set_box(b, workloads.smooth_init_varp(prob, b, problems.SEED_TARGET))
    T.init_post()
    for dt in dts:
        T.advance(dt)
    rho_t = [T.get_box(b)[0] for b in range(nb)]
    # GPU adjoint
    for b in range(nb):
        L.set_box(b, S0b[b])
    L.adjoint_forward(dts, every)
    for b in range(nb):
        g = np.zeros_like(S0b[b])
        g[0] = L.get_box(b)[0] - rho_t[b]
        L.set_adjoint(b, g)
    L.adjoint_backward()
    gpu = _gather(prob, [L.get_adjoint(b) for b in range(nb)])
    # autograd through the one-box restatement on the whole domain
    S0 = _gather(prob, S0b)
    U0 = torch.tensor(S0[:6], requires_grad=True)
    ST = _torch_forward(U0, dts, _prob())
    rt = _gather(prob, [r[None] for r in rho_t])[0]
    loss = 0.5 * ((ST[0] - torch.as_tensor(rt)) ** 2).sum()
    loss.backward()
    ref = U0.grad.numpy()
    scale = float(np.abs(ref).max())
    diff = float(np.abs(gpu[:6] - ref).max())
    print(f"{nb} boxes, {nsteps} steps: max |adjoint - autograd| = {diff:.3e} "
          f"(max |grad| = {scale:.3e})", flush=True)
    assert diff <= 1e-12 * scale
    assert np.all(gpu[6:] == 0.0)
