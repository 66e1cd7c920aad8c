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


@pytest.mark.gpu
def test_adjoint_matches_torch_autograd():
    from camr_amd.hydro import Level
    prob = _prob()
    nsteps = 3
    L = Level(prob)
    L.set_box(0, workloads.smooth_init_varp(prob, 0))
    L.init_post()
    S0 = L.get_box(0)
    dts = L.run(nsteps)
    T = Level(prob)
    T.set_box(0, workloads.smooth_init_varp(prob, 0, problems.SEED_TARGET))
    T.init_post()
    for dt in dts:
        T.advance(dt)
    rho_t = T.get_box(0)[0]
    # GPU adjoint
    L.set_box(0, S0)
    L.adjoint_forward(dts, 2)
    g = np.zeros_like(S0)
    g[0] = L.get_box(0)[0] - rho_t
    L.set_adjoint(0, g)
    L.adjoint_backward()
    gpu = L.get_adjoint(0)
    # autograd through the restatement
    U0 = torch.tensor(S0[:6], requires_grad=True)
    ST = _torch_forward(U0, dts, prob)
    loss = 0.5 * ((ST[0] - torch.as_tensor(rho_t)) ** 2).sum()
    loss.backward()
    ref = U0.grad.numpy()
    scale = float(np.abs(ref).max())
    diff = float(np.abs(gpu[:6] - ref).max())
    print(f"max |adjoint - autograd| = {diff:.3e} (max |grad| = {scale:.3e})", flush=True)
    assert diff <= 1e-12 * scale
    assert np.all(gpu[6:] == 0.0)
