"""Forward-mode tangent of the rollout (cd_tangent_step) against the adjoint
(cd_adjoint_step): the dot-product test

    <dL/dX, J t>  =  <J^T dL/dX, t>

for random tangents t of the per-agent parameters (tau_0, mu, kappa,
damping, target_velocity) and of the initial position and velocity, where
J is the derivative of the whole trajectory X.  Both sides are exact
derivatives of the same floating-point rollout computed by different
sweeps (one forward, one reverse), so they agree to rounding at any horizon,
however large the gradient grows; an error in either sweep (a dropped term,
a transposed Jacobian, a wrong sign) breaks the identity.  The tangent is
also checked against central differences of the GPU forward on a short
contact-free window, and the identity is checked on the 4000-agent
differentiable-rollout workload at 10, 25, 50 and 100 steps."""
import numpy as np
import pytest
import torch

from crowddynamics_amd import examples as E
from crowddynamics_amd.adjoint import Rollout, rollout_positions
from tests import adjoint_cases as AC

pytestmark = pytest.mark.gpu

DT = 0.005
CELL = 3.6
#: typical size of each parameter (tau_0, mu, kappa, damping, v0): tangents
#: are drawn relative to it so that every parameter moves the trajectory
PSCALE = np.array([3.0, 1.2e5, 4e4, 500.0, 1.15])


def say(*a):
    print(*a, flush=True)


def _params(arr, device):
    ps = [torch.tensor(3.0), torch.tensor(1.2e5), torch.tensor(4e4), torch.tensor(500.0),
          torch.as_tensor(np.array(arr['target_velocity'], dtype=np.float64))]
    return [p.to(device=device, dtype=torch.float64) for p in ps]


def _targets(R, arr, seed=0):
    rng = np.random.default_rng(seed)
    p = _params(arr, 'cuda')
    p[0] = p[0] * 1.05
    p[2] = p[2] * 1.05
    p[4] = p[4] * torch.as_tensor(1.0 + 0.05 * rng.standard_normal(len(arr)), device='cuda')
    with torch.no_grad():
        return rollout_positions(R, *p)


def _loss(X, Xhat):
    return ((X[1:] - Xhat[1:]) ** 2).sum(-1).mean()


def dot_product_test(R, arr, seed=0, state=True):
    """(tangent side, adjoint side, scale) of the identity for one random
    tangent; `scale` bounds the terms summed on the tangent side."""
    n = len(arr)
    Xhat = _targets(R, arr, seed)
    R.set_parameters(*_params(arr, 'cuda'))
    X = R.forward().clone().requires_grad_(True)
    _loss(X, Xhat).backward()
    gX = X.grad.detach()
    g, lx0, lv0 = R.backward(gX)
    rng = np.random.default_rng(seed + 100)
    dp = torch.as_tensor(rng.standard_normal((n, 5)) * (1e-3 * PSCALE), device='cuda')
    dx0 = torch.as_tensor(rng.standard_normal((n, 2)) * 1e-3 if state else np.zeros((n, 2)),
                          device='cuda')
    dv0 = torch.as_tensor(rng.standard_normal((n, 2)) * 1e-3 if state else np.zeros((n, 2)),
                          device='cuda')
    dX = R.tangent(dp, dx0, dv0)
    terms = gX * dX
    lhs = float(terms.sum())
    rhs = float((g * dp).sum() + (lx0 * dx0).sum() + (lv0 * dv0).sum())
    scale = float(terms.abs().sum())
    return lhs, rhs, scale


@pytest.mark.parametrize('case', ['sparse', 'dimers', 'packed'])
def test_tangent_adjoint_dot_product(case):
    arr = {'sparse': lambda: AC.sparse(40, seed=3),
           'dimers': lambda: AC.dimers(15, seed=2),
           'packed': lambda: AC.packed(36, seed=2)}[case]()
    R = Rollout(arr, steps=30, dt=DT, cell_size=CELL, checkpoint=7)
    lhs, rhs, scale = dot_product_test(R, arr, seed=1)
    say(case, 'tangent', lhs, 'adjoint', rhs, 'scale', scale)
    assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), scale), (lhs, rhs, scale)


def test_tangent_matches_central_fd():
    """The tangent's loss derivative <dL/dX, dX> along tau_0 and along one
    agent's desired speed equals central differences of the loss computed
    by the GPU forward (h = 1e-6, and the Richardson extrapolation from h
    and h / 2), on a contact-free window of a sparse room."""
    arr = AC.sparse(300, seed=5, density=0.5)
    n = len(arr)
    R = Rollout(arr, steps=30, dt=DT, cell_size=CELL, checkpoint=25)
    Xhat = _targets(R, arr, seed=1)
    R.set_parameters(*_params(arr, 'cuda'))
    X = R.forward().clone().requires_grad_(True)
    L = _loss(X, Xhat)
    L.backward()
    gX = X.grad.detach()
    L = float(L.detach())
    for k, index in ((0, None), (4, 0), (4, 17), (4, 299)):
        dp = torch.zeros((n, 5), dtype=torch.float64, device='cuda')
        if index is None:
            dp[:, k] = 1.0
        else:
            dp[index, k] = 1.0
        R.set_parameters(*_params(arr, 'cuda'))
        dL = float((gX * R.tangent(dp)).sum())

        def loss_at(delta):
            p = _params(arr, 'cuda')
            if index is None:
                p[k] = p[k] + delta
            else:
                p[k] = p[k].clone()
                p[k][index] += delta
            with torch.no_grad():
                return float(_loss(rollout_positions(R, *p), Xhat))

        h = 1e-6
        f1 = (loss_at(h) - loss_at(-h)) / (2 * h)
        f2 = (loss_at(h / 2) - loss_at(-h / 2)) / h
        rich = (4.0 * f2 - f1) / 3.0
        floor = 100 * np.finfo(np.float64).eps * abs(L) / (h / 2)
        say('param', k, index, 'tangent', dL, 'fd', f1, 'richardson', rich)
        assert abs(dL - f1) <= 1e-4 * abs(f1) + floor
        assert abs(dL - rich) <= 3e-5 * abs(rich) + floor


@pytest.mark.parametrize('steps', [10, 25, 50, 100])
def test_rollout4k_dot_product(steps):
    """The 4000-agent differentiable-rollout workload (fixed dt 0.005 s,
    checkpoint 25): the identity holds where finite differences stop
    converging, so the gradient growth of long windows is the dynamics'
    own, not an error of the adjoint."""
    scn = E.rollout(n=4000, steps=steps)
    arr = scn.agents
    R = Rollout(arr, steps=steps, dt=scn.dt_max, cell_size=scn.cell_size, checkpoint=25,
                bounds=scn.bounds)
    lhs, rhs, scale = dot_product_test(R, arr, seed=steps, state=True)
    say('rollout4k steps', steps, 'tangent', lhs, 'adjoint', rhs, 'scale', scale,
        'rel', abs(lhs - rhs) / max(abs(lhs), scale))
    assert abs(lhs - rhs) <= 1e-8 * max(abs(lhs), scale), (lhs, rhs, scale)
