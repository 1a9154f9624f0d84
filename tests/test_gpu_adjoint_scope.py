"""The adjoint and the tangent over the whole standard tree: wall contacts
(core/interactions.py:107-141), the navigation lookup of each step
(core/steering/navigation.py:59-78, the bilinear potential gradient and the
nearest node) and the adaptive dt (core/integrator.py:29-97), with
fluctuation on.

Checks, on a walled room with exits and a distance-to-exit potential, some
agents pressed into the walls and one fast agent (so that dt lies between
its clamps and follows the fastest agent):
* the dot-product test <dL/dX, J t> = <J^T dL/dX, t> for random tangents of
  every parameter and of the initial state: the tangent and the adjoint are
  each other's transpose to rounding;
* the tangent against central differences of the GPU forward along tau_0,
  mu (through the wall contacts), the target_velocity of the agent with the
  largest one (through dt) and another agent's (through the navigation
  direction it adjusts to), on a short window."""
import numpy as np
import pytest
import torch

from crowddynamics_amd import _lib
from crowddynamics_amd import examples as E
from crowddynamics_amd.adjoint import Rollout, rollout_positions

pytestmark = pytest.mark.gpu

PSCALE = np.array([3.0, 1.2e5, 4e4, 500.0, 1.15])


def say(*a):
    print(*a, flush=True)


def walled_room(n=160, seed=4, nav_mode=1):
    """A 16 x 8 m room with four exits, agents at least 1.5 m from every
    exit, eight of them pressed 2 cm into the lower and upper walls, and one
    moving at 3 m/s (dt = 1.1 v0_max dt_max / v_max is then below dt_max)."""
    scn = E.room(n=n, seed=seed, width=16.0, height=8.0, nav_mode=nav_mode)
    arr = scn.agents.copy()
    p = arr['position']
    exits = np.array([[8.0, 0.0], [16.0, 4.0], [8.0, 8.0], [0.0, 4.0]])
    far = np.min(np.hypot(p[:, None, 0] - exits[None, :, 0], p[:, None, 1] - exits[None, :, 1]),
                 axis=1) > 1.5
    arr = arr[far].copy()
    xs = [2.0, 4.0, 11.0, 13.5]
    for k in range(8):
        r = arr['radius'][k]
        x = xs[k % 4] + 0.3 * (k // 4)
        y = r - 0.02 if k < 4 else 8.0 - r + 0.02
        arr['position'][k] = (x, y)
        arr['velocity'][k] = (0.2, 0.0)
    arr['velocity'][20] = (3.0, 0.5)
    return scn, arr


def _params(arr):
    ps = [torch.tensor(3.0), torch.tensor(1.2e5), torch.tensor(4e4), torch.tensor(500.0),
          torch.as_tensor(np.array(arr['target_velocity'], dtype=np.float64))]
    return [q.to(device='cuda', dtype=torch.float64) for q in ps]


def _rollout(scn, arr, steps, checkpoint=7):
    return Rollout(arr, steps=steps, cell_size=3.6, checkpoint=checkpoint, seed=3,
                   fluctuation=True, segments=scn.segments, nav=scn.nav, dt_min=0.001,
                   dt_max=0.01, bounds=scn.bounds)


def _targets(R, arr):
    p = _params(arr)
    p[0] = p[0] * 1.05
    p[1] = p[1] * 1.05
    p[4] = p[4] * 1.03
    with torch.no_grad():
        return rollout_positions(R, *p)


def _loss(X, Xhat):
    return ((X[1:] - Xhat[1:]) ** 2).sum(-1).mean()


def _status(R):
    return R.st.ws.state().status


@pytest.mark.parametrize('nav_mode', [1, 0])
def test_dot_product_walls_navigation_adaptive_dt(nav_mode):
    scn, arr = walled_room(nav_mode=nav_mode)
    n = len(arr)
    R = _rollout(scn, arr, steps=24)
    Xhat = _targets(R, arr)
    R.set_parameters(*_params(arr))
    X = R.forward().clone().requires_grad_(True)
    _loss(X, Xhat).backward()
    gX = X.grad.detach()
    g, lx0, lv0 = R.backward(gX)
    assert not (_status(R) & _lib.CD_STATUS_NAV_KEPT), 'an agent left the navigation map'
    rng = np.random.default_rng(11)
    dp = torch.as_tensor(rng.standard_normal((n, 5)) * (1e-3 * PSCALE), device='cuda')
    dx0 = torch.as_tensor(rng.standard_normal((n, 2)) * 1e-3, device='cuda')
    dv0 = torch.as_tensor(rng.standard_normal((n, 2)) * 1e-3, device='cuda')
    dX = R.tangent(dp, dx0, dv0)
    terms = gX * dX
    lhs = float(terms.sum())
    rhs = float((g * dp).sum() + (lx0 * dx0).sum() + (lv0 * dv0).sum())
    scale = float(terms.abs().sum())
    say('nav mode', nav_mode, 'tangent', lhs, 'adjoint', rhs, 'scale', scale)
    assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), scale), (lhs, rhs, scale)
    # the wall contacts and dt reach the gradient
    assert float(g[:8, 1].abs().sum()) > 0.0, 'no wall-contact gradient in mu'
    k = int(np.argmax(arr['target_velocity']))
    assert float(g[k, 4].abs()) > 0.0


def test_tangent_matches_central_fd_walls_navigation_adaptive_dt():
    scn, arr = walled_room()
    n = len(arr)
    R = _rollout(scn, arr, steps=5, checkpoint=4)
    Xhat = _targets(R, arr)
    R.set_parameters(*_params(arr))
    X = R.forward().clone().requires_grad_(True)
    L = _loss(X, Xhat)
    L.backward()
    gX = X.grad.detach()
    L = float(L.detach())
    vmax_agent = int(np.argmax(arr['target_velocity']))
    cases = ((0, None, 1e-6), (1, None, 1e-6 * 1.2e5), (4, vmax_agent, 1e-6), (4, 30, 1e-6))
    for k, index, h in cases:
        dp = torch.zeros((n, 5), dtype=torch.float64, device='cuda')
        if index is None:
            dp[:, k] = 1.0
        else:
            dp[index, k] = 1.0
        R.set_parameters(*_params(arr))
        dL = float((gX * R.tangent(dp)).sum())

        def loss_at(delta):
            p = _params(arr)
            if index is None:
                p[k] = p[k] + delta
            else:
                p[k] = p[k].clone()
                p[k][index] += delta
            with torch.no_grad():
                return float(_loss(rollout_positions(R, *p), Xhat))

        f1 = (loss_at(h) - loss_at(-h)) / (2 * h)
        f2 = (loss_at(h / 2) - loss_at(-h / 2)) / h
        rich = (4.0 * f2 - f1) / 3.0
        floor = 100 * np.finfo(np.float64).eps * abs(L) / (h / 2)
        say('param', k, index, 'tangent', dL, 'fd', f1, 'richardson', rich)
        assert abs(dL - rich) <= 1e-4 * abs(rich) + floor, (k, index, dL, f1, rich)
