"""Hand-written adjoint (cd_adjoint_step behind RolloutFunction) against
torch autograd through the dense reference, and against central finite
differences of the GPU forward itself (fp64; h = 1e-6 for tau_0 and the
desired speeds, 1e-4 of the value for mu, kappa and damping).

Finite differences are taken on windows where they are meaningful: tau_0
and target_velocity on contact-free windows (sparse rooms), mu, kappa and
damping on windows whose contacts persist throughout (pairs pressed
together), since the loss is flat in them without contacts."""
import numpy as np
import pytest
import torch

from crowddynamics_amd.adjoint import Rollout, rollout_positions
from tests import adjoint_cases as AC
from tests import dense_reference as D

pytestmark = pytest.mark.gpu

DT = 0.005
CELL = 3.6


def say(*a):
    print(*a, flush=True)


def _params(arr, device, grad=False):
    ps = [torch.tensor(3.0), torch.tensor(1.2e5), torch.tensor(4e4), torch.tensor(500.0),
          torch.as_tensor(np.array(arr['target_velocity'], dtype=np.float64))]
    return [p.to(device=device, dtype=torch.float64).requires_grad_(grad) for p in ps]


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


def _adjoint_grads(R, arr, Xhat):
    p = _params(arr, 'cuda', grad=True)
    L = _loss(rollout_positions(R, *p), Xhat)
    L.backward()
    return float(L.detach()), [t.grad.detach().cpu() for t in p]


@pytest.mark.parametrize('case', ['sparse', 'packed'])
def test_adjoint_step_matches_dense_autograd(case):
    """One step's adjoint on the same state: lam_x, lam_v, lam_g and the
    per-agent parameter gradients equal torch autograd of the dense step
    (random adjoint seeds, contacts and truncated social forces included)."""
    arr = AC.sparse(40, seed=4) if case == 'sparse' else AC.packed(36, seed=2)
    rng = np.random.default_rng(5)
    arr['velocity'] += rng.normal(0.0, 0.5, (len(arr), 2))
    arr['force_prev'] = rng.normal(0.0, 50.0, (len(arr), 2))
    n = len(arr)
    R = Rollout(arr, steps=1, dt=DT, cell_size=CELL)
    R.set_parameters(*_params(arr, 'cuda'))
    R.agents.upload(arr)
    R.set_parameters(*_params(arr, 'cuda'))
    seeds = {k: rng.normal(0.0, 1.0, (n, 2)) for k in ('lam_x', 'lam_v', 'lam_g', 'gx')}
    for k in ('lam_x', 'lam_v', 'lam_g'):
        R.adj[k].copy_(torch.as_tensor(seeds[k]))
    R.grad.zero_()
    gx = torch.as_tensor(seeds['gx'], device='cuda')
    R.adjoint_step(gx)
    torch.cuda.synchronize()
    got = {k: R.adj[k].cpu() for k in ('lam_x', 'lam_v', 'lam_g')}
    got_grad = R.grad.cpu()
    # dense autograd of S = lam_x . x' + lam_v . v' + lam_g . F
    K = D.constants(arr)
    x = torch.as_tensor(np.array(arr['position'])).requires_grad_(True)
    v = torch.as_tensor(np.array(arr['velocity'])).requires_grad_(True)
    fp = torch.as_tensor(np.array(arr['force_prev'])).requires_grad_(True)
    p = [t.expand(n).clone() if t.dim() == 0 else t.clone() for t in _params(arr, 'cpu')]
    p = [t.requires_grad_(True) for t in p]
    x1, v1, F = D.step(x, v, fp, K, DT, CELL, *p)
    S = ((torch.as_tensor(seeds['lam_x']) * x1).sum() + (torch.as_tensor(seeds['lam_v']) * v1).sum()
         + (torch.as_tensor(seeds['lam_g']) * F).sum())
    S.backward()
    want = {'lam_x': x.grad + torch.as_tensor(seeds['gx']), 'lam_v': v.grad, 'lam_g': fp.grad}
    for k in want:
        scale = float(torch.max(torch.abs(want[k])))
        err = float(torch.max(torch.abs(got[k] - want[k])))
        say(case, k, 'rel err', err / scale)
        assert err <= 1e-11 * scale, (k, err, scale)
    for c, t in enumerate(p):
        scale = float(torch.max(torch.abs(t.grad))) + 1e-300
        err = float(torch.max(torch.abs(got_grad[:, c] - t.grad)))
        say(case, 'param', c, 'rel err', err / scale)
        assert err <= 1e-10 * scale, (c, err, scale)


@pytest.mark.parametrize('case', ['sparse', 'dimers'])
def test_adjoint_matches_dense_autograd(case):
    """Whole rollouts (40 steps, checkpoint 7): the gradient of the
    trajectory loss equals torch autograd through the dense reference."""
    arr = AC.sparse(28, seed=2) if case == 'sparse' else AC.dimers(15, seed=2)
    steps = 40
    say('case', case, 'n', len(arr), 'steps', steps)
    R = Rollout(arr, steps=steps, dt=DT, cell_size=CELL, checkpoint=7)
    Xhat = _targets(R, arr)
    L, g = _adjoint_grads(R, arr, Xhat)
    p = _params(arr, 'cpu', grad=True)
    Xd = D.rollout(arr, steps, DT, CELL, *p)
    Ld = _loss(Xd, Xhat.cpu())
    Ld.backward()
    with torch.no_grad():
        X = rollout_positions(R, *_params(arr, 'cuda')).cpu()
    assert torch.max(torch.abs(X - Xd)).item() < 1e-10
    if case == 'dimers':
        assert float(AC.min_skin(Xd.detach(), arr['radius']).max()) < 0.0
    assert abs(L - float(Ld.detach())) <= 1e-9 * abs(L)
    for k, (a, b) in enumerate(zip(g, (t.grad for t in p))):
        scale = float(torch.max(torch.abs(b))) + 1e-300
        err = float(torch.max(torch.abs(a - b)))
        say('param', k, 'max |adjoint - autograd| / max|autograd| =', err / scale)
        assert err <= 1e-8 * scale, (k, err, scale)


def _fd_checked(R, arr, Xhat, k, h, L, g, index=None, rtol=1e-4, rtol_rich=3e-5):
    """Central differences at h (default within 1e-4) and their Richardson
    extrapolation from h and h / 2 (default within 3e-5).  The power-law force has
    near-singular terms (time to collision at a small discriminant), so
    the truncation error of a single difference reaches the 1e-5 level on
    windows with many interacting pairs; the dense reference shows the same
    finite-difference errors on the same window while agreeing with the
    adjoint to 1e-8 (test_adjoint_matches_dense_autograd)."""
    f1 = _fd(R, arr, Xhat, k, h, index)
    f2 = _fd(R, arr, Xhat, k, h / 2, index)
    rich = (4.0 * f2 - f1) / 3.0
    say('param', k, index, 'adjoint', g, 'fd', f1, 'richardson', rich, 'loss', L)
    assert _close(g, f1, L, h / 2, rtol=rtol)
    assert _close(g, rich, L, h / 2, rtol=rtol_rich)


def _close(g, fd, L, h, rtol=1e-5):
    """|g - fd| within rtol plus the rounding floor of the central
    difference (eps L / h)."""
    floor = 100 * np.finfo(np.float64).eps * abs(L) / h
    return abs(g - fd) <= rtol * abs(fd) + floor


def _fd(R, arr, Xhat, k, h, index=None):
    def loss_at(delta):
        p = _params(arr, 'cuda')
        if index is None:
            p[k] = p[k] + delta
        else:
            p[k] = p[k].clone()
            p[k][index] += delta
        with torch.no_grad():
            return float(_loss(rollout_positions(R, *p), Xhat))
    return (loss_at(h) - loss_at(-h)) / (2 * h)


def test_adjoint_central_fd_contact_free():
    arr = AC.sparse(300, seed=5, density=0.5)
    steps = 30
    R = Rollout(arr, steps=steps, dt=DT, cell_size=CELL, checkpoint=25)
    Xhat = _targets(R, arr, seed=1)
    with torch.no_grad():
        X = rollout_positions(R, *_params(arr, 'cuda'))
    assert float(AC.min_skin(X, arr['radius']).min()) > 0.0
    L, g = _adjoint_grads(R, arr, Xhat)
    _fd_checked(R, arr, Xhat, 0, 1e-6, L, float(g[0]))
    for i in (0, 17, 299):
        _fd_checked(R, arr, Xhat, 4, 1e-6, L, float(g[4][i]), index=i)


def test_adjoint_central_fd_persistent_contacts():
    arr = AC.dimers(30, seed=3)
    steps = 20
    R = Rollout(arr, steps=steps, dt=DT, cell_size=CELL, checkpoint=6)
    Xhat = _targets(R, arr, seed=2)
    with torch.no_grad():
        X = rollout_positions(R, *_params(arr, 'cuda'))
    # the set of touching pairs is the same at every step of the window
    rr = torch.as_tensor(np.array(arr['radius']), device='cuda')
    touch = (torch.cdist(X, X) - (rr[:, None] + rr[None, :])) < 0
    assert bool(torch.all(touch == touch[0])) and int(touch[0].sum()) > 0
    L, g = _adjoint_grads(R, arr, Xhat)
    for k, name in ((1, 'mu'), (2, 'kappa'), (3, 'damping')):
        # mu, kappa and damping are O(1e2..1e5): a step of 1e-6 of their
        # value moves the trajectory by less than a few hundred ulps of the
        # positions, so the differences use 1e-4 of the value (the window is
        # smooth in them: truncation error below 1e-7 of the gradient)
        h = 1e-4 * float(_params(arr, 'cpu')[k])
        _fd_checked(R, arr, Xhat, k, h, L, float(g[k]), rtol=1e-5, rtol_rich=1e-5)


def test_adjoint_lanes_per_agent():
    """2, 16, 32 and 64 lanes per agent (the stencil's candidates split in
    fewer or more parts) give the gradients of 4 lanes up to summation
    order."""
    arr = AC.packed(40, seed=8)
    out = []
    for lanes in (4, 2, 16, 32, 64):
        R = Rollout(arr, steps=20, dt=DT, cell_size=CELL, checkpoint=6, lanes=lanes)
        Xhat = _targets(R, arr, seed=4)
        out.append(_adjoint_grads(R, arr, Xhat)[1])
    for other in out[1:]:
        for a, b in zip(out[0], other):
            scale = float(torch.max(torch.abs(a))) + 1e-300
            assert float(torch.max(torch.abs(a - b))) <= 1e-12 * scale


def test_adjoint_processing_order_is_exact():
    """Agents served in cell order (rebuilt every forward step, or every
    third) give the index-order gradients bit for bit."""
    arr = AC.packed(40, seed=8)
    out = []
    for every in (0, 1, 3):
        R = Rollout(arr, steps=20, dt=DT, cell_size=CELL, checkpoint=6, lanes=16,
                    order_every=every)
        Xhat = _targets(R, arr, seed=4)
        out.append(_adjoint_grads(R, arr, Xhat)[1])
    for g in out[1:]:
        for a, b in zip(g, out[0]):
            assert torch.equal(a, b)


def test_adjoint_checkpoint_interval_is_exact():
    arr = AC.packed(40, seed=6)
    out = []
    for ck in (1, 9, 30):
        R = Rollout(arr, steps=30, dt=DT, cell_size=CELL, checkpoint=ck)
        Xhat = _targets(R, arr, seed=3)
        out.append(_adjoint_grads(R, arr, Xhat)[1])
    for g in out[1:]:
        for a, b in zip(g, out[0]):
            assert torch.equal(a, b)
