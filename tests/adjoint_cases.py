"""Agent sets for the adjoint tests: a sparse open room (social forces
only) and a compressed lattice whose neighbours overlap (persistent
contacts)."""
import math

import numpy as np

from crowddynamics_amd import examples as E


def sparse(n=24, seed=0, density=0.6):
    arr = E.rollout(n=n, seed=seed, density=density).agents.copy()
    rng = np.random.default_rng(seed + 7)
    arr['velocity'] += rng.normal(0.0, 0.3, (n, 2))
    return arr


def packed(n=30, seed=1, squeeze=0.995):
    """A crowd pressing inward: hex lattice at spacing 2 r_max shrunk by
    `squeeze` (equal radii, so lattice neighbours overlap), desired
    directions toward the centre, nearly at rest.  The adjusting force holds
    the overlaps, so the touching pairs persist over short windows."""
    rng = np.random.default_rng(seed)
    arr = E._attributes(rng, n)
    arr['radius'] = 0.25
    side = math.sqrt(n) * 0.6
    while True:
        try:
            pos = E.place_hex(rng, n, arr['radius'], side, side)
            break
        except ValueError:
            side *= 1.1
    pos = pos * squeeze
    arr['position'] = pos
    c = pos.mean(0)
    e = c - pos
    e /= np.maximum(np.hypot(e[:, 0], e[:, 1]), 1e-9)[:, None]
    arr['target_direction'] = e
    arr['velocity'] = rng.normal(0.0, 0.02, (n, 2))
    arr['target'] = 0
    return arr


def min_skin(X, r):
    """Smallest h = |x_i - x_j| - r_i - r_j over all pairs and steps (torch)."""
    import torch
    rr = torch.as_tensor(np.array(r, dtype=np.float64), dtype=X.dtype, device=X.device)
    d = torch.cdist(X, X)
    h = d - (rr[:, None] + rr[None, :])
    n = h.shape[-1]
    h = h + torch.eye(n, dtype=X.dtype, device=X.device) * 1e9
    return h.amin(dim=(-1, -2))


def dimers(n_pairs=20, seed=2, spacing=5.0, overlap=1.5e-3, slide=0.05):
    """Isolated pairs pressed together: each pair overlaps by about
    `overlap` with desired directions toward each other (the adjusting force
    balances the contact force), sliding tangentially at +-`slide` so the
    friction term is active; pairs sit `spacing` apart, beyond sight of
    each other.  The touching pairs persist over short windows."""
    rng = np.random.default_rng(seed)
    n = 2 * n_pairs
    arr = E._attributes(rng, n)
    arr['radius'] = 0.25
    side = int(math.ceil(math.sqrt(n_pairs)))
    pos = np.zeros((n, 2))
    e0 = np.zeros((n, 2))
    vel = np.zeros((n, 2))
    for k in range(n_pairs):
        c = spacing * np.array([k % side, k // side], dtype=np.float64)
        th = rng.uniform(0, 2 * np.pi)
        u = np.array([math.cos(th), math.sin(th)])
        t = np.array([-u[1], u[0]])
        half = 0.5 * (0.5 - overlap * rng.uniform(0.8, 1.2))
        pos[2 * k] = c - half * u
        pos[2 * k + 1] = c + half * u
        e0[2 * k] = u
        e0[2 * k + 1] = -u
        vel[2 * k] = slide * t
        vel[2 * k + 1] = -slide * t
    arr['position'] = pos
    arr['target_direction'] = e0
    arr['velocity'] = vel
    arr['target'] = 0
    return arr
