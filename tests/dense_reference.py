"""Dense torch fp64 restatement of the rollout step, for checking the
hand-written adjoint against torch autograd on small inputs.

Same step as the fused GPU step with the Adjusting, AgentAgentInteractions,
Integrator and Reset nodes (circular agents, fixed dt): every ordered pair
(i, j), i != j, whose cells are adjacent under the binning contract (origin
= per-axis minimum, cell index = trunc((p - origin) / cell_size)) adds
agent i's side of interaction_agent_agent_circular
(crowddynamics/core/interactions.py:51-70: power-law social force,
power_law.py:213-259, with truncation to 2 kN, and contact.py:12-48 when
the bodies overlap); force_adjust (adjusting.py:16-51); the reference's
Verlet order (integrator.py:164-176).  The pair set and the branch masks
are taken from the current (detached) state; the arithmetic is
differentiable.  Pinned against the CPU oracle in tests/test_adjoint_host.py.
"""
from __future__ import annotations

import numpy as np
import torch

SIGHT = 3.0
F_MAX = 2e3
TAU_MAX = 30.0


def pair_mask(x, cell_size):
    """[n, n] bool: adjacent cells, i != j (binning contract)."""
    with torch.no_grad():
        o = x.min(0).values
        lxy = torch.trunc((x - o) / cell_size).to(torch.int64)
        dl = (lxy[:, None, :] - lxy[None, :, :]).abs()
        m = (dl[..., 0] <= 1) & (dl[..., 1] <= 1)
        m.fill_diagonal_(False)
    return m


def pair_forces(x, v, r, mass, k_soc, tau_0, mu, kappa, damping, mask):
    """Sum over j of agent i's side of the circular pair term, [n, 2]."""
    dx = x[:, None, :] - x[None, :, :]
    dv = v[:, None, :] - v[None, :, :]
    r_tot = r[:, None] + r[None, :]
    one = torch.ones_like(dx)
    dx_s = torch.where(mask[..., None], dx, one)
    d = torch.hypot(dx_s[..., 0], dx_s[..., 1])
    h = d - r_tot
    sight = mask & (h < SIGHT)
    a = (dv * dv).sum(-1)
    b = -(dx_s * dv).sum(-1)
    c = (dx_s * dx_s).sum(-1) - r_tot * r_tot
    disc = b * b - a * c
    with torch.no_grad():
        ok = sight & (disc > 0) & (a != 0)
    a_s = torch.where(ok, a, torch.ones_like(a))
    disc_s = torch.where(ok, disc, torch.ones_like(disc))
    ds = torch.sqrt(disc_s)
    b_s = torch.where(ok, b, torch.zeros_like(b))
    tau = (b_s - ds) / a_s
    with torch.no_grad():
        ok = ok & (tau > 0) & (tau <= TAU_MAX)
    tau_s = torch.where(ok, tau, torch.ones_like(tau))
    t0 = tau_0[:, None]
    mag = ((2.0 / tau_s + 1.0 / t0) * torch.exp(-tau_s / t0)) / (tau_s * tau_s)
    g = (dv - (dv * b_s[..., None] + dx_s * a_s[..., None]) / ds[..., None]) / a_s[..., None]
    neg_mk = (-mass * k_soc)[:, None, None]
    s = (neg_mk * g) * mag[..., None]
    slen = torch.hypot(s[..., 0], s[..., 1])
    with torch.no_grad():
        big = slen > F_MAX
    scale = torch.where(big, F_MAX / torch.where(big, slen, torch.ones_like(slen)),
                        torch.ones_like(slen))
    s = torch.where(ok[..., None], s * scale[..., None], torch.zeros_like(s))
    with torch.no_grad():
        touch = sight & (h < 0)
    n = dx_s / d[..., None]
    t = torch.stack([n[..., 1], -n[..., 0]], -1)
    kvt = kappa[:, None] * (dv * t).sum(-1)
    cvn = damping[:, None] * (dv * n).sum(-1)
    fc = (-h)[..., None] * (mu[:, None, None] * n - kvt[..., None] * t) + cvn[..., None] * n
    fc = torch.where(touch[..., None], fc, torch.zeros_like(fc))
    return (s + fc).sum(1)


def constants(arr):
    """Per-agent constants of the step as torch fp64 tensors."""
    def col(name):
        return torch.as_tensor(np.array(arr[name], dtype=np.float64, order="C"),
                               dtype=torch.float64)
    return dict(e0=col('target_direction'), m=col('mass'), r=col('radius'),
                k_soc=col('k_soc'), tau_adj=col('tau_adj'))


def per_agent(p, n):
    p = torch.as_tensor(p, dtype=torch.float64)
    return p.expand(n) if p.dim() == 0 else p


def step(x, v, fp, K, dt, cell_size, tau_0, mu, kappa, damping, v0):
    """One step from (x, v, force_prev); returns (x', v', F)."""
    n = x.shape[0]
    m = K['m']
    t0, mu_, ka, da, v0_ = (per_agent(p, n) for p in (tau_0, mu, kappa, damping, v0))
    mask = pair_mask(x, cell_size)
    F = (m / K['tau_adj'])[:, None] * (v0_[:, None] * K['e0'] - v)
    F = F + pair_forces(x, v, K['r'], m, K['k_soc'], t0, mu_, ka, da, mask)
    a_old = fp / m[:, None]
    a_new = F / m[:, None]
    v = v + (a_old + a_new) / 2.0 * dt
    x = x + v * dt + a_new / 2.0 * (dt * dt)
    return x, v, F


def rollout(arr, steps, dt, cell_size, tau_0, mu, kappa, damping, v0):
    """Positions [steps + 1, n, 2] from the records `arr` (numpy); the
    parameters are torch fp64 tensors (scalars or per agent)."""
    K = constants(arr)

    def col(name):
        return torch.as_tensor(np.array(arr[name], dtype=np.float64, order="C"),
                               dtype=torch.float64)

    x, v, fp = col('position'), col('velocity'), col('force_prev')
    out = [x]
    for _ in range(steps):
        x, v, fp = step(x, v, fp, K, dt, cell_size, tau_0, mu, kappa, damping, v0)
        out.append(x)
    return torch.stack(out)
