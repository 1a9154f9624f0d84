"""Seeded scenario generators for the benchmark and parity workloads.

The reference builds scenarios with shapely geometry and scikit-fmm fields
(crowddynamics/examples/simulations.py, crowddynamics/examples/fields.py);
those are setup-time tools outside the per-step path.  These generators
produce the same kind of inputs directly: agent records, wall segments
(obstacle_type_linear) and a precomputed navigation field.

Workloads (all parameters the reference's trait defaults unless stated):

* ``hallway``    200 Circular agents, 20 x 5 m hallway with two line walls,
                 desired direction +x, cell 3.6 m, dt in [0.001, 0.01].
* ``room``       20 000 Circular agents in a 200 x 100 m room with four
                 1.2 m exits (one centred on each wall), potential = distance
                 to the nearest exit on a 0.1 m grid, cell 0.6 m.
* ``corridor``   50 000 ThreeCircle agents in a 300 x 150 m corridor with 200
                 random wall segments, cell 0.6 m.
* ``ensemble``   S independent Circular rooms with log-uniform agent counts.
* ``rollout``    4 000 Circular agents for the differentiable rollout.

Agent attributes: radius ~ U[0.22, 0.28], mass ~ N(80, 8) and desired speed
~ N(1.15, 0.2), both truncated at +-3 sigma (as the reference's
truncnorm(-3, 3), crowddynamics/simulation/agents.py:131-134); inertia
4*pi*(m/80)*(r/0.27)**2 (agents.py:185-192); initial velocity v0 * e0.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .agents import empty_agents, inertia_rot, shoulders_host
from .core import obstacle_type_linear


@dataclass
class Scenario:
    name: str
    agents: np.ndarray
    segments: np.ndarray
    cell_size: float
    dt_min: float
    dt_max: float
    steps: int
    seed: int = 1
    nav: dict | None = None
    bounds: tuple = (0.0, 0.0, 1.0, 1.0)
    extra: dict = field(default_factory=dict)

    @property
    def n(self):
        return len(self.agents)


def segments_array(seg_list):
    seg = np.zeros(len(seg_list), dtype=obstacle_type_linear)
    for k, (p0, p1) in enumerate(seg_list):
        seg[k]['p0'] = p0
        seg[k]['p1'] = p1
    return seg


def _truncnormal(rng, mean, sd, size):
    out = rng.normal(mean, sd, size)
    bad = np.abs(out - mean) > 3 * sd
    while bad.any():
        out[bad] = rng.normal(mean, sd, bad.sum())
        bad = np.abs(out - mean) > 3 * sd
    return out


def _seg_dist(px, py, seg):
    """Distance from points (px, py) to every segment: [K, W]."""
    p0 = seg[:, 0:2]
    p1 = seg[:, 2:4]
    d = p1 - p0
    l2 = (d ** 2).sum(1)
    l2 = np.where(l2 == 0, 1.0, l2)
    t = ((px[:, None] - p0[None, :, 0]) * d[None, :, 0] +
         (py[:, None] - p0[None, :, 1]) * d[None, :, 1]) / l2[None, :]
    t = np.clip(t, 0.0, 1.0)
    qx = p0[None, :, 0] + t * d[None, :, 0]
    qy = p0[None, :, 1] + t * d[None, :, 1]
    return np.hypot(px[:, None] - qx, py[:, None] - qy)


def place_rejection(rng, n, radii, box, segs=None, chunk=2048, max_rounds=10000):
    """Non-overlapping placement by seeded rejection sampling.

    Candidates are drawn uniformly in the box shrunk by each radius, in
    chunks; a candidate is rejected when it overlaps an agent placed earlier,
    an earlier candidate of its chunk, or a wall segment.  Returns positions
    for the first `placed` radii (placed < n when the room is too full)."""
    from scipy.spatial import cKDTree
    x0, y0, x1, y1 = box
    rmax = float(radii.max()) if n else 0.0
    seg = None
    if segs is not None and len(segs):
        seg = np.concatenate([segs['p0'], segs['p1']], axis=1)
    pos = np.zeros((n, 2))
    placed = 0
    rounds = 0
    while placed < n and rounds < max_rounds:
        rounds += 1
        k = min(chunk, n - placed)
        r = radii[placed:placed + k]
        cx = rng.uniform(x0 + r, x1 - r)
        cy = rng.uniform(y0 + r, y1 - r)
        c = np.stack([cx, cy], 1)
        ok = np.ones(k, dtype=bool)
        if placed:
            tree = cKDTree(pos[:placed])
            nbrs = tree.query_ball_point(c, 2 * rmax)
            for a, lst in enumerate(nbrs):
                if lst:
                    d = np.hypot(*(pos[lst] - c[a]).T)
                    if np.any(d < r[a] + radii[lst]):
                        ok[a] = False
        ct = cKDTree(c)
        for a, b in ct.query_pairs(2 * rmax):
            lo, hi = (a, b) if a < b else (b, a)
            if np.hypot(*(c[lo] - c[hi])) < r[lo] + r[hi]:
                ok[hi] = False
        if seg is not None:
            d = _seg_dist(cx, cy, seg)
            ok &= ~np.any(d < r[:, None], axis=1)
        acc = np.nonzero(ok)[0]
        # accepted candidates take the next radii in order
        m = len(acc)
        if m == 0:
            continue
        # keep radius/position pairing: only accept a prefix-consistent set
        chosen = c[acc]
        chosen_r = r[acc]
        # reorder the unplaced radii so accepted ones come first
        rest = np.setdiff1d(np.arange(k), acc)
        radii[placed:placed + k] = np.concatenate([chosen_r, r[rest]])
        pos[placed:placed + m] = chosen
        placed += m
    return pos[:placed], placed


def place_hex(rng, n, radii, side_x, side_y, origin=(0.0, 0.0)):
    """Jittered hexagonal lattice: n sites chosen at random from a lattice of
    spacing a >= 2*max(r); each agent is displaced by at most
    (a - 2 r_max)/2, so no two agents overlap."""
    rmax = float(radii.max())
    a = 2.0 * rmax * 1.0001
    dy = a * math.sqrt(3.0) / 2.0
    nx = int((side_x - 2 * rmax) // a) + 1
    ny = int((side_y - 2 * rmax) // dy) + 1
    jj, ii = np.meshgrid(np.arange(nx), np.arange(ny))
    sx = rmax + jj * a + (ii % 2) * a / 2.0
    sy = rmax + ii * dy
    keep = (sx <= side_x - rmax) & (sy <= side_y - rmax)
    sites = np.stack([sx[keep], sy[keep]], 1)
    if len(sites) < n:
        raise ValueError('room too small for %d agents' % n)
    sel = rng.choice(len(sites), n, replace=False)
    p = sites[sel]
    slack = max(a - 2.0 * rmax, 0.0) / 2.0
    ang = rng.uniform(0, 2 * np.pi, n)
    rad = rng.uniform(0, slack, n)
    p = p + np.stack([rad * np.cos(ang), rad * np.sin(ang)], 1)
    return p + np.asarray(origin)


def _attributes(rng, n, model='circular'):
    arr = empty_agents(n, model)
    arr['radius'] = rng.uniform(0.22, 0.28, n)
    arr['mass'] = _truncnormal(rng, 80.0, 8.0, n)
    arr['target_velocity'] = _truncnormal(rng, 1.15, 0.2, n)
    arr['inertia_rot'] = inertia_rot(arr['mass'], arr['radius'])
    return arr


def hallway(n=200, seed=0, width=20.0, height=5.0, steps=1000, cell_size=3.6):
    """200 Circular agents in a 20 x 5 m hallway bounded by two line walls."""
    rng = np.random.default_rng(seed)
    arr = _attributes(rng, n)
    segs = segments_array([((0.0, 0.0), (width, 0.0)), ((0.0, height), (width, height))])
    r = arr['radius'].copy()
    pos, placed = place_rejection(rng, n, r, (0.0, 0.0, width, height), segs)
    if placed < n:
        raise ValueError('hallway placement failed')
    arr['radius'] = r
    arr['inertia_rot'] = inertia_rot(arr['mass'], arr['radius'])
    arr['position'] = pos
    arr['target_direction'] = (1.0, 0.0)
    arr['velocity'] = arr['target_velocity'][:, None] * arr['target_direction']
    arr['target'] = 0
    return Scenario('hallway', arr, segs, cell_size, 0.001, 0.01, steps, seed=1,
                    bounds=(0.0, 0.0, width, height))


def exit_potential(exits, minx, miny, maxx, maxy, step):
    """Distance to the nearest exit segment on the meshgrid
    (crowddynamics/core/steering/quickest_path.py:22-49 grid: inclusive ends)
    and the unit direction toward it: potential [rows, cols], direction
    [rows, cols, 2] (zero where the distance vanishes)."""
    xs = np.arange(minx, maxx + step, step)
    ys = np.arange(miny, maxy + step, step)
    X, Y = np.meshgrid(xs, ys, indexing='xy')
    px, py = X.ravel(), Y.ravel()
    seg = np.asarray(exits, dtype=np.float64).reshape(-1, 4)
    best = np.full(px.shape, np.inf)
    qx = np.zeros_like(px)
    qy = np.zeros_like(py)
    for p0x, p0y, p1x, p1y in seg:
        dx, dy = p1x - p0x, p1y - p0y
        l2 = dx * dx + dy * dy
        t = np.clip(((px - p0x) * dx + (py - p0y) * dy) / l2, 0.0, 1.0)
        cx, cy = p0x + t * dx, p0y + t * dy
        d = np.hypot(px - cx, py - cy)
        m = d < best
        best[m] = d[m]
        qx[m] = cx[m]
        qy[m] = cy[m]
    ex, ey = qx - px, qy - py
    nrm = np.hypot(ex, ey)
    with np.errstate(invalid='ignore', divide='ignore'):
        ex = np.where(nrm > 0, ex / nrm, 0.0)
        ey = np.where(nrm > 0, ey / nrm, 0.0)
    shape = X.shape
    return best.reshape(shape), np.stack([ex, ey], -1).reshape(shape + (2,))


def room(n=20000, seed=0, width=200.0, height=100.0, exit_width=1.2, steps=500,
         cell_size=0.6, nav_step=0.1, nav_mode=1):
    """Room with four exits, one centred on each wall; agents read the
    distance-to-exit potential (nav_mode 1: bilinear gradient; 0: nearest
    node of the direction map)."""
    rng = np.random.default_rng(seed)
    arr = _attributes(rng, n)
    w, h, e = width, height, exit_width / 2.0
    walls = [((0, 0), (w / 2 - e, 0)), ((w / 2 + e, 0), (w, 0)),
             ((w, 0), (w, h / 2 - e)), ((w, h / 2 + e), (w, h)),
             ((w, h), (w / 2 + e, h)), ((w / 2 - e, h), (0, h)),
             ((0, h), (0, h / 2 + e)), ((0, h / 2 - e), (0, 0))]
    segs = segments_array(walls)
    exits = [(w / 2 - e, 0, w / 2 + e, 0), (w, h / 2 - e, w, h / 2 + e),
             (w / 2 - e, h, w / 2 + e, h), (0, h / 2 - e, 0, h / 2 + e)]
    r = arr['radius'].copy()
    pos, placed = place_rejection(rng, n, r, (0.0, 0.0, w, h), segs)
    if placed < n:
        raise ValueError('room placement failed (%d of %d)' % (placed, n))
    arr['radius'] = r
    arr['inertia_rot'] = inertia_rot(arr['mass'], arr['radius'])
    arr['position'] = pos
    pot, dirs = exit_potential(exits, 0.0, 0.0, w, h, nav_step)
    arr['target'] = 0
    # initial desired direction from the direction map at the nearest node
    jj = np.clip(np.rint(pos[:, 0] / nav_step).astype(int), 0, dirs.shape[1] - 1)
    ii = np.clip(np.rint(pos[:, 1] / nav_step).astype(int), 0, dirs.shape[0] - 1)
    e0 = dirs[ii, jj]
    arr['target_direction'] = e0
    arr['velocity'] = arr['target_velocity'][:, None] * e0
    nav = dict(maps=(pot[None] if nav_mode == 1 else dirs[None]), minx=0.0, miny=0.0,
               step=nav_step, mode=nav_mode, potential=pot, directions=dirs)
    return Scenario('room', arr, segs, cell_size, 0.001, 0.01, steps, seed=1, nav=nav,
                    bounds=(0.0, 0.0, w, h), extra=dict(exits=exits))


def corridor(n=50000, seed=0, width=300.0, height=150.0, n_obstacles=200, steps=300,
             cell_size=0.6):
    """ThreeCircle agents in a corridor with random line obstacles; desired
    direction +x; target orientation from the velocity (extra['orient'];
    the reference's Orientation node takes it from the desired direction)."""
    rng = np.random.default_rng(seed)
    arr = empty_agents(n, 'three_circle')
    arr['r_t'] = 0.15
    arr['r_s'] = 0.10
    arr['r_ts'] = 0.15
    arr['radius'] = 0.25
    arr['mass'] = _truncnormal(rng, 80.0, 8.0, n)
    arr['target_velocity'] = _truncnormal(rng, 1.15, 0.2, n)
    arr['inertia_rot'] = inertia_rot(arr['mass'], arr['radius'])
    arr['target_angular_velocity'] = 2.0 * np.pi
    orng = np.random.default_rng(seed + 2)
    obs = []
    for _ in range(n_obstacles):
        c = orng.uniform((5.0, 5.0), (width - 5.0, height - 5.0))
        L = orng.uniform(1.0, 5.0)
        a = orng.uniform(0, np.pi)
        d = 0.5 * L * np.array([np.cos(a), np.sin(a)])
        obs.append((tuple(c - d), tuple(c + d)))
    walls = [((0, 0), (width, 0)), ((width, 0), (width, height)),
             ((width, height), (0, height)), ((0, height), (0, 0))]
    segs = segments_array(walls + obs)
    # three-circle footprint: torso and shoulders fit in the 0.25 m radius
    r = np.full(n, 0.25)
    pos, placed = place_rejection(rng, n, r, (0.0, 0.0, width, height), segs)
    if placed < n:
        raise ValueError('corridor placement failed (%d of %d)' % (placed, n))
    arr['position'] = pos
    arr['orientation'] = 0.0
    arr['target_orientation'] = 0.0
    arr['target_direction'] = (1.0, 0.0)
    arr['velocity'] = arr['target_velocity'][:, None] * arr['target_direction']
    arr['target'] = 0
    shoulders_host(arr)
    # the workload aligns each body with its velocity (target_orientation =
    # angle(v)); the reference's Orientation node aligns it with the desired
    # direction (core/steering/orientation.py:17-21)
    return Scenario('corridor', arr, segs, cell_size, 0.001, 0.01, steps, seed=1,
                    bounds=(0.0, 0.0, width, height), extra=dict(orient='velocity'))


def ensemble(n_scenarios=256, seed=0, n_min=500, n_max=8000, rho=(0.5, 3.0), steps=1000,
             cell_size=3.6):
    """Independent square rooms: N_s log-uniform in [n_min, n_max], density
    U[rho], side sqrt(N_s / rho_s), four walls, desired direction +x.
    Placement on a jittered hexagonal lattice (dense rooms exceed what random
    sequential placement reaches).  Returns a list of Scenario."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(n_scenarios):
        srng = np.random.default_rng(seed * 100003 + s)
        ns = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        rho_s = rng.uniform(*rho)
        side = math.sqrt(ns / rho_s)
        arr = _attributes(srng, ns)
        while True:
            try:
                pos = place_hex(srng, ns, arr['radius'], side, side)
                break
            except ValueError:
                side *= 1.05
        arr['position'] = pos
        arr['target_direction'] = (1.0, 0.0)
        arr['velocity'] = arr['target_velocity'][:, None] * arr['target_direction']
        arr['target'] = 0
        segs = segments_array([((0, 0), (side, 0)), ((side, 0), (side, side)),
                               ((side, side), (0, side)), ((0, side), (0, 0))])
        out.append(Scenario('ensemble-%d' % s, arr, segs, cell_size, 0.001, 0.01, steps,
                            seed=1 + s, bounds=(0.0, 0.0, side, side),
                            extra=dict(density=rho_s)))
    return out


def rollout(n=4000, seed=0, steps=500, dt=0.005, density=1.0):
    """Circular agents for the differentiable rollout: open square room at the
    given density, desired direction toward the room centre's +x side."""
    rng = np.random.default_rng(seed)
    arr = _attributes(rng, n)
    side = math.sqrt(n / density)
    r = arr['radius'].copy()
    pos, placed = place_rejection(rng, n, r, (0.0, 0.0, side, side), None)
    if placed < n:
        raise ValueError('rollout placement failed')
    arr['radius'] = r
    arr['inertia_rot'] = inertia_rot(arr['mass'], arr['radius'])
    arr['position'] = pos
    arr['target_direction'] = (1.0, 0.0)
    arr['velocity'] = arr['target_velocity'][:, None] * arr['target_direction']
    arr['target'] = 0
    return Scenario('rollout', arr, segments_array([]), 3.6, dt, dt, steps, seed=1,
                    bounds=(0.0, 0.0, side, side))


class Hallway:
    """Factory of the reference's Hallway simulation
    (crowddynamics/examples/simulations.py:74-163) on the drop-in classes:
    two line walls, targets at both ends, two groups walking in opposite
    directions (target 1 toward +x, target 0 toward -x), and the same logic
    tree.  Navigation reads nearest-node direction maps of the distance to
    each end (the reference derives them by FMM; for a straight hallway the
    distance maps are the closed forms used here)."""

    def __new__(cls, size=50, agent_type='circular', width=40.0, height=5.0, ratio=1.0 / 3,
                 seed=0, fused='auto', precision='f64'):
        from .logic import (Adjusting, AgentAgentInteractions, AgentObstacleInteractions,
                            Fluctuation, InsideDomain, Integrator, Navigation, Orientation,
                            Reset)
        from .simulation import Agents, Field, MultiAgentSimulation
        rng = np.random.default_rng(seed)
        n = size // 2 * 2
        arr = _attributes(rng, n, agent_type)
        if agent_type == 'three_circle':
            r = arr['radius']
            arr['r_t'], arr['r_s'], arr['r_ts'] = 0.5882 * r, 0.3725 * r, 0.6275 * r
        half = n // 2
        segs = segments_array([((0.0, 0.0), (width, 0.0)), ((0.0, height), (width, height))])
        r = arr['radius'].copy()
        p0, k0 = place_rejection(rng, half, r[:half].copy(),
                                 (0.0, 0.0, ratio * width, height), segs)
        p1, k1 = place_rejection(rng, n - half, r[half:].copy(),
                                 ((1 - ratio) * width, 0.0, width, height), segs)
        if k0 < half or k1 < n - half:
            raise ValueError('hallway spawn areas too small for %d agents' % n)
        arr['position'][:half] = p0
        arr['position'][half:] = p1
        for sl, phi, target in ((slice(0, half), 0.0, 1), (slice(half, n), np.pi, 0)):
            e = np.array([np.cos(phi), np.sin(phi)])
            arr['target_direction'][sl] = e
            arr['velocity'][sl] = e
            arr['target'][sl] = target
            if agent_type == 'three_circle':
                arr['orientation'][sl] = phi
                arr['target_orientation'][sl] = phi
        if agent_type == 'three_circle':
            shoulders_host(arr)
        step = 0.1
        cols = int(np.floor(width / step + 1e-9)) + 1
        rows = int(np.floor(height / step + 1e-9)) + 1
        maps = np.zeros((2, rows, cols, 2))
        maps[0, ..., 0] = -1.0   # target 0: the x = 0 end
        maps[1, ..., 0] = 1.0    # target 1: the x = width end
        field = Field(obstacles=segs, targets=2, navigation=maps, nav_origin=(0.0, 0.0),
                      nav_step=step, nav_mode=0,
                      domain=[(0, 0), (width, 0), (width, height), (0, height)])
        sim = MultiAgentSimulation(field=field, agents=Agents(agent_type, arr, precision),
                                   name='Hallway', fused=fused,
                                   cell_capacity=int((width * 2 / 3.6 + 4) * (height * 4 / 3.6 + 4)))
        sim.logic = Reset(sim) << InsideDomain(sim) << (
            Integrator(sim) << (
                Fluctuation(sim),
                Adjusting(sim) << (Navigation(sim), Orientation(sim)),
                AgentAgentInteractions(sim),
                AgentObstacleInteractions(sim)))
        return sim
