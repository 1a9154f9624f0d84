"""Autotuner of the batched step: picks the block-list cell size and the
force kernel's block size for a given ensemble by timing candidates on the
device.

Only physics-neutral cell sizes are candidates.  An agent pair interacts
when its skin distance is below the social-force sight (3 m,
crowddynamics/core/interactions.py:15-16 via
crowddynamics/simulation/logic.py:97-119) or the agents touch, so every
interacting pair of agents with radii <= r_max lies within
sight + 2 r_max; with cells at least that wide all such pairs sit in
adjacent cells and the cell size changes only the candidate set and the
summation order, not which pairs contribute.  Results of two candidates
therefore agree to rounding.
"""
from __future__ import annotations

import numpy as np

from .ensemble import Ensemble

SIGHT_SOC = 3.0


def max_body_radius(scenarios):
    r = [float(np.max(s.agents['radius'])) for s in scenarios if len(s.agents)]
    return max(r) if r else 0.0


def neutral_cell_sizes(r_max, factors=(1.0, 1.15, 1.35, 1.6), sight=SIGHT_SOC):
    """Candidate cell sizes >= sight + 2 r_max (rounded up to 1 mm)."""
    base = np.ceil((sight + 2.0 * r_max) * 1000.0) / 1000.0
    return [float(np.ceil(base * f * 1000.0) / 1000.0) for f in factors]


def tune_ensemble(scenarios, precision='f32', cell_sizes=None, force_threads=(64, 128, 256),
                  steps=10, warmup=3, flags=None):
    """Time every (cell size, force block) pair over `steps` batched steps
    (after `warmup` steps) and return the fastest.

    Returns dict(cell_size, force_threads, ms_per_step, table) where table
    lists (ms_per_step, cell_size, force_threads, per-kernel ms)."""
    if cell_sizes is None:
        cell_sizes = neutral_cell_sizes(max_body_radius(scenarios))
    kw = {} if flags is None else dict(flags=flags)
    table = []
    for cs in cell_sizes:
        for ft in force_threads:
            ens = Ensemble(scenarios, precision=precision, cell_size=cs, force_threads=ft, **kw)
            ens.step(warmup)
            ms = ens.timed(steps)
            ens.check()
            table.append((sum(ms) / steps, cs, ft, [m / steps for m in ms]))
            del ens
    best = min(table, key=lambda t: t[0])
    return dict(cell_size=best[1], force_threads=best[2], ms_per_step=best[0], table=table)
