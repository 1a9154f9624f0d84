"""Autotuner of the step: picks the block-list cell size and the force
kernel's tile shape (grid cells per workgroup) by timing candidates on the
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


# Default force-kernel candidates.  The LDS-tiled kernel (tile != (0, 0))
# is not among them: it stages a region's neighbour records through LDS
# with a block-wide scan, a binary search per staged record and barriers
# between the phases, and the pair arithmetic it then runs is the same as
# the agent-parallel kernel's, which already finds its candidates' records
# in L1 (96 % hits in cell order) and is bound by that arithmetic.  Timed on
# the 256-scenario ensemble it took 2.3 to 6.3 ms per step against 0.75 ms,
# so it only costs tuning time; it stays selectable by passing a tile.
ENSEMBLE_KERNELS = ((0, 0, 4), (0, 0, 8), (0, 0, 4, 10), (0, 0, 8, 10))
ROOM_KERNELS = ((0, 0, 4), (0, 0, 8), (0, 0, 16))


def tune_ensemble(scenarios, precision='f32', cell_sizes=None, kernels=ENSEMBLE_KERNELS,
                  steps=10, warmup=3, flags=None):
    """Time every (cell size, force kernel) pair over `steps` batched steps
    (after `warmup` steps) and return the fastest.

    A force kernel is (tile_x, tile_y, lanes[, order_every]): tile (0, 0)
    selects the agent-parallel kernel with `lanes` lanes per agent, serving
    the agents in index order, or in cell order rebuilt every `order_every`
    steps; otherwise LDS tiles of tile_x * tile_y cells (not a default
    candidate, see ENSEMBLE_KERNELS).  The choice never changes results (pair
    and wall terms are summed exactly).  Returns
    dict(cell_size, tile, lanes, order_every, ms_per_step, table) where table
    lists (ms_per_step, cell_size, kernel, per-kernel ms)."""
    if cell_sizes is None:
        cell_sizes = neutral_cell_sizes(max_body_radius(scenarios))
    kw = {} if flags is None else dict(flags=flags)
    table = []
    for cs in cell_sizes:
        for kern in kernels:
            ens = Ensemble(scenarios, precision=precision, cell_size=cs, tile=kern[:2],
                           lanes=kern[2], order_every=kern[3] if len(kern) > 3 else 0, **kw)
            ens.step(warmup)
            ms = ens.timed(steps)
            ens.check()
            table.append((ms[3] / steps, cs, tuple(kern), [m / steps for m in ms]))
            del ens
    best = min(table, key=lambda t: t[0])
    return dict(cell_size=best[1], tile=best[2][:2], lanes=best[2][2],
                order_every=best[2][3] if len(best[2]) > 3 else 0, ms_per_step=best[0],
                table=table)


def tune_kernel(make_stepper, kernels=ROOM_KERNELS, steps=20, warmup=5):
    """Fastest force kernel (tile_x, tile_y, lanes) for one simulation:
    make_stepper(tile, lanes) builds a Stepper on a fresh copy of the state.
    Returns (kernel, table)."""
    table = []
    for kern in kernels:
        st = make_stepper(kern[:2], kern[2])
        st.step(warmup)
        ms = st.timed(steps)
        st.check()
        table.append((ms[3] / steps, tuple(kern), [m / steps for m in ms]))
        del st
    best = min(table, key=lambda t: t[0])
    return best[1], table
