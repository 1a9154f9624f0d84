"""Batched ensemble: independent scenarios of uneven agent counts advanced
together, two kernel launches per step over all of them (capturable in a
HIP graph).

The reference runs one simulation per process
(crowddynamics/simulation/multiagent.py:199-236, ``MultiAgentSimulation.run``
on one ``Agents.array``).  ``Ensemble`` concatenates the agent records of
many scenarios into one device array; each scenario keeps its own grid cells,
walls, adaptive dt, elapsed time and fluctuation stream, so every scenario
evolves exactly as its standalone ``Stepper`` run with the same parameters.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .core import segments_to_device
from .device import (DeviceAgents, Workspace, cell_capacity_for, cell_slots_for,
                     min_body_radius, stream_handle)
from .exceptions import InvalidType, InvalidValue
from .stepper import ALL_NODES, Stepper, make_params


def _bounds(scn):
    p = scn.agents['position']
    if scn.bounds is not None:
        return scn.bounds
    if len(p) == 0:
        return (0.0, 0.0, 1.0, 1.0)
    return (*p.min(0), *p.max(0))


class Ensemble(Stepper):
    """Scenarios (objects with ``agents``, ``segments``, ``dt_min``,
    ``dt_max``, ``seed``, ``bounds``, ``nav`` and ``cell_size``, e.g.
    :class:`crowddynamics_amd.examples.Scenario`) stepped as one batch.

    All scenarios share the agent model, precision, node flags, cell size and
    navigation maps (``nav`` of every scenario is the same object, or None);
    they differ in agents, walls, dt bounds and seed.  ``cell_size`` defaults
    to the scenarios' common value.
    """

    _entry = 'cd_batch_step'

    def __init__(self, scenarios, precision='f32', flags=ALL_NODES, cell_size=None,
                 cell_slots=None, tile=None, lanes=0, order_every=0, first_step=0, device=None,
                 margin=0.5):
        scenarios = list(scenarios)
        if not scenarios:
            raise InvalidValue('an ensemble needs at least one scenario')
        dtype = scenarios[0].agents.dtype
        if any(s.agents.dtype != dtype for s in scenarios):
            raise InvalidType('all scenarios of an ensemble hold the same agent model')
        navs = [s.nav for s in scenarios]
        if any(v is not navs[0] for v in navs):
            raise InvalidValue('the scenarios of an ensemble share one navigation map set')
        if cell_size is None:
            sizes = {float(s.cell_size) for s in scenarios}
            if len(sizes) != 1:
                raise InvalidValue('scenarios disagree on cell_size; pass cell_size=')
            cell_size = sizes.pop()
        self.scenarios = scenarios
        self.S = len(scenarios)
        self.counts = np.array([len(s.agents) for s in scenarios], dtype=np.int64)
        self.offsets = np.concatenate([[0], np.cumsum(self.counts)]).astype(np.int64)
        array = np.concatenate([s.agents for s in scenarios])
        self.agents = DeviceAgents(array, precision=precision, device=device)
        dev = self.agents.device
        self.n = self.agents.n
        self.cell_size = float(cell_size)

        segs = [np.asarray(segments_to_device(s.segments, 'cpu').numpy())
                if s.segments is not None and len(s.segments) else np.zeros((0, 4))
                for s in scenarios]
        self.seg_offsets = np.concatenate([[0], np.cumsum([len(g) for g in segs])])
        allsegs = np.concatenate(segs) if segs else np.zeros((0, 4))
        self.segments = torch.from_numpy(np.ascontiguousarray(allsegs)).to(dev)
        if len(allsegs) == 0:
            flags &= ~_lib.CD_USE_OBSTACLE
        nav = navs[0]
        self.nav = nav
        if nav is not None:
            self.maps = torch.as_tensor(np.ascontiguousarray(nav['maps']), dtype=torch.float64,
                                        device=dev)
        else:
            self.maps = None
            flags &= ~_lib.CD_USE_NAV

        caps = []
        for s in scenarios:
            x0, y0, x1, y1 = _bounds(s)
            caps.append(cell_capacity_for(x1 - x0, y1 - y0, self.cell_size, margin)
                        if len(s.agents) else 1)
        self.cell_offsets = np.concatenate([[0], np.cumsum(caps)]).astype(np.int64)
        if cell_slots is None:
            cell_slots = max(cell_slots_for(self.cell_size, min_body_radius(s.agents))
                             for s in scenarios)
        self.ws = Workspace(self.n, int(self.cell_offsets[-1]), self.agents.model,
                            cell_slots=cell_slots, device=dev, max_scenarios=self.S,
                            tile=tile, lanes=lanes, order_every=order_every)
        self.table = (_lib.CDScenario * self.S)()
        for k, s in enumerate(scenarios):
            t = self.table[k]
            t.agent_begin, t.agent_end = int(self.offsets[k]), int(self.offsets[k + 1])
            t.cell_begin, t.cell_capacity = int(self.cell_offsets[k]), int(caps[k])
            t.seg_begin, t.seg_end = int(self.seg_offsets[k]), int(self.seg_offsets[k + 1])
            t.seed = int(s.seed)
            t.dt_min, t.dt_max = float(s.dt_min), float(s.dt_max)
        self.params = make_params(self.cell_size, 0.0, 0.0, flags, 0, self.segments,
                                  self.maps, nav)
        self.flags = flags
        self.prepare(first_step)

    def prepare(self, first_step=0):
        """Upload the scenario table and seed the per-scenario reductions
        from the current state (synchronous)."""
        _lib.check(_lib.lib().cd_batch_prepare(self.agents.ref(), C.byref(self.params),
                                               self.table, self.S, self.ws.ref(),
                                               int(first_step), stream_handle()),
                   'batch_prepare')

    def state(self):
        """State of scenario 0 with the batch's status and step counter."""
        return self.ws.states(1)[0]

    def states(self):
        """Per-scenario state scalars (dt, time_tot, grid origin and shape)."""
        return self.ws.states(self.S)

    def download_scenarios(self):
        """Per-scenario host copies of the agent records."""
        arr = np.concatenate([s.agents for s in self.scenarios])
        self.download(arr)
        return [arr[self.offsets[k]:self.offsets[k + 1]].copy() for k in range(self.S)]
