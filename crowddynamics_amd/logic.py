"""LogicNodes with the reference's names and parameters
(crowddynamics/simulation/logic.py:31-387), each running its operator on the
device.  Trees are built exactly as in the reference, e.g. the Hallway tree
(crowddynamics/examples/simulations.py:123-136)::

    Reset(sim) << InsideDomain(sim) << (
        Integrator(sim) << (
            Fluctuation(sim),
            Adjusting(sim) << (Navigation(sim), Orientation(sim)),
            AgentAgentInteractions(sim),
            AgentObstacleInteractions(sim)))
"""
from __future__ import annotations

import numpy as np

from . import core
from .agents import is_model
from .device import Workspace, cell_capacity_for, cell_slots_for, min_body_radius
from .exceptions import InvalidType
from .simulation import LogicNodeBase


class LogicNode(LogicNodeBase):
    def __init__(self, simulation, name=None, **traits):
        super().__init__(name)
        self.simulation = simulation
        for k, v in traits.items():
            if not hasattr(self, k):
                raise TypeError('%s has no trait %r' % (self.__class__.__name__, k))
            setattr(self, k, v)

    @property
    def device_agents(self):
        a = self.simulation.agents
        a.ensure_device()
        return a.device

    def _model(self):
        arr = self.simulation.agents._array
        if is_model(arr, 'circular'):
            return 0
        if is_model(arr, 'three_circle'):
            return 1
        raise InvalidType


class Reset(LogicNode):
    """Zero force (and torque), logic.py:59-64."""

    def update(self):
        core.reset(self.device_agents)


class Integrator(LogicNode):
    """velocity_verlet_integrator, logic.py:67-75; dt and time_tot stay on the
    device and are read through simulation.data on access."""
    dt_min = 0.01
    dt_max = 0.01

    def update(self):
        d = self.device_agents
        if self.simulation._integrator_ws is None:
            self.simulation._integrator_ws = Workspace(d.n, 16, d.model, cell_slots=1)
            ws = self.simulation._integrator_ws
            data = self.simulation.data
            data._time_offset = float(dict.__getitem__(data, 'time_tot'))
            data.bind(ws.state)
        core.velocity_verlet_integrator(d, self.dt_min, self.dt_max,
                                        self.simulation._integrator_ws)


class Fluctuation(LogicNode):
    """Random force (and torque) of logic.py:78-86 from the counter-based
    stream keyed by `seed` and counted by (agent, iteration)."""
    seed = 0

    def update(self):
        core.force_fluctuation(self.device_agents, self.seed,
                               dict.__getitem__(self.simulation.data, 'iterations'))


class Adjusting(LogicNode):
    """force_adjust_agents (+ torque_adjust_agents), logic.py:89-94."""

    def update(self):
        d = self.device_agents
        core.force_adjust_agents(d)
        if self._model() == 1:
            core.torque_adjust_agents(d)


class AgentAgentInteractions(LogicNode):
    """agent_agent_block_list, logic.py:97-119.  As in the reference, the
    sight and force cut-offs are the fixed 3.0 m and 2 kN; cell_size defaults
    to sight_soc + 2 * max_agent_radius."""
    sight_soc = 3.0
    max_agent_radius = 0.3
    f_soc_max = 2e3
    cell_size = None

    def __init__(self, simulation, name=None, **traits):
        super().__init__(simulation, name, **traits)
        if self.cell_size is None:
            self.cell_size = self.sight_soc + 2 * self.max_agent_radius
        self._ws = None

    def update(self):
        d = self.device_agents
        if self._ws is None:
            cap = self.simulation.cell_capacity
            if cap is None:
                p = self.simulation.agents._array['position']
                span = p.max(0) - p.min(0) if len(p) else np.ones(2)
                cap = cell_capacity_for(span[0], span[1], self.cell_size)
            arr = self.simulation.agents._array
            self._ws = Workspace(d.n, cap, d.model,
                                 cell_slots=cell_slots_for(self.cell_size, min_body_radius(arr)))
        core.agent_agent_block_list(d, self.cell_size, self._ws)


class AgentObstacleInteractions(LogicNode):
    """agent_obstacle, logic.py:122-130 (segments converted once and kept on
    the device instead of every step)."""

    def update(self):
        core.agent_obstacle(self.device_agents, self.simulation.field.device_segments())


class Navigation(LogicNode):
    """Navigation.update, logic.py:135-165: target direction from the field's
    navigation map of each agent's target."""
    step = 0.1
    radius = 0.5
    strength = 0.3

    def update(self):
        f = self.simulation.field
        maps = f.device_navigation()
        if maps is None:
            return
        core.navigation(self.device_agents, maps, f.nav_origin[0], f.nav_origin[1],
                        f.nav_step, f.nav_mode)


class Orientation(LogicNode):
    """orient_towards_target_direction for three-circle agents, logic.py:258-261."""

    def update(self):
        if self._model() == 1:
            core.orient_towards_target_direction(self.device_agents)


class InsideDomain(LogicNode):
    """Sets agents not inside the domain inactive (logic.py:343-358): on the
    device, active = position inside the domain polygon (the crossings test
    of Path.contains_points), and data['inactive'] counts the changes (read
    from the device when accessed)."""

    def __init__(self, simulation, name=None, **traits):
        super().__init__(simulation, name, **traits)
        self._counter = None
        simulation.data['inactive'] = 0

    def update(self):
        dom = self.simulation.field.domain
        if dom is None:
            return
        import torch
        from . import state
        agents = self.device_agents
        if self._counter is None:
            self._counter = torch.zeros(1, dtype=torch.int64, device=agents.device)
            counter = self._counter
            self.simulation.data.bind_value('inactive', lambda: int(counter.item()))
        state.inside_domain(agents, dom, self._counter)


class TargetReached(LogicNode):
    """Detects agents that reached the polygon targets (logic.py:361-387):
    for each polygon target i of field.target_polygons, reached_by |= inside
    and data['target_<i>'] = the number of agents that reached it (read from
    the device when accessed).  The reference passes the whole record array
    to contains_points (logic.py:385); the positions are meant and used."""
    prefix = 'target_{index}'

    def __init__(self, simulation, name=None, **traits):
        super().__init__(simulation, name, **traits)
        self._reached = {}
        for i in simulation.field.target_polygons:
            simulation.data[self.prefix.format(index=i)] = 0

    def update(self):
        import torch
        from . import state
        agents = self.device_agents
        for i, poly in self.simulation.field.target_polygons.items():
            if i not in self._reached:
                reached = torch.zeros(agents.n, dtype=torch.uint8, device=agents.device)
                count = torch.zeros(1, dtype=torch.int64, device=agents.device)
                self._reached[i] = (reached, count)
                self.simulation.data.bind_value(self.prefix.format(index=i),
                                                lambda c=count: int(c.item()))
            reached, count = self._reached[i]
            state.target_reached(agents, poly, reached, count)
