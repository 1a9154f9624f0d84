"""Simulation objects with the reference's interfaces, running on the GPU.

* ``LogicNodeBase``: tree DSL of crowddynamics/simulation/base.py:27-84
  (``<<`` returns the left operand, ``__getitem__`` finds a node by name,
  ``inject_before`` / ``inject_after`` / ``add_children``).
* ``Agents``: owns the agent records (crowddynamics/simulation/agents.py:605-680).
  ``agents.array`` is the reference's host structured array; the step runs on
  a device mirror and the array is synchronised lazily: reading it after a
  step downloads the state, and a read also marks it for re-upload before
  the next step, so in-place edits keep the reference's semantics.
* ``Field``: obstacles as obstacle_type_linear segments, targets and the
  precomputed navigation fields (crowddynamics/simulation/field.py:18-164
  builds them with shapely and scikit-fmm at setup).
* ``MultiAgentSimulation``: update() runs the logic tree in post-order
  (crowddynamics/simulation/multiagent.py:51-55).  When the tree consists of
  the standard nodes in a compatible order, the whole update runs as one
  fused cd_step (six kernels); otherwise node by node.
"""
from __future__ import annotations

from collections import OrderedDict

import numpy as np

from .agents import AgentModelToType, is_model, model_of
from .core import obstacle_type_linear
from .exceptions import CrowdDynamicsException, InvalidType


class LogicNodeBase:
    """Tree node with the reference's composition operators."""

    def __init__(self, name=None):
        self.name = name or self.__class__.__name__
        self._parent = None
        self.children = []

    @property
    def parent(self):
        return self._parent

    @parent.setter
    def parent(self, node):
        if self._parent is not None:
            self._parent.children.remove(self)
        self._parent = node
        if node is not None:
            node.children.append(self)

    @property
    def root(self):
        n = self
        while n._parent is not None:
            n = n._parent
        return n

    def update(self):
        raise NotImplementedError

    def inject_before(self, node):
        parent = self.parent
        self.parent = node
        node.parent = parent

    def inject_after(self, node):
        for child in list(self.children):
            child.parent = node
        node.parent = self

    def add_children(self, node):
        node.parent = self
        return self

    def __lshift__(self, other):
        if isinstance(other, LogicNodeBase):
            self.add_children(other)
        else:
            for o in other:
                self.add_children(o)
        return self

    def __repr__(self):
        return self.name

    def __getitem__(self, item):
        for node in pre_order(self.root):
            if node.name == item:
                return node
        raise KeyError('Key: "{}" not in the tree.'.format(item))


def pre_order(node):
    yield node
    for c in node.children:
        yield from pre_order(c)


def post_order(node):
    for c in node.children:
        yield from post_order(c)
    yield node


class Agents:
    """Agent records plus their device mirror."""

    def __init__(self, agent_type='circular', array=None, precision='f64'):
        if isinstance(agent_type, str):
            if agent_type not in AgentModelToType:
                raise InvalidType('unknown agent type %r' % agent_type)
            dtype = AgentModelToType[agent_type]
        else:
            dtype = np.dtype(agent_type)
        self._array = np.zeros(0, dtype=dtype) if array is None else array
        model_of(self._array)
        self.precision = precision
        self.device = None
        self._host_stale = False    # device holds newer state
        self._host_touched = False  # host array handed out since the last upload
        self.on_download = []       # callables(array) run after a download

    @property
    def array(self):
        if self._host_stale and self.device is not None:
            import torch
            torch.cuda.current_stream().synchronize()
            self.device.download(self._array)
            self._host_stale = False
            for fn in self.on_download:
                fn(self._array)
        self._host_touched = True
        return self._array

    @array.setter
    def array(self, value):
        model_of(value)
        self._array = value
        self.device = None
        self._host_stale = False
        self._host_touched = True

    def add(self, records):
        """Append records (an alternative to add_non_overlapping_group, whose
        spawn sampling is setup-time code)."""
        arr = self.array
        if records.dtype != arr.dtype:
            raise CrowdDynamicsException('agent type mismatch')
        self.array = np.concatenate((arr, records))

    def ensure_device(self):
        """Device mirror, uploaded when the host array may have changed.
        Returns True when an upload happened."""
        from .device import DeviceAgents
        if self.device is None or self.device.n != len(self._array):
            self.device = DeviceAgents(self._array, precision=self.precision)
            self._host_touched = False
            return True
        if self._host_touched:
            self.device.upload(self._array)
            self._host_touched = False
            return True
        return False

    def mark_device_newer(self):
        self._host_stale = True


class Field:
    """Obstacles, targets and navigation fields of a simulation.

    `obstacles`: obstacle_type_linear[W] (or [W, 4]); `targets`: number of
    targets (or a list); navigation: per target either a direction map
    [rows, cols, 2] (reference, nearest node) or a potential [rows, cols]
    (bilinear gradient), on a grid with origin (minx, miny) and spacing step.
    `domain`: polygon vertices [K, 2] for InsideDomain (optional)."""

    def __init__(self, obstacles=None, targets=0, domain=None, navigation=None,
                 nav_origin=(0.0, 0.0), nav_step=0.1, nav_mode=0):
        if obstacles is None:
            obstacles = np.zeros(0, dtype=obstacle_type_linear)
        self.obstacles = obstacles
        self.targets = list(range(targets)) if isinstance(targets, int) else list(targets)
        self.domain = None if domain is None else np.asarray(domain, dtype=np.float64)
        #: polygon targets (vertex arrays [k, 2]) for TargetReached, by target index
        self.target_polygons = {}
        self.navigation = navigation
        self.nav_origin = tuple(nav_origin)
        self.nav_step = float(nav_step)
        self.nav_mode = int(nav_mode)
        self._dev = {}

    def device_segments(self):
        from .core import segments_to_device
        if 'segments' not in self._dev:
            self._dev['segments'] = segments_to_device(self.obstacles)
        return self._dev['segments']

    def device_navigation(self):
        if self.navigation is None:
            return None
        if 'nav' not in self._dev:
            import torch
            self._dev['nav'] = torch.as_tensor(np.ascontiguousarray(self.navigation),
                                               dtype=torch.float64, device='cuda')
        return self._dev['nav']

    def nav_dict(self):
        if self.navigation is None:
            return None
        return dict(maps=self.navigation, minx=self.nav_origin[0], miny=self.nav_origin[1],
                    step=self.nav_step, mode=self.nav_mode)


class SimulationData(OrderedDict):
    """simulation.data whose 'dt' and 'time_tot' live on the device and are
    read (synchronously) only when accessed."""

    def __init__(self, *a, **kw):
        super().__init__(*a, **kw)
        self._source = None
        self._time_offset = 0.0
        self._device_values = {}

    def bind(self, source):
        """source() -> cd_state of the workspace that integrates."""
        self._source = source

    def bind_value(self, key, fn):
        """data[key] is read (synchronously) from the device by fn()."""
        self._device_values[key] = fn
        if key not in self:
            super().__setitem__(key, 0)

    def __getitem__(self, key):
        if key in ('dt', 'time_tot') and self._source is not None:
            st = self._source()
            if st is not None:
                return float(st.dt) if key == 'dt' else self._time_offset + float(st.time_tot)
        if key in self._device_values:
            return self._device_values[key]()
        return super().__getitem__(key)


class MultiAgentSimulation:
    """Post-order update of a LogicNode tree over device-resident agents."""

    def __init__(self, field=None, agents=None, logic=None, name=None, fused='auto',
                 cell_capacity=None):
        self.name = name or self.__class__.__name__
        self.field = field if field is not None else Field()
        self.agents = agents
        self.logic = logic
        self.fused = fused
        self.cell_capacity = cell_capacity
        self.data = SimulationData()
        self.data['iterations'] = 0
        self.data['time_tot'] = 0.0
        self.data['dt'] = 0.0
        self.exit_condition = None
        self._plan = None
        self._stepper = None
        self._integrator_ws = None

    # -- execution ---------------------------------------------------------

    def _make_plan(self):
        from . import logic as L
        order = list(post_order(self.logic.root))
        plan = dict(order=order, fused_flags=None)
        if self.fused is False:
            return plan
        fusable = [L.Fluctuation, L.Navigation, L.Orientation, L.Adjusting,
                   L.AgentAgentInteractions, L.AgentObstacleInteractions, L.Integrator, L.Reset]
        state = (L.InsideDomain, L.TargetReached)
        seq = [n for n in order if not isinstance(n, state)]
        kinds = [type(n) for n in seq]
        if any(k not in fusable for k in kinds) or len(set(kinds)) != len(kinds):
            return plan
        pos = [fusable.index(k) for k in kinds]
        if pos != sorted(pos) or L.Integrator not in kinds:
            return plan
        model = model_of(self.agents._array)
        has = set(kinds)
        if model == 1 and ((L.Orientation in has) != (L.Adjusting in has)):
            return plan
        from . import _lib
        flags = 0
        flags |= _lib.CD_USE_FLUCT if L.Fluctuation in has else 0
        flags |= _lib.CD_USE_NAV if L.Navigation in has else 0
        flags |= _lib.CD_USE_ADJUST if L.Adjusting in has else 0
        flags |= _lib.CD_USE_AGENT_AGENT if L.AgentAgentInteractions in has else 0
        flags |= _lib.CD_USE_OBSTACLE if L.AgentObstacleInteractions in has else 0
        flags |= _lib.CD_USE_RESET if L.Reset in has else 0
        plan['fused_flags'] = flags
        plan['nodes'] = {type(n): n for n in seq}
        # state nodes read positions only: after the fused step (post-order)
        plan['inside'] = [n for n in order if isinstance(n, state)]
        return plan

    def _fused_stepper(self):
        from . import logic as L
        from .stepper import Stepper
        nodes = self._plan['nodes']
        integ = nodes[L.Integrator]
        aa = nodes.get(L.AgentAgentInteractions)
        nav = self.field.nav_dict() if L.Navigation in nodes else None
        seg = self.field.obstacles if L.AgentObstacleInteractions in nodes else None
        st = Stepper(self.agents._array, segments=seg, nav=nav,
                     cell_size=aa.cell_size if aa else 1.0, dt_min=integ.dt_min,
                     dt_max=integ.dt_max, seed=nodes[L.Fluctuation].seed
                     if L.Fluctuation in nodes else 0, flags=self._plan['fused_flags'],
                     precision=self.agents.precision, cell_capacity=self.cell_capacity,
                     first_step=self.data['iterations'], device_agents=self.agents.device)
        return st

    def update(self):
        """Execute one iteration: every node's update() in post-order."""
        if self._plan is None:
            self._plan = self._make_plan()
        uploaded = self.agents.ensure_device()
        if self._plan['fused_flags'] is not None:
            if self._stepper is None:
                self._stepper = self._fused_stepper()
                base = float(super(SimulationData, self.data).__getitem__('time_tot'))
                self.data._time_offset = base
                self.data.bind(lambda: self._stepper.ws.state())
            elif uploaded:
                self._stepper.prepare(self.data['iterations'])
            self._stepper.step(1)
            for n in self._plan['inside']:
                n.update()
        else:
            for node in self._plan['order']:
                node.update()
        self.agents.mark_device_newer()
        OrderedDict.__setitem__(self.data, 'iterations',
                                OrderedDict.__getitem__(self.data, 'iterations') + 1)

    def run(self, steps=None):
        """Update until exit_condition(self) is true (base.py:123-126) or for
        `steps` iterations."""
        if steps is not None:
            for _ in range(steps):
                self.update()
            return
        while self.exit_condition is None or not self.exit_condition(self):
            self.update()

    def check(self):
        """Raise if the device recorded an overflow or a non-finite position."""
        if self._stepper is not None:
            return self._stepper.check()
        return None
