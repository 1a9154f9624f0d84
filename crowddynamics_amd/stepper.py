"""Fused multi-step driver over cd_step (two kernels per step) with HIP
graph capture.

``Stepper`` runs one MultiAgentSimulation.update() of the standard logic
tree (crowddynamics/examples/simulations.py:123-136:
Reset << (Integrator << (Fluctuation, Adjusting << (Navigation, Orientation),
AgentAgentInteractions, AgentObstacleInteractions)), in the reference's
post-order, crowddynamics/simulation/multiagent.py:51-55) per ``step()``.
All step state (dt, time_tot, the fluctuation counter, the reductions that
feed the next step's binning and dt) stays on the device, so a captured
graph of K steps replays without host involvement.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .core import segments_to_device
from .device import (DeviceAgents, Workspace, cell_capacity_for, cell_slots_for,
                     min_body_radius, stream_handle)
from .exceptions import CrowdDynamicsException, InvalidValue

ALL_NODES = (_lib.CD_USE_FLUCT | _lib.CD_USE_NAV | _lib.CD_USE_ADJUST |
             _lib.CD_USE_AGENT_AGENT | _lib.CD_USE_OBSTACLE | _lib.CD_USE_RESET)


class GridOverflow(CrowdDynamicsException):
    """The binning grid outgrew the workspace's cell capacity."""


def make_params(cell_size, dt_min, dt_max, flags, seed, segments, maps, nav):
    """cd_step_params_t of a step; `segments` [W, 4] and `maps` are device
    tensors that must outlive the parameters."""
    p = _lib.CDStepParams()
    p.cell_size = float(cell_size)
    p.dt_min = float(dt_min)
    p.dt_max = float(dt_max)
    p.flags = int(flags)
    p.seed = int(seed)
    p.segments = C.c_void_p(segments.data_ptr() if len(segments) else 0)
    p.n_segments = len(segments)
    if maps is not None:
        p.nav_mode = int(nav.get('mode', 0))
        p.nav_maps = C.c_void_p(maps.data_ptr())
        p.nav_targets, p.nav_rows, p.nav_cols = maps.shape[:3]
        p.nav_minx = float(nav['minx'])
        p.nav_miny = float(nav['miny'])
        p.nav_step = float(nav['step'])
    return p


#: the fused step bins the walls into cells of this side (m) when a scenario
#: has at least WALL_GRID_MIN_SEGMENTS segments (results are the same either way)
WALL_GRID_CELL = 1.0
WALL_GRID_MIN_SEGMENTS = 32


def wall_reach(array):
    """Largest distance from an agent's centre at which a wall contact can
    occur: the radius of circular agents, r_ts + max(r_t, r_s) of
    three-circle agents."""
    if len(array) == 0:
        return 0.0
    if 'position_ls' in array.dtype.names:
        return float(np.max(array['r_ts'] + np.maximum(array['r_t'], array['r_s'])))
    return float(np.max(array['radius']))


def build_wall_grid(params, segments, reach, cell):
    """Bins the segments ([W, 4] tensor) into cells of side `cell`
    (cd_wall_grid_build) and points `params` at the device copy, which the
    caller keeps alive as long as the parameters."""
    L = _lib.lib()
    seg = np.ascontiguousarray(segments.detach().cpu().numpy(), dtype=np.float64)
    ptr = C.c_void_p(seg.ctypes.data)
    need = L.cd_wall_grid_build(ptr, len(seg), float(reach), float(cell), None, 0,
                                C.byref(params))
    if need < 0:
        _lib.check(int(need), 'wall_grid_build')
    out = np.zeros(int(need), dtype=np.int32)
    got = L.cd_wall_grid_build(ptr, len(seg), float(reach), float(cell),
                               C.c_void_p(out.ctypes.data), int(need), C.byref(params))
    if got != need:
        _lib.check(int(got) if got < 0 else _lib.CD_ERR_ARG, 'wall_grid_build')
    dev = torch.from_numpy(out).to(segments.device)
    params.wall_cells = C.c_void_p(dev.data_ptr())
    return dev


class Stepper:
    def __init__(self, array, segments=None, nav=None, cell_size=3.6, dt_min=0.01,
                 dt_max=0.01, seed=1, flags=ALL_NODES, precision='f64', cell_capacity=None,
                 bounds=None, first_step=0, device=None, device_agents=None,
                 cell_slots=None, force_threads=0, tile=None, lanes=0, order_every=0,
                 wall_grid='auto'):
        if device_agents is not None:
            self.agents = device_agents
        else:
            self.agents = DeviceAgents(array, precision=precision, device=device)
        dev = self.agents.device
        self.n = self.agents.n
        if segments is not None and len(segments):
            self.segments = segments_to_device(segments, dev)
        else:
            self.segments = torch.zeros((0, 4), dtype=torch.float64, device=dev)
            flags &= ~_lib.CD_USE_OBSTACLE
        self.nav = nav
        if nav is not None:
            self.maps = torch.as_tensor(np.ascontiguousarray(nav['maps']), dtype=torch.float64,
                                        device=dev)
        else:
            self.maps = None
            flags &= ~_lib.CD_USE_NAV
        if cell_capacity is None:
            if bounds is not None:
                x0, y0, x1, y1 = bounds
            else:
                p = array['position']
                x0, y0 = p.min(0) if len(p) else (0.0, 0.0)
                x1, y1 = p.max(0) if len(p) else (1.0, 1.0)
            cell_capacity = cell_capacity_for(x1 - x0, y1 - y0, cell_size)
        if cell_slots is None:
            cell_slots = cell_slots_for(cell_size, min_body_radius(array))
        self.ws = Workspace(self.n, cell_capacity, self.agents.model, cell_slots=cell_slots,
                            device=dev, force_threads=force_threads, tile=tile, lanes=lanes,
                            order_every=order_every)
        self.params = make_params(cell_size, dt_min, dt_max, flags, seed, self.segments,
                                  self.maps, nav)
        self.wall_cells = None
        if wall_grid == 'auto':
            wall_grid = WALL_GRID_CELL if len(self.segments) >= WALL_GRID_MIN_SEGMENTS else None
        if wall_grid and (flags & _lib.CD_USE_OBSTACLE):
            self.wall_cells = build_wall_grid(self.params, self.segments, wall_reach(array),
                                              float(wall_grid))
        self.flags = flags
        self.prepare(first_step)

    def prepare(self, first_step=0):
        """Seed the device reductions from the current state (also after any
        host-side change of positions or velocities)."""
        _lib.check(_lib.lib().cd_step_prepare(self.agents.ref(), C.byref(self.params),
                                              self.ws.ref(), int(first_step), stream_handle()),
                   'step_prepare')

    _entry = 'cd_step'

    def step(self, k=1):
        f = getattr(_lib.lib(), self._entry)
        s = stream_handle()
        for _ in range(k):
            _lib.check(f(self.agents.ref(), C.byref(self.params), self.ws.ref(), s), 'step')

    def capture(self, k):
        """HIP graph of k consecutive steps (launched on the current stream)."""
        if torch.cuda.current_stream() == torch.cuda.default_stream():
            raise InvalidValue('capture needs a non-default stream (use torch.cuda.stream)')
        L = _lib.lib()
        s = stream_handle()
        _lib.check(L.cd_graph_begin(s), 'graph_begin')
        try:
            self.step(k)
        except Exception:
            ex = C.c_void_p()
            L.cd_graph_end(s, C.byref(ex))
            if ex.value:
                L.cd_graph_destroy(ex)
            raise
        ex = C.c_void_p()
        _lib.check(L.cd_graph_end(s, C.byref(ex)), 'graph_end')
        return Graph(ex)

    def state(self):
        return self.ws.state()

    def timed(self, steps):
        """GPU milliseconds of the first kernel (per-agent terms and binning),
        the force kernel (forces and integration), an empty slot (0) and the
        whole step, summed over `steps` steps, from HIP events on the current
        stream."""
        ms = (C.c_double * 4)()
        _lib.check(_lib.lib().cd_step_timed(self.agents.ref(), C.byref(self.params),
                                            self.ws.ref(), int(steps), ms, stream_handle()),
                   'step_timed')
        return list(ms)

    def kernel_timed(self, kernel, reps=20):
        """Average GPU milliseconds of `reps` back-to-back launches of one
        step kernel ('bin': the first kernel, per-agent terms and binning;
        'forces': the force kernel, which also integrates).  Leaves the state
        meaningless: time a copy, or prepare() again afterwards."""
        k = {'bin': 0, 'forces': 1}[kernel]
        ms = C.c_double()
        _lib.check(_lib.lib().cd_kernel_timed(self.agents.ref(), C.byref(self.params),
                                              self.ws.ref(), k, int(reps), C.byref(ms),
                                              stream_handle()), 'kernel_timed')
        return ms.value

    def check(self):
        """Raise if the device recorded a grid overflow or non-finite positions."""
        st = self.state()
        if st.status & _lib.CD_STATUS_NONFINITE:
            raise CrowdDynamicsException('non-finite agent position')
        if st.status & _lib.CD_STATUS_SLOTS_FULL:
            import warnings
            warnings.warn('a cell held more agents than its %d slots; the exact slow path '
                          'was used' % self.ws.cell_slots)
        if st.status & _lib.CD_STATUS_GRID_OVERFLOW:
            raise GridOverflow('agents spread beyond the cell capacity (%d cells)'
                               % self.ws.cell_capacity)
        return st

    def download(self, array):
        torch.cuda.current_stream().synchronize()
        self.agents.download(array)
        return array


def capture_many(steppers, k):
    """One HIP graph of k rounds of one step of every stepper in turn (a
    per-scenario loop of plain steps, launched on the current stream)."""
    if torch.cuda.current_stream() == torch.cuda.default_stream():
        raise InvalidValue('capture needs a non-default stream (use torch.cuda.stream)')
    L = _lib.lib()
    s = stream_handle()
    _lib.check(L.cd_graph_begin(s), 'graph_begin')
    try:
        for _ in range(k):
            for st in steppers:
                st.step(1)
    except Exception:
        ex = C.c_void_p()
        L.cd_graph_end(s, C.byref(ex))
        if ex.value:
            L.cd_graph_destroy(ex)
        raise
    ex = C.c_void_p()
    _lib.check(L.cd_graph_end(s, C.byref(ex)), 'graph_end')
    return Graph(ex)


class Graph:
    """An instantiated HIP graph (cd_graph_* wrapper)."""

    def __init__(self, exec_ptr):
        self.ptr = exec_ptr

    def launch(self):
        _lib.check(_lib.lib().cd_graph_launch(self.ptr, stream_handle()), 'graph_launch')

    def __del__(self):
        if getattr(self, 'ptr', None) is not None and self.ptr.value:
            try:
                _lib.lib().cd_graph_destroy(self.ptr)
            except Exception:
                pass
            self.ptr = None
