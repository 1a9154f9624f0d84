"""Per-timestep operators with the reference's names, running on the device.

Each function takes a :class:`~crowddynamics_amd.device.DeviceAgents` and
enqueues HIP kernels on the current torch stream (no host synchronisation).
The reference functions they replace are cited per function; they mutate the
agent state in place exactly like the numba originals.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .device import DeviceAgents, Workspace, stream_handle
from .exceptions import InvalidType, InvalidValue

#: obstacle_type_linear (crowddynamics/core/structures.py:6-9)
obstacle_type_linear = np.dtype([('p0', np.float64, 2), ('p1', np.float64, 2)])


def _check_agents(agents):
    if not isinstance(agents, DeviceAgents):
        raise InvalidType('expected DeviceAgents')


def _on_agents_device(t, agents, what):
    """`t` as a contiguous fp64 tensor on the agents' device.  A host
    pointer or a tensor of another dtype must never reach a kernel (the
    kernels read fp64 device memory), so other tensors are converted here;
    anything that is not a tensor of real numbers raises InvalidValue."""
    if not isinstance(t, torch.Tensor):
        raise InvalidValue('%s must be a torch.Tensor' % what)
    if t.is_complex() or t.dtype == torch.bool:
        raise InvalidValue('%s must hold real numbers, not %s' % (what, t.dtype))
    dev = agents.device
    if dev.type == 'cuda' and dev.index is None:
        dev = torch.device('cuda', torch.cuda.current_device())
    return t.to(device=dev, dtype=torch.float64).contiguous()


def segments_to_device(obstacles, device='cuda'):
    """obstacle_type_linear[W] (or [W, 4]) -> fp64 [W, 4] device tensor."""
    if isinstance(obstacles, torch.Tensor):
        t = obstacles.to(device=device, dtype=torch.float64).reshape(-1, 4).contiguous()
        return t
    s = np.asarray(obstacles)
    if s.dtype.names:
        arr = np.zeros((len(s), 4))
        arr[:, 0:2] = s['p0']
        arr[:, 2:4] = s['p1']
    else:
        arr = np.asarray(s, dtype=np.float64).reshape(-1, 4)
    return torch.from_numpy(np.ascontiguousarray(arr)).to(device)


def add_to_cells(agents, cell_size, workspace):
    """cell_lists.add_to_cells (called at crowddynamics/core/interactions.py:192-193).
    Results stay in `workspace` (see Workspace.block_list)."""
    _check_agents(agents)
    _lib.check(_lib.lib().cd_add_to_cells(agents.ref(), float(cell_size), workspace.ref(),
                                          stream_handle()), 'add_to_cells')


def agent_agent_block_list(agents, cell_size, workspace):
    """agent_agent_block_list (crowddynamics/core/interactions.py:191-205)."""
    _check_agents(agents)
    _lib.check(_lib.lib().cd_agent_agent_block_list(
        agents.ref(), float(cell_size), workspace.ref(), stream_handle()),
        'agent_agent_block_list')


def agent_obstacle(agents, obstacles):
    """agent_obstacle (crowddynamics/core/interactions.py:208-214); `obstacles`
    is a [W, 4] fp64 device tensor (segments_to_device)."""
    _check_agents(agents)
    if not isinstance(obstacles, torch.Tensor):
        obstacles = segments_to_device(obstacles, agents.device)
    if obstacles.dim() == 3 and tuple(obstacles.shape[1:]) == (2, 2):
        obstacles = obstacles.reshape(-1, 4)
    if obstacles.dim() != 2 or obstacles.shape[1] != 4:
        raise InvalidValue('obstacles must be [W, 4] (p0x, p0y, p1x, p1y), got %s'
                           % (tuple(obstacles.shape),))
    obstacles = _on_agents_device(obstacles, agents, 'obstacles')
    n_seg = obstacles.shape[0]
    ptr = C.c_void_p(obstacles.data_ptr() if n_seg else 0)
    _lib.check(_lib.lib().cd_agent_obstacle(agents.ref(), ptr, n_seg, stream_handle()),
               'agent_obstacle')


def force_adjust_agents(agents):
    """force_adjust_agents (crowddynamics/core/motion/adjusting.py:98-108)."""
    _check_agents(agents)
    _lib.check(_lib.lib().cd_force_adjust_agents(agents.ref(), stream_handle()),
               'force_adjust_agents')


def torque_adjust_agents(agents):
    """torque_adjust_agents (crowddynamics/core/motion/adjusting.py:111-121)."""
    _check_agents(agents)
    _lib.check(_lib.lib().cd_torque_adjust_agents(agents.ref(), stream_handle()),
               'torque_adjust_agents')


def orient_towards_target_direction(agents):
    """orient_towards_target_direction (crowddynamics/core/steering/orientation.py:17-21)."""
    _check_agents(agents)
    _lib.check(_lib.lib().cd_orient_towards_target_direction(agents.ref(), stream_handle()),
               'orient_towards_target_direction')


def navigation(agents, maps, minx, miny, step, mode=0):
    """Navigation.update (crowddynamics/simulation/logic.py:149-165) with
    getdefault (crowddynamics/core/steering/navigation.py:67-78).

    mode 0: `maps` is a [T, rows, cols, 2] fp64 device tensor of direction
    maps read at the truncated grid index (reference behaviour).  mode 1:
    `maps` is [T, rows, cols] potentials read by bilinear gradient."""
    _check_agents(agents)
    if mode not in (0, 1):
        raise InvalidValue('navigation mode must be 0 (nearest node) or 1 (bilinear)')
    if not isinstance(maps, torch.Tensor):
        maps = torch.as_tensor(np.asarray(maps, dtype=np.float64))
    if mode == 0 and (maps.dim() != 4 or maps.shape[-1] != 2):
        raise InvalidValue('direction maps must be [T, rows, cols, 2]')
    if mode == 1 and maps.dim() != 3:
        raise InvalidValue('potentials must be [T, rows, cols]')
    if min(maps.shape[:3]) < 1 or not step > 0:
        raise InvalidValue('navigation maps must be non-empty with a positive step')
    maps = _on_agents_device(maps, agents, 'navigation maps')
    T_, rows, cols = maps.shape[:3]
    _lib.check(_lib.lib().cd_navigation(agents.ref(), C.c_void_p(maps.data_ptr()), T_, rows,
                                        cols, float(minx), float(miny), float(step), int(mode),
                                        stream_handle()), 'navigation')


def force_fluctuation(agents, seed, step):
    """Fluctuation.update (crowddynamics/simulation/logic.py:78-86): adds the
    force (and torque) noise of `step` from the seeded counter-based stream."""
    _check_agents(agents)
    _lib.check(_lib.lib().cd_force_fluctuation(agents.ref(), int(seed), int(step),
                                               stream_handle()), 'force_fluctuation')


def velocity_verlet_integrator(agents, dt_min, dt_max, workspace):
    """velocity_verlet_integrator (crowddynamics/core/integrator.py:209-256).
    dt stays on the device (Workspace.state().dt)."""
    _check_agents(agents)
    _lib.check(_lib.lib().cd_velocity_verlet_integrator(
        agents.ref(), float(dt_min), float(dt_max), workspace.ref(), stream_handle()),
        'velocity_verlet_integrator')


def reset(agents):
    """Reset.update (crowddynamics/simulation/logic.py:59-64)."""
    _check_agents(agents)
    _lib.check(_lib.lib().cd_reset(agents.ref(), stream_handle()), 'reset')
