"""Agent state, herding neighbours and snapshots on the device.

Operators with the reference's names and meaning, running in libcdamd.so
(csrc/cd_state.hip), so that a full logic tree keeps host synchronisation
off its step:

* inside_domain / target_reached: the InsideDomain and TargetReached nodes
  (crowddynamics/simulation/logic.py:343-387), point in polygon by the
  crossings test of matplotlib's Path.contains_points;
* find_nearest_neighbors / herding_interaction
  (core/steering/collective_motion.py:69-154): the nearest visible agents
  over the block list and the followers' herding direction;
* Snapshots: the agent records of selected steps kept on the device and
  written as save_npy does (crowddynamics/io.py:19-45): one .npy of the
  stacked records per dump.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np
import torch

from . import _lib
from .device import DeviceAgents, Workspace, stream_handle
from .exceptions import InvalidType, InvalidValue


def _dev_f64(a, device, shape_tail):
    t = torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=device)
    if t.dim() != 1 + len(shape_tail) or tuple(t.shape[1:]) != shape_tail:
        raise InvalidValue('expected an array of shape [k, %s]' % ', '.join(map(str, shape_tail)))
    return t.contiguous()


def _agents(agents):
    if not isinstance(agents, DeviceAgents):
        raise InvalidType('expected DeviceAgents')
    return agents


def inside_domain(agents, vertices, counter):
    """InsideDomain.update (logic.py:351-357): agents['active'] = inside the
    polygon `vertices` [k, 2]; counter (uint64 device tensor of one element)
    += the number of agents whose active flag changed."""
    a = _agents(agents)
    v = _dev_f64(vertices, a.device, (2,))
    _lib.check(_lib.lib().cd_inside_domain(
        C.c_void_p(a['position'].data_ptr()), a.n, C.c_void_p(v.data_ptr()), v.shape[0],
        C.c_void_p(a['active'].data_ptr()), C.c_void_p(counter.data_ptr()), stream_handle()),
        'inside_domain')


def target_reached(agents, vertices, reached_by, count):
    """TargetReached.update (logic.py:383-387) for one polygon target:
    reached_by (uint8 [n]) |= inside; count (uint64 [1]) = sum(reached_by)."""
    a = _agents(agents)
    v = _dev_f64(vertices, a.device, (2,))
    _lib.check(_lib.lib().cd_target_reached(
        C.c_void_p(a['position'].data_ptr()), a.n, C.c_void_p(v.data_ptr()), v.shape[0],
        C.c_void_p(reached_by.data_ptr()), C.c_void_p(count.data_ptr()), stream_handle()),
        'target_reached')


def find_nearest_neighbors(agents, sight, size_nearest_other, obstacles, workspace):
    """find_nearest_neighbors (collective_motion.py:69-107): [n, k] int64,
    the k nearest agents strictly within sight whose line of sight crosses
    no obstacle segment ([W, 4]), nearest first, -1 where missing.  The block
    list is built in `workspace` at cell size = sight."""
    a = _agents(agents)
    k = int(size_nearest_other)
    segs = _dev_f64(obstacles, a.device, (4,)) if obstacles is not None and len(obstacles) \
        else torch.zeros((0, 4), dtype=torch.float64, device=a.device)
    out = torch.empty((a.n, k), dtype=torch.int64, device=a.device)
    _lib.check(_lib.lib().cd_nearest_neighbors(
        a.ref(), float(sight), k, C.c_void_p(segs.data_ptr()) if len(segs) else None,
        len(segs), C.c_void_p(out.data_ptr()), workspace.ref(), stream_handle()),
        'find_nearest_neighbors')
    return out


def herding_interaction(agents, is_herding, neighbors, weight_position, phi):
    """herding_interaction (collective_motion.py:110-154): (direction [n, 2],
    has_direction [n] bool tensors) of the agents flagged in is_herding."""
    a = _agents(agents)
    if a.precision != 'f64':
        raise InvalidValue('herding runs on fp64 agent state')
    flags = torch.as_tensor(is_herding, device=a.device).to(torch.uint8).contiguous()
    nb = neighbors.contiguous()
    out = torch.empty((a.n, 2), dtype=torch.float64, device=a.device)
    has = torch.empty(a.n, dtype=torch.uint8, device=a.device)
    _lib.check(_lib.lib().cd_herding_interaction(
        C.c_void_p(a['position'].data_ptr()), C.c_void_p(a['velocity'].data_ptr()), a.n,
        C.c_void_p(flags.data_ptr()), C.c_void_p(nb.data_ptr()), nb.shape[1],
        float(weight_position), float(np.cos(phi)), C.c_void_p(out.data_ptr()),
        C.c_void_p(has.data_ptr()), stream_handle()), 'herding_interaction')
    return out, has.bool()


class Snapshots:
    """save_npy (crowddynamics/io.py:19-45) of device-resident agents.

    record() stores the current agent records in a device buffer (a copy on
    the current stream; no host synchronisation), dump() writes the stored
    records, stacked, to `<directory>/<basename>_<index>.npy` and starts a
    new buffer.  Each record is a copy of the state at its step (the
    reference appends the same array object each time, so its dumps repeat
    the last state)."""

    def __init__(self, agents, template, directory, basename='agents', capacity=64,
                 fields=None):
        self.agents = _agents(agents)
        self.template = np.array(template, copy=True)
        if len(self.template) != self.agents.n:
            raise InvalidValue('template has %d records for %d agents'
                               % (len(self.template), self.agents.n))
        self.directory = directory
        self.basename = basename
        self.capacity = int(capacity)
        names = self.template.dtype.names
        self.fields = [f for f in (fields or self.agents.fields) if f in self.agents.fields
                       and f in names]
        self.buf = {f: torch.empty((self.capacity,) + tuple(self.agents.fields[f].shape),
                                   dtype=self.agents.fields[f].dtype, device=self.agents.device)
                    for f in self.fields}
        self.count = 0
        self.index = 0
        os.makedirs(directory, exist_ok=True)

    def record(self):
        if self.count == self.capacity:
            raise InvalidValue('snapshot buffer full (%d); dump() first' % self.capacity)
        for f in self.fields:
            self.buf[f][self.count].copy_(self.agents.fields[f])
        self.count += 1

    def dump(self):
        """Writes the buffered records; returns the file path (None when empty)."""
        if self.count == 0:
            return None
        out = np.repeat(self.template[None, :], self.count, axis=0)
        for f in self.fields:
            host = self.buf[f][:self.count].to('cpu')
            if out.dtype[f] == np.bool_:
                out[f] = host.numpy().astype(bool)
            else:
                out[f] = host.to(torch.float64 if out[f].dtype.kind == 'f' else host.dtype).numpy()
        path = os.path.join(self.directory, '%s_%d.npy' % (self.basename, self.index))
        np.save(path, out)
        self.index += 1
        self.count = 0
        return path
