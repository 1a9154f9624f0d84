"""Device-resident agent state (structure of arrays) and workspace.

torch provides the HBM allocations and the stream; every computation runs in
libcdamd.so.  ``DeviceAgents`` mirrors a structured agent array
(crowddynamics/simulation/agents.py:605-680 keeps it as ``Agents.array``) as
one tensor per field; ``Workspace`` owns the block-list arrays, cell-sorted
copies and the device scalars (dt, time_tot, reductions, status).
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib
from .agents import (BOOL_FIELDS, INT_FIELDS, MUTABLE_FIELDS, REAL_FIELDS, SCALAR_FIELDS,
                     VECTOR_FIELDS, model_of)
from .exceptions import DeviceError, InvalidValue

_F64_ONLY_VECTORS = ('position', 'position_ls', 'position_rs')


def require_gpu():
    if not torch.cuda.is_available():
        raise DeviceError('no HIP device available: the agent update runs only on the GPU')


def stream_handle(stream=None):
    s = stream if stream is not None else torch.cuda.current_stream()
    return C.c_void_p(s.cuda_stream)


class DeviceAgents:
    """SoA mirror of an agent record array on a HIP device.

    precision 'f64' keeps every field fp64; 'f32' stores velocity,
    target_direction, force and force_prev in fp32 (position stays fp64).
    """

    def __init__(self, array, precision='f64', device=None):
        require_gpu()
        if precision not in ('f64', 'f32'):
            raise InvalidValue("precision must be 'f64' or 'f32'")
        self.model = model_of(array)
        self.precision = precision
        self.real = torch.float64 if precision == 'f64' else torch.float32
        self.device = torch.device(device) if device is not None else torch.device('cuda')
        self.n = len(array)
        self.fields = {}
        names = array.dtype.names
        for f in VECTOR_FIELDS:
            dt = self.real if f in REAL_FIELDS else torch.float64
            if f in names or f not in _F64_ONLY_VECTORS[1:]:
                self.fields[f] = torch.zeros((self.n, 2), dtype=dt, device=self.device)
        for f in SCALAR_FIELDS:
            if f in names:
                self.fields[f] = torch.zeros(self.n, dtype=torch.float64, device=self.device)
        for f in INT_FIELDS:
            self.fields[f] = torch.zeros(self.n, dtype=torch.int64, device=self.device)
        for f in BOOL_FIELDS:
            if f in names:
                self.fields[f] = torch.zeros(self.n, dtype=torch.uint8, device=self.device)
        self.upload(array)
        self.struct = _lib.CDAgents()
        s = self.struct
        s.n = self.n
        s.model = self.model
        s.real = 0 if precision == 'f64' else 1
        for f, t in self.fields.items():
            setattr(s, f, C.c_void_p(t.data_ptr()))

    def ref(self):
        return C.byref(self.struct)

    def upload(self, array, fields=None):
        """Copy fields of the host record array to the device."""
        if len(array) != self.n:
            raise InvalidValue('agent count changed (%d != %d)' % (len(array), self.n))
        for f, t in self.fields.items():
            if fields is not None and f not in fields:
                continue
            src = np.array(array[f], order='C', copy=True)
            if src.dtype == np.bool_:
                src = src.astype(np.uint8)
            t.copy_(torch.from_numpy(src).to(t.dtype))

    def download(self, array, fields=MUTABLE_FIELDS + ('active', 'target_reached')):
        """Copy the (mutated) fields back into the host record array."""
        for f in fields:
            if f in self.fields and f in array.dtype.names:
                t = self.fields[f].detach()
                if array.dtype[f] == np.bool_:
                    array[f] = t.to('cpu').numpy().astype(bool)
                else:
                    array[f] = t.to('cpu', torch.float64).numpy()

    def __getitem__(self, name):
        return self.fields[name]


def cell_capacity_for(extent_x, extent_y, cell_size, margin=0.5):
    """Grid capacity for a region extent_x x extent_y (m) with a margin for
    agents that leave it."""
    nx = int(np.ceil(extent_x * (1.0 + margin) / cell_size)) + 3
    ny = int(np.ceil(extent_y * (1.0 + margin) / cell_size)) + 3
    return max(nx * ny, 16)


def cell_slots_for(cell_size, r_min=0.2, slack=1.25):
    """Agent slots per cell: the most non-overlapping discs of radius r_min
    whose centres fit in a cell (hexagonal packing density over the cell
    grown by r_min on each side), with slack for compressed crowds.  Cells
    holding more take the exact slow path of the force kernel."""
    r = max(float(r_min), 0.05)
    k = slack * (cell_size + 2 * r) ** 2 * 0.9069 / (np.pi * r * r)
    return int(min(max(4 * int(np.ceil(k / 4.0)), 4), 1024))


def min_body_radius(array):
    """Smallest radius bounding how close two agents' centres get."""
    if 'position_ls' in array.dtype.names:
        return float(np.min(array['r_t'])) if len(array) else 0.2
    return float(np.min(array['radius'])) if len(array) else 0.2


class Workspace:
    """Caller-owned workspace of the C ABI (slot lists, block lists, state)."""

    def __init__(self, n, cell_capacity, model, cell_slots=16, device=None, stream=None,
                 max_scenarios=0, force_threads=0, tile=None, lanes=0, order_every=0):
        require_gpu()
        L = _lib.lib()
        self.desc = _lib.CDWorkspace()
        d = self.desc
        d.n = int(n)
        d.cell_capacity = int(cell_capacity)
        d.cell_slots = int(cell_slots)
        d.model = int(model)
        d.max_scenarios = int(max_scenarios)
        d.force_threads = int(force_threads)
        if tile is not None:
            d.tile_x, d.tile_y = int(tile[0]), int(tile[1])
        d.agent_lanes = int(lanes)
        d.order_every = int(order_every)
        self.layout = _lib.CDLayout()
        nbytes = L.cd_workspace_bytes(C.byref(d), C.byref(self.layout))
        if nbytes <= 0:
            raise InvalidValue('invalid workspace sizes (n=%d, cells=%d, slots=%d)'
                               % (n, cell_capacity, cell_slots))
        dev = torch.device(device) if device is not None else torch.device('cuda')
        self.buf = torch.empty(int(nbytes), dtype=torch.uint8, device=dev)
        d.ptr = self.buf.data_ptr()
        d.bytes = int(nbytes)
        _lib.check(L.cd_workspace_init(C.byref(d), stream_handle(stream)), 'workspace_init')

    n = property(lambda self: self.desc.n)
    cell_capacity = property(lambda self: self.desc.cell_capacity)
    cell_slots = property(lambda self: self.desc.cell_slots)
    nbytes = property(lambda self: self.desc.bytes)

    def ref(self):
        return C.byref(self.desc)

    def state(self, stream=None):
        """Synchronous read of the device scalars."""
        st = _lib.CDState()
        _lib.check(_lib.lib().cd_read_state(self.ref(), C.byref(st), stream_handle(stream)),
                   'read_state')
        return st

    def states(self, count, stream=None):
        """Synchronous read of the device scalars of scenarios [0, count)."""
        out = (_lib.CDState * int(count))()
        _lib.check(_lib.lib().cd_batch_read_states(self.ref(), out, int(count),
                                                   stream_handle(stream)), 'batch_read_states')
        return list(out)

    def clear_status(self, stream=None):
        _lib.check(_lib.lib().cd_clear_status(self.ref(), stream_handle(stream)), 'clear_status')

    def _i32(self, off, count):
        return self.buf[off:off + 4 * count].view(torch.int32)

    def block_list(self, stream=None):
        """(points_indices, cells_count, cells_offset, grid_shape, origin, cell_of)
        of the last cd_add_to_cells, as numpy int64 (synchronous)."""
        st = self.state(stream)
        nx, ny = int(st.shape[0]), int(st.shape[1])
        cells = nx * ny
        lay = self.layout
        n = self.n
        pidx = self._i32(lay.points_indices, n).cpu().numpy().astype(np.int64)
        cnt = self._i32(lay.cells_count, cells).cpu().numpy().astype(np.int64)
        off = self._i32(lay.cells_offset, cells).cpu().numpy().astype(np.int64)
        cell_of = self._i32(lay.cell_of, n).cpu().numpy().astype(np.int64)
        return pidx, cnt, off, (nx, ny), np.array(st.origin[:]), cell_of
