"""The allocation-free AgentAgentInteractions entry point
(cd_agent_agent_block_list_noalloc) against the workspace entry point
(cd_agent_agent_block_list), through the C ABI.

Both run the same kernels, whose pair sums are exact, so the forces and
torques must agree bit for bit.  Two calls in a row share the library's
scratch array; the second has fewer agents (and another agent model), so it
also checks that the first call's contents do not leak into the second.
The workspace entry point is itself held to the oracle by the pair-force
parity tests; here the oracle is checked once more at the same tolerance.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from crowddynamics_amd import _lib, core
from crowddynamics_amd import examples as E
from crowddynamics_amd.device import (DeviceAgents, Workspace, cell_capacity_for,
                                      cell_slots_for, min_body_radius, stream_handle)
from crowddynamics_amd.exceptions import InvalidValue
from oracle import oracle as O
from tests.helpers import assert_vec_close

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _sizes(arr, cell, model):
    p = arr['position']
    span = p.max(0) - p.min(0)
    cap = cell_capacity_for(span[0], span[1], cell, margin=0.1)
    slots = cell_slots_for(cell, min_body_radius(arr))
    w = _lib.CDWorkspace()
    w.n, w.cell_capacity, w.cell_slots, w.model = len(arr), cap, slots, model
    need = _lib.lib().cd_workspace_bytes(C.byref(w), None)
    assert 0 < need <= _lib.NOALLOC_SCRATCH_BYTES, (need, cap, slots)
    return cap, slots


def _noalloc(arr, cell, precision='f64'):
    """forces (and torques) of both entry points, and both status words"""
    model = 1 if 'position_ls' in arr.dtype.names else 0
    arr = arr.copy()
    arr['force'] = 0.0
    if model:
        arr['torque'] = 0.0
    cap, slots = _sizes(arr, cell, model)
    say('noalloc n=%d cell=%g model=%d cells=%d slots=%d %s'
        % (len(arr), cell, model, cap, slots, precision))
    # the workspace entry point
    dev_ws = DeviceAgents(arr, precision=precision)
    ws = Workspace(len(arr), cap, model, cell_slots=slots)
    core.agent_agent_block_list(dev_ws, cell, ws)
    # the allocation-free entry point
    dev_na = DeviceAgents(arr, precision=precision)
    status = torch.full((1,), -1, dtype=torch.int64, device='cuda')
    rc = _lib.lib().cd_agent_agent_block_list_noalloc(
        dev_na.ref(), float(cell), int(cap), int(slots), C.c_void_p(status.data_ptr()),
        stream_handle())
    _lib.check(rc, 'agent_agent_block_list_noalloc')
    torch.cuda.synchronize()
    out_ws, out_na = arr.copy(), arr.copy()
    dev_ws.download(out_ws)
    dev_na.download(out_na)
    return out_ws, out_na, int(ws.state().status), int(status.item())


def _same(out_ws, out_na, fields):
    for f in fields:
        assert np.array_equal(out_ws[f], out_na[f]), f


def test_noalloc_matches_workspace_entry_point_twice():
    # first call: the 20k room's agents left of x = 100 m (about 10k), fp64
    room = E.room()
    first = room.agents[room.agents['position'][:, 0] < 100.0].copy()
    out_ws, out_na, st_ws, st_na = _noalloc(first, 0.6)
    assert st_ws == st_na == 0
    _same(out_ws, out_na, ('force',))
    ref = first.copy()
    ref['force'] = 0.0
    fs, _ = O.agent_agent_scale(ref, 0.6)
    assert_vec_close(out_na['force'], ref['force'], fs, 1e-12, 'force vs oracle')
    # second call, smaller: 2000 three-circle agents of a corridor
    cor = E.corridor(n=2000, width=60.0, height=30.0, n_obstacles=10)
    out_ws, out_na, st_ws, st_na = _noalloc(cor.agents, 0.6)
    assert st_ws == st_na
    assert st_na & ~4 == 0
    _same(out_ws, out_na, ('force', 'torque'))
    # third, smaller again and fp32
    hall = E.hallway()
    out_ws, out_na, st_ws, st_na = _noalloc(hall.agents, 3.6, precision='f32')
    assert st_ws == st_na == 0
    _same(out_ws, out_na, ('force',))


def test_noalloc_status_and_limits():
    hall = E.hallway()
    arr = hall.agents.copy()
    dev = DeviceAgents(arr)
    status = torch.full((1,), -1, dtype=torch.int64, device='cuda')
    L = _lib.lib()
    # a grid too small for the agents: pairs skipped and reported
    rc = L.cd_agent_agent_block_list_noalloc(dev.ref(), 0.6, 4, 8,
                                             C.c_void_p(status.data_ptr()), stream_handle())
    _lib.check(rc, 'noalloc')
    torch.cuda.synchronize()
    assert int(status.item()) & _lib.CD_STATUS_GRID_OVERFLOW
    # sizes beyond the scratch array are refused before any launch
    rc = L.cd_agent_agent_block_list_noalloc(dev.ref(), 0.6, 1 << 26, 16, None, stream_handle())
    assert rc == _lib.CD_ERR_WORKSPACE
    with pytest.raises(InvalidValue):
        _lib.check(rc, 'noalloc')
