"""The C ABI library loads on a host without a GPU and exports every symbol
include/cdamd.h declares (no compute calls)."""
import ctypes
import os
import re

from crowddynamics_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, 'include', 'cdamd.h')).read()
    src = re.sub(r'/\*.*?\*/', '', src, flags=re.S)
    return sorted(set(re.findall(r'\b(cd_[a-z0-9_]+)\s*\(', src)))


def test_header_symbols_exported():
    L = _lib.lib()
    names = _declared()
    assert len(names) >= 20
    for name in names:
        assert hasattr(L, name), name


def test_bindings_cover_header():
    assert set(_declared()) == set(_lib.SIGNATURES)


def test_workspace_sizing_is_host_only():
    L = _lib.lib()
    lay = _lib.CDLayout()
    w = _lib.CDWorkspace()
    w.n, w.cell_capacity, w.cell_slots, w.model = 20000, 60000, 12, 0
    n = L.cd_workspace_bytes(ctypes.byref(w), ctypes.byref(lay))
    assert n == lay.total_bytes > 0
    assert lay.points_indices % 256 == 0 and lay.cells_offset % 256 == 0
    w.model = 1
    assert L.cd_workspace_bytes(ctypes.byref(w), None) > n      # three-circle slot data
    w.n = -1
    assert L.cd_workspace_bytes(ctypes.byref(w), None) < 0
    w.n, w.cell_slots = 10, 0
    assert L.cd_workspace_bytes(ctypes.byref(w), None) < 0
    w.cell_slots, w.order_every = 12, 7   # the processing order adds 4 (n + cells) bytes
    m = L.cd_workspace_bytes(ctypes.byref(w), None)
    assert m > 0
    w.order_every = -1
    assert L.cd_workspace_bytes(ctypes.byref(w), None) < 0
    assert L.cd_version().startswith(b'cdamd')


def test_cell_slot_sizing():
    from crowddynamics_amd.device import cell_slots_for
    assert cell_slots_for(0.6, 0.22) >= 8
    assert cell_slots_for(3.6, 0.22) >= 100
    assert cell_slots_for(3.6, 0.22) % 4 == 0


def test_argument_validation_without_device():
    """Null agent pointers are rejected before any HIP call."""
    L = _lib.lib()
    assert L.cd_force_adjust_agents(None, None) == _lib.CD_ERR_ARG
    a = _lib.CDAgents()
    a.n, a.model, a.real = 10, 7, 0
    assert L.cd_force_adjust_agents(ctypes.byref(a), None) == _lib.CD_ERR_TYPE
    a.model, a.real = 0, 5
    assert L.cd_force_adjust_agents(ctypes.byref(a), None) == _lib.CD_ERR_ARG
    for model, real in ((0, 0), (0, 1), (1, 0), (1, 1)):  # every model in fp64 and fp32
        a.model, a.real = model, real
        assert L.cd_force_adjust_agents(ctypes.byref(a), None) == _lib.CD_ERR_ARG
    # the adjoint covers circular agents in fp64 only
    a.model, a.real, a.n = 1, 0, 0
    assert L.cd_adjoint_step(ctypes.byref(a), None, None, None, None, None) == _lib.CD_ERR_ARG


def _fake_circular(n):
    """Descriptor with non-null placeholder pointers (never dereferenced:
    every call below fails validation before touching the device)."""
    a = _lib.CDAgents()
    a.n, a.model, a.real = n, 0, 0
    for f in ('position', 'velocity', 'target_direction', 'force', 'force_prev', 'radius',
              'mass', 'target_velocity', 'tau_adj', 'k_soc', 'tau_0', 'mu', 'kappa',
              'damping', 'std_rand_force', 'target'):
        setattr(a, f, 0x1000)
    return a


def test_batch_table_validation_without_device():
    """cd_batch_prepare rejects inconsistent scenario tables on the host."""
    L = _lib.lib()
    a = _fake_circular(100)
    w = _lib.CDWorkspace()
    w.n, w.cell_capacity, w.cell_slots, w.model, w.max_scenarios = 100, 400, 8, 0, 3
    w.bytes = L.cd_workspace_bytes(ctypes.byref(w), None)
    assert w.bytes > 0
    w.ptr = 0x2000
    p = _lib.CDStepParams()
    p.cell_size, p.flags = 3.6, _lib.CD_USE_AGENT_AGENT

    def table(rows):
        t = (_lib.CDScenario * len(rows))()
        for k, (a0, a1, c0, cc) in enumerate(rows):
            t[k].agent_begin, t[k].agent_end = a0, a1
            t[k].cell_begin, t[k].cell_capacity = c0, cc
            t[k].dt_min, t[k].dt_max = 0.001, 0.01
        return t

    def prep(rows, count=None):
        t = table(rows)
        return L.cd_batch_prepare(ctypes.byref(a), ctypes.byref(p), t,
                                  len(rows) if count is None else count,
                                  ctypes.byref(w), 0, None)

    bad = [
        [(0, 50, 0, 100), (40, 100, 100, 100)],        # overlapping agent ranges
        [(0, 50, 0, 100), (50, 101, 100, 100)],        # beyond the agent arrays
        [(0, 50, 0, 100), (50, 100, 350, 100)],        # beyond the cell capacity
        [(0, 50, 0, 0), (50, 100, 100, 100)],          # empty cell range
        [(0, 30, 0, 100), (30, 60, 100, 100), (60, 90, 200, 100), (90, 100, 300, 50)],
    ]
    for rows in bad:
        assert prep(rows) == _lib.CD_ERR_ARG, rows
    assert prep([(0, 100, 0, 400)], count=0) == _lib.CD_ERR_ARG
    w.force_threads = 96
    assert L.cd_workspace_bytes(ctypes.byref(w), None) == -1
