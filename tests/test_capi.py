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
    a.model, a.real = 1, 1
    assert L.cd_force_adjust_agents(ctypes.byref(a), None) == _lib.CD_ERR_UNSUPPORTED
    a.model, a.real = 0, 0
    assert L.cd_force_adjust_agents(ctypes.byref(a), None) == _lib.CD_ERR_ARG
