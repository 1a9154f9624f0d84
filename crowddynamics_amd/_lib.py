"""ctypes binding of libcdamd.so (the C ABI in include/cdamd.h).

The library is required: there is no CPU fallback.  Importing this module
does not touch the GPU; ``lib()`` raises DeviceError when the library is
missing or cannot be loaded.
"""
from __future__ import annotations

import ctypes as C
import os

from .exceptions import DeviceError, InvalidType, InvalidValue

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libcdamd.so')

CD_OK = 0
CD_ERR_ARG = -1
CD_ERR_TYPE = -2
CD_ERR_WORKSPACE = -3
CD_ERR_UNSUPPORTED = -4
CD_ERR_HIP = -5

#: bytes of the library's device scratch array of the allocation-free entry point
NOALLOC_SCRATCH_BYTES = 64 << 20

CD_STATUS_GRID_OVERFLOW = 1
CD_STATUS_NONFINITE = 2
CD_STATUS_SLOTS_FULL = 4
CD_STATUS_NAV_KEPT = 8

CD_USE_FLUCT = 1
CD_USE_NAV = 2
CD_USE_ADJUST = 4
CD_USE_AGENT_AGENT = 8
CD_USE_OBSTACLE = 16
CD_USE_RESET = 32
CD_USE_ORIENT_VELOCITY = 64

_vp = C.c_void_p


class CDAgents(C.Structure):
    _fields_ = [('n', C.c_int64), ('model', C.c_int32), ('real', C.c_int32),
                ('position', _vp), ('velocity', _vp), ('target_direction', _vp),
                ('force', _vp), ('force_prev', _vp), ('position_ls', _vp),
                ('position_rs', _vp), ('radius', _vp), ('r_t', _vp), ('r_s', _vp),
                ('r_ts', _vp), ('mass', _vp), ('inertia_rot', _vp),
                ('target_velocity', _vp), ('target_angular_velocity', _vp),
                ('tau_adj', _vp), ('k_soc', _vp), ('tau_0', _vp), ('mu', _vp),
                ('kappa', _vp), ('damping', _vp), ('std_rand_force', _vp),
                ('orientation', _vp), ('angular_velocity', _vp),
                ('target_orientation', _vp), ('torque', _vp), ('torque_prev', _vp),
                ('tau_rot', _vp), ('std_rand_torque', _vp), ('target', _vp)]


class CDWorkspace(C.Structure):
    _fields_ = [('ptr', _vp), ('bytes', C.c_int64), ('n', C.c_int64),
                ('cell_capacity', C.c_int64), ('cell_slots', C.c_int64), ('model', C.c_int32),
                ('max_scenarios', C.c_int32), ('n_scenarios', C.c_int32),
                ('blocks_f', C.c_int32), ('blocks_e', C.c_int32),
                ('force_threads', C.c_int32), ('tile_x', C.c_int32), ('tile_y', C.c_int32),
                ('blocks_t', C.c_int32), ('agent_lanes', C.c_int32),
                ('order_every', C.c_int32)]


class CDScenario(C.Structure):
    _fields_ = [('agent_begin', C.c_int64), ('agent_end', C.c_int64),
                ('cell_begin', C.c_int64), ('cell_capacity', C.c_int64),
                ('seg_begin', C.c_int64), ('seg_end', C.c_int64), ('seed', C.c_uint64),
                ('dt_min', C.c_double), ('dt_max', C.c_double)]


class CDAdjoint(C.Structure):
    _fields_ = [('lam_x', _vp), ('lam_v', _vp), ('lam_g', _vp), ('lam_f', _vp),
                ('lam_vt', _vp), ('grad', _vp), ('x_next', _vp), ('v_next', _vp)]


class CDTangent(C.Structure):
    _fields_ = [('dx', _vp), ('dv', _vp), ('dg', _vp), ('df', _vp), ('dp', _vp),
                ('x_next', _vp), ('v_next', _vp)]


class CDState(C.Structure):
    _fields_ = [('dt', C.c_double), ('time_tot', C.c_double),
                ('origin', C.c_double * 2), ('shape', C.c_int64 * 2),
                ('status', C.c_int64), ('step', C.c_int64), ('v_max', C.c_double),
                ('v0_max', C.c_double)]


class CDLayout(C.Structure):
    _fields_ = [('points_indices', C.c_int64), ('cells_count', C.c_int64),
                ('cells_offset', C.c_int64), ('cell_of', C.c_int64),
                ('total_bytes', C.c_int64)]


class CDStepParams(C.Structure):
    _fields_ = [('cell_size', C.c_double), ('dt_min', C.c_double), ('dt_max', C.c_double),
                ('flags', C.c_int32), ('nav_mode', C.c_int32), ('seed', C.c_uint64),
                ('segments', _vp), ('n_segments', C.c_int64), ('nav_maps', _vp),
                ('nav_targets', C.c_int64), ('nav_rows', C.c_int64), ('nav_cols', C.c_int64),
                ('nav_minx', C.c_double), ('nav_miny', C.c_double), ('nav_step', C.c_double),
                ('wall_cells', _vp), ('wall_nx', C.c_int64), ('wall_ny', C.c_int64),
                ('wall_minx', C.c_double), ('wall_miny', C.c_double), ('wall_cell', C.c_double)]


#: every symbol include/cdamd.h declares, with (restype, argtypes)
_P = C.POINTER
_W = _P(CDWorkspace)
SIGNATURES = {
    'cd_workspace_bytes': (C.c_int64, [_W, _P(CDLayout)]),
    'cd_workspace_init': (C.c_int, [_W, _vp]),
    'cd_read_state': (C.c_int, [_W, _P(CDState), _vp]),
    'cd_clear_status': (C.c_int, [_W, _vp]),
    'cd_add_to_cells': (C.c_int, [_P(CDAgents), C.c_double, _W, _vp]),
    'cd_agent_agent_block_list': (C.c_int, [_P(CDAgents), C.c_double, _W, _vp]),
    'cd_agent_agent_block_list_noalloc': (C.c_int, [_P(CDAgents), C.c_double, C.c_int64,
                                                    C.c_int64, _vp, _vp]),
    'cd_agent_obstacle': (C.c_int, [_P(CDAgents), _vp, C.c_int64, _vp]),
    'cd_force_adjust_agents': (C.c_int, [_P(CDAgents), _vp]),
    'cd_torque_adjust_agents': (C.c_int, [_P(CDAgents), _vp]),
    'cd_orient_towards_target_direction': (C.c_int, [_P(CDAgents), _vp]),
    'cd_navigation': (C.c_int, [_P(CDAgents), _vp, C.c_int64, C.c_int64, C.c_int64,
                                C.c_double, C.c_double, C.c_double, C.c_int32, _vp]),
    'cd_force_fluctuation': (C.c_int, [_P(CDAgents), C.c_uint64, C.c_uint64, _vp]),
    'cd_velocity_verlet_integrator': (C.c_int, [_P(CDAgents), C.c_double, C.c_double, _W, _vp]),
    'cd_reset': (C.c_int, [_P(CDAgents), _vp]),
    'cd_wall_grid_build': (C.c_int64, [_vp, C.c_int64, C.c_double, C.c_double, _vp,
                                       C.c_int64, _P(CDStepParams)]),
    'cd_step_prepare': (C.c_int, [_P(CDAgents), _P(CDStepParams), _W, C.c_uint64, _vp]),
    'cd_step': (C.c_int, [_P(CDAgents), _P(CDStepParams), _W, _vp]),
    'cd_step_tape': (C.c_int, [_P(CDAgents), _P(CDStepParams), _W, _vp, _vp, _vp]),
    'cd_batch_prepare': (C.c_int, [_P(CDAgents), _P(CDStepParams), _P(CDScenario), C.c_int32,
                                   _W, C.c_uint64, _vp]),
    'cd_batch_step': (C.c_int, [_P(CDAgents), _P(CDStepParams), _W, _vp]),
    'cd_batch_read_states': (C.c_int, [_W, _P(CDState), C.c_int32, _vp]),
    'cd_adjoint_step': (C.c_int, [_P(CDAgents), _P(CDStepParams), _vp, _P(CDAdjoint), _W,
                                  _vp]),
    'cd_tangent_step': (C.c_int, [_P(CDAgents), _P(CDStepParams), _P(CDTangent), _W, _vp]),
    'cd_rollout_record': (C.c_int, [_P(CDAgents), _P(CDStepParams), _W, C.c_int64, C.c_int64,
                                    _vp, _vp, _vp, _vp, _vp]),
    'cd_adjoint_segment': (C.c_int, [_P(CDAgents), _P(CDStepParams), _W, C.c_int64, C.c_int64,
                                     _vp, _vp, _vp, _vp, _vp, _vp, _P(CDAdjoint), _vp]),
    'cd_step_timed': (C.c_int, [_P(CDAgents), _P(CDStepParams), _W, C.c_int64,
                                _P(C.c_double), _vp]),
    'cd_kernel_timed': (C.c_int, [_P(CDAgents), _P(CDStepParams), _W, C.c_int32, C.c_int64,
                                  _P(C.c_double), _vp]),
    'cd_field_work_bytes': (C.c_int64, [C.c_int64, C.c_int64]),
    'cd_distance_map': (C.c_int, [_vp, _vp, C.c_int64, C.c_int64, C.c_double, _vp, _vp,
                                  C.c_int64, _vp]),
    'cd_direction_map': (C.c_int, [_vp, C.c_int64, C.c_int64, _vp, _vp, _vp]),
    'cd_obstacle_handling': (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_int64, C.c_double,
                                       C.c_double, _vp]),
    'cd_fill_missing': (C.c_int, [_vp, _vp, _vp, C.c_int64, C.c_int64, C.c_int32, _vp, _vp,
                                  _vp]),
    'cd_inside_domain': (C.c_int, [_vp, C.c_int64, _vp, C.c_int64, _vp, _vp, _vp]),
    'cd_target_reached': (C.c_int, [_vp, C.c_int64, _vp, C.c_int64, _vp, _vp, _vp]),
    'cd_nearest_neighbors': (C.c_int, [_P(CDAgents), C.c_double, C.c_int32, _vp, C.c_int64,
                                       _vp, _W, _vp]),
    'cd_herding_interaction': (C.c_int, [_vp, _vp, C.c_int64, _vp, _vp, C.c_int32, C.c_double,
                                         C.c_double, _vp, _vp, _vp]),
    'cd_graph_begin': (C.c_int, [_vp]),
    'cd_graph_end': (C.c_int, [_vp, _P(_vp)]),
    'cd_graph_launch': (C.c_int, [_vp, _vp]),
    'cd_graph_destroy': (C.c_int, [_vp]),
    'cd_hypot': (C.c_int, [_vp, _vp, C.c_int64, _vp]),
    'cd_philox4x32': (C.c_int, [_vp, _vp, _vp]),
    'cd_version': (C.c_char_p, []),
}

_lib = None


def lib():
    """The loaded library; raises DeviceError if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise DeviceError('libcdamd.so is not built (%s); run '
                              '`python -m crowddynamics_amd.build`' % LIB_PATH)
        try:
            L = C.CDLL(LIB_PATH)
        except OSError as e:
            raise DeviceError('cannot load %s: %s' % (LIB_PATH, e)) from e
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, what=''):
    """Map a C status code to the reference's exception types."""
    if rc == CD_OK:
        return
    if rc == CD_ERR_TYPE:
        raise InvalidType('%s: agents must be circular or three-circle records' % what)
    if rc == CD_ERR_ARG:
        raise InvalidValue('%s: invalid argument' % what)
    if rc == CD_ERR_WORKSPACE:
        raise InvalidValue('%s: workspace too small' % what)
    if rc == CD_ERR_UNSUPPORTED:
        raise InvalidValue('%s: unsupported model/precision combination' % what)
    raise DeviceError('%s: HIP error (%d)' % (what, rc))
