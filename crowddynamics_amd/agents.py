"""Agent record types and the device-side structure-of-arrays mirror.

The host-visible API keeps the reference's packed structured numpy records
(``agent_type_circular`` / ``agent_type_three_circle``).  Their field order is
the one ``class_to_struct_dtype`` derives from the trait classes
(crowddynamics/traits.py:158-219) through the MRO of ``Circular`` and
``ThreeCircle`` (crowddynamics/simulation/agents.py:348,373,447-448): packed,
align=False, 228 and 316 bytes.

Kernels never touch that layout (most f8 fields sit at offsets that are not
multiples of 8).  :class:`DeviceAgents` holds one contiguous device array per
field instead, vectors as interleaved ``[n, 2]``; ``upload``/``download``
convert between the two.
"""
from __future__ import annotations

import numpy as np

from .exceptions import InvalidType

_STATES = [
    ('active', np.bool_),
    ('target_reached', np.bool_),
    ('target', np.int64),
    ('is_leader', np.bool_),
    ('is_follower', np.bool_),
    ('index_leader', np.int64),
    ('familiar_exit', np.int64),
]
_BODY = [
    ('radius', np.float64),
    ('r_t', np.float64),
    ('r_s', np.float64),
    ('r_ts', np.float64),
    ('mass', np.float64),
    ('inertia_rot', np.float64),
    ('target_velocity', np.float64),
    ('target_angular_velocity', np.float64),
]
_TRANSLATIONAL = [
    ('position', np.float64, (2,)),
    ('velocity', np.float64, (2,)),
    ('target_direction', np.float64, (2,)),
    ('force', np.float64, (2,)),
    ('force_prev', np.float64, (2,)),
    ('tau_adj', np.float64),
    ('k_soc', np.float64),
    ('tau_0', np.float64),
    ('mu', np.float64),
    ('kappa', np.float64),
    ('damping', np.float64),
    ('std_rand_force', np.float64),
]
_ROTATIONAL = [
    ('orientation', np.float64),
    ('angular_velocity', np.float64),
    ('target_orientation', np.float64),
    ('torque', np.float64),
    ('torque_prev', np.float64),
    ('tau_rot', np.float64),
    ('std_rand_torque', np.float64),
]
_SHOULDERS = [
    ('position_ls', np.float64, (2,)),
    ('position_rs', np.float64, (2,)),
]

#: Circular record, 228 bytes (crowddynamics/simulation/agents.py:447)
agent_type_circular = np.dtype(_STATES + _BODY + _TRANSLATIONAL)
#: ThreeCircle record, 316 bytes (crowddynamics/simulation/agents.py:448)
agent_type_three_circle = np.dtype(
    _SHOULDERS + _STATES + _BODY + _TRANSLATIONAL + _ROTATIONAL)

AgentModelToType = {
    'circular': agent_type_circular,
    'three_circle': agent_type_three_circle,
}

#: Defaults of the motion traits (crowddynamics/simulation/agents.py:223-292)
DEFAULTS = dict(
    tau_adj=0.5, k_soc=1.5, tau_0=3.0, mu=1.2e5, kappa=4e4, damping=500.0,
    std_rand_force=0.1, tau_rot=0.2, std_rand_torque=0.1,
)

NO_TARGET = -1
NO_LEADER = -1


def is_model(agents, model):
    """True when ``agents`` has the record type named ``model``.

    Same test as crowddynamics/simulation/agents.py:460-470 (dtype hash).
    """
    return hash(agents.dtype) == hash(AgentModelToType[model])


def model_of(agents):
    """0 for circular, 1 for three-circle; InvalidType otherwise."""
    if is_model(agents, 'circular'):
        return 0
    if is_model(agents, 'three_circle'):
        return 1
    raise InvalidType('agents must be agent_type_circular or '
                      'agent_type_three_circle records')


def inertia_rot(mass, radius):
    """Rotational inertia 4*pi*(m/80)*(r/0.27)**2
    (crowddynamics/simulation/agents.py:185-192)."""
    return 4.0 * np.pi * (np.asarray(mass) / 80.0) * (np.asarray(radius) / 0.27) ** 2


def empty_agents(n, model='circular'):
    """Zero-initialised records with the reference trait defaults filled in."""
    arr = np.zeros(n, dtype=AgentModelToType[model])
    arr['active'] = True
    arr['target'] = NO_TARGET
    arr['index_leader'] = NO_LEADER
    arr['familiar_exit'] = NO_TARGET
    for k, v in DEFAULTS.items():
        if k in arr.dtype.names:
            arr[k] = v
    return arr


def shoulders_host(arr):
    """position_ls / position_rs from position, orientation and r_ts
    (crowddynamics/simulation/agents.py:473-486)."""
    phi = arr['orientation']
    off = np.stack((np.sin(phi) * arr['r_ts'], (-np.cos(phi)) * arr['r_ts']), axis=1)
    arr['position_ls'] = arr['position'] - off
    arr['position_rs'] = arr['position'] + off


#: Fields mirrored on the device, in the order of the C-ABI struct cd_agents_t.
VECTOR_FIELDS = ('position', 'velocity', 'target_direction', 'force', 'force_prev',
                 'position_ls', 'position_rs')
SCALAR_FIELDS = ('radius', 'r_t', 'r_s', 'r_ts', 'mass', 'inertia_rot',
                 'target_velocity', 'target_angular_velocity', 'tau_adj', 'k_soc',
                 'tau_0', 'mu', 'kappa', 'damping', 'std_rand_force',
                 'orientation', 'angular_velocity', 'target_orientation', 'torque',
                 'torque_prev', 'tau_rot', 'std_rand_torque')
INT_FIELDS = ('target',)
#: Boolean fields mirrored on the device as uint8 (state nodes, herding).
BOOL_FIELDS = ('active', 'target_reached', 'is_leader', 'is_follower')
#: Fields the step writes (downloaded by ``DeviceAgents.download``).
MUTABLE_FIELDS = ('position', 'velocity', 'target_direction', 'force', 'force_prev',
                  'position_ls', 'position_rs', 'orientation', 'angular_velocity',
                  'target_orientation', 'torque', 'torque_prev')
#: Fields held in the compute precision in fp32 mode (position stays fp64).
REAL_FIELDS = ('velocity', 'target_direction', 'force', 'force_prev')
