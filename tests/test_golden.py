"""Golden fixtures (tests/golden/*.npz, written by make_golden.py): the
oracle reproduces them on this host, and the HIP path (gpu marker) matches
them through the C ABI."""
import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.helpers import assert_vec_close

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def _load(name):
    return np.load(os.path.join(HERE, name), allow_pickle=False)


def _cfg(g):
    return O.Scenario(cell_size=float(g['cell_size']), dt_min=0.001, dt_max=0.01, seed=1,
                      segments=g['segments'])


def test_oracle_reproduces_hallway_golden():
    g = _load('hallway200.npz')
    a = g['initial'].copy()
    O.run(a, _cfg(g), 10)
    for f in ('position', 'velocity', 'force_prev'):
        assert np.allclose(a[f], g['step10'][f], rtol=1e-13, atol=1e-13)


def test_oracle_reproduces_threecircle_golden():
    g = _load('threecircle400.npz')
    a = g['initial'].copy()
    a['force'] = 0.0
    a['torque'] = 0.0
    O.agent_agent_block_list(a, float(g['cell_size']))
    assert np.allclose(a['force'], g['pair_force'], rtol=1e-13, atol=1e-12)
    assert np.allclose(a['torque'], g['pair_torque'], rtol=1e-13, atol=1e-12)


@pytest.mark.gpu
def test_gpu_matches_golden():
    from crowddynamics_amd import core
    from crowddynamics_amd.device import DeviceAgents, Workspace
    from crowddynamics_amd.stepper import ALL_NODES, Stepper
    from crowddynamics_amd import _lib
    g = _load('hallway200.npz')
    print('golden hallway pair forces', flush=True)
    a = g['initial'].copy()
    a['force'] = 0.0
    dev = DeviceAgents(a)
    ws = Workspace(len(a), 4096, 0, cell_slots=128)
    core.agent_agent_block_list(dev, float(g['cell_size']), ws)
    out = a.copy()
    dev.download(out)
    assert_vec_close(out['force'], g['pair_force'], g['pair_scale'], 1e-12, 'golden force')
    print('golden hallway 10 steps', flush=True)
    st = Stepper(g['initial'].copy(), segments=g['segments'], cell_size=float(g['cell_size']),
                 dt_min=0.001, dt_max=0.01, seed=1, flags=ALL_NODES & ~_lib.CD_USE_NAV,
                 bounds=(0, 0, 20, 5))
    st.step(10)
    out = st.download(g['initial'].copy())
    assert np.abs(out['position'] - g['step10']['position']).max() < 1e-11
    t = _load('threecircle400.npz')
    print('golden three-circle pair forces', flush=True)
    a = t['initial'].copy()
    a['force'] = 0.0
    a['torque'] = 0.0
    dev = DeviceAgents(a)
    ws = Workspace(len(a), 4096, 1, cell_slots=32)
    core.agent_agent_block_list(dev, float(t['cell_size']), ws)
    out = a.copy()
    dev.download(out)
    assert_vec_close(out['force'], t['pair_force'], t['pair_scale'], 1e-12, 'golden tc force')
    assert_vec_close(out['torque'], t['pair_torque'], t['torque_scale'], 1e-12, 'golden torque')
