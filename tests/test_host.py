"""Host-side logic: record layouts, type dispatch, scenario generators."""
import numpy as np
import pytest
import torch

from crowddynamics_amd import agents as A
from crowddynamics_amd import examples as E
from crowddynamics_amd.exceptions import DeviceError, InvalidType


def test_record_layouts_match_reference_derivation():
    c = A.agent_type_circular
    t = A.agent_type_three_circle
    assert c.itemsize == 228 and t.itemsize == 316
    offs = {k: v[1] for k, v in c.fields.items()}
    assert offs['target'] == 2 and offs['radius'] == 28 and offs['position'] == 92
    assert offs['force'] == 140 and offs['std_rand_force'] == 220
    toffs = {k: v[1] for k, v in t.fields.items()}
    assert toffs['position_ls'] == 0 and toffs['position'] == 124
    assert toffs['orientation'] == 260 and toffs['std_rand_torque'] == 308


def test_is_model_and_invalid_type():
    assert A.is_model(np.zeros(1, A.agent_type_circular), 'circular')
    assert A.is_model(np.zeros(1, A.agent_type_three_circle), 'three_circle')
    assert not A.is_model(np.zeros(1, A.agent_type_circular), 'three_circle')
    with pytest.raises(InvalidType):
        A.model_of(np.zeros(3, dtype=[('position', 'f8', 2)]))


def test_inertia_formula():
    assert np.isclose(A.inertia_rot(80.0, 0.27), 4 * np.pi)


@pytest.mark.parametrize('gen', [E.hallway, lambda: E.room(n=3000, width=80, height=40),
                                 lambda: E.corridor(n=2000, width=60, height=30,
                                                    n_obstacles=20)])
def test_generators_place_without_overlap(gen):
    from scipy.spatial import cKDTree
    s = gen()
    p = s.agents['position']
    r = s.agents['radius'] if 'position_ls' not in s.agents.dtype.names else \
        np.full(len(p), 0.25)
    for i, j in cKDTree(p).query_pairs(0.6):
        assert np.hypot(*(p[i] - p[j])) >= r[i] + r[j]
    x0, y0, x1, y1 = s.bounds
    assert np.all(p[:, 0] - r >= x0 - 1e-12) and np.all(p[:, 0] + r <= x1 + 1e-12)


def test_ensemble_generator():
    sc = E.ensemble(n_scenarios=6, n_min=100, n_max=400)
    assert len(sc) == 6
    for s in sc:
        assert 100 <= s.n <= 400
        from scipy.spatial import cKDTree
        p, r = s.agents['position'], s.agents['radius']
        for i, j in cKDTree(p).query_pairs(0.6):
            assert np.hypot(*(p[i] - p[j])) >= r[i] + r[j]


@pytest.mark.skipif(torch.cuda.is_available(), reason='checks the no-GPU error path')
def test_device_path_fails_loudly_without_gpu():
    from crowddynamics_amd.device import DeviceAgents
    with pytest.raises(DeviceError):
        DeviceAgents(A.empty_agents(4))
