"""The drop-in LogicNode layer on the GPU: Hallway runs unchanged and matches
the oracle; the fused path equals the node-by-node path bit for bit; the
reference's validation runs pass through the drop-in API."""
import math

import numpy as np
import pytest

from crowddynamics_amd import examples as E
from crowddynamics_amd import logic as L
from crowddynamics_amd.agents import empty_agents, inertia_rot, shoulders_host
from crowddynamics_amd.simulation import Agents, Field, MultiAgentSimulation
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _oracle_for(sim):
    f = sim.field
    return O.Scenario(cell_size=3.6, dt_min=0.01, dt_max=0.01, seed=0, segments=f.obstacles,
                      nav_maps=f.navigation, nav_mode=f.nav_mode, nav_origin=f.nav_origin,
                      nav_step=f.nav_step)


@pytest.mark.parametrize('model', ['circular', 'three_circle'])
def test_hallway_nodes_vs_oracle_and_fused(model):
    say('hallway drop-in', model)
    sims = {}
    for fused in (False, 'auto'):
        sim = E.Hallway(size=60, agent_type=model, fused=fused)
        start = sim.agents.array.copy()
        sim.run(5)
        sims[fused] = sim
    ref = start.copy()
    O.run(ref, _oracle_for(sims[False]), 5)
    a = sims[False].agents.array
    b = sims['auto'].agents.array
    assert sims['auto']._plan['fused_flags'] is not None
    assert sims[False]._plan['fused_flags'] is None
    # the fused kernel splits the wall sum over four lanes, the node path adds
    # walls one by one: same terms, different order
    for f in ('position', 'velocity', 'force_prev', 'target_direction'):
        assert np.allclose(a[f], b[f], rtol=1e-12, atol=1e-12), f
        assert np.allclose(a[f], ref[f], rtol=1e-12, atol=1e-12), f
    assert np.all(a['force'] == 0.0)
    if model == 'three_circle':
        for f in ('orientation', 'angular_velocity', 'position_ls'):
            assert np.allclose(a[f], b[f], rtol=1e-12, atol=1e-12), f
            assert np.allclose(a[f], ref[f], rtol=1e-12, atol=1e-12), f
    for s in sims.values():
        assert s.data['iterations'] == 5
        assert abs(s.data['time_tot'] - 0.05) < 1e-15
        assert s.data['dt'] == 0.01


def test_host_edits_between_steps_are_seen():
    sim = E.Hallway(size=20)
    sim.run(2)
    arr = sim.agents.array
    # an isolated spot 10 m outside the hallway wall: one step moves the agent
    # by the adjusting force only (a / 2 * dt^2 from rest)
    arr['position'][0] = (20.0, -10.0)
    arr['velocity'][0] = (0.0, 0.0)
    sim.update()
    got = sim.agents.array['position'][0]
    assert np.hypot(*(got - (20.0, -10.0))) < 1e-3


def _validation_agents(model, placements):
    a = empty_agents(len(placements), model)
    for k, (pos, phi) in enumerate(placements):
        a['radius'][k] = 0.255
        a['mass'][k] = 73.5
        a['target_velocity'][k] = 1.0
        a['position'][k] = pos
        a['velocity'][k] = (math.cos(phi), math.sin(phi))
        a['target_direction'][k] = (math.cos(phi), math.sin(phi))
        if model == 'three_circle':
            a['r_t'][k], a['r_s'][k], a['r_ts'][k] = 0.5882 * 0.255, 0.3725 * 0.255, \
                0.6275 * 0.255
            a['orientation'][k] = phi
            a['target_orientation'][k] = phi
    a['inertia_rot'] = inertia_rot(a['mass'], a['radius'])
    if model == 'three_circle':
        shoulders_host(a)
    return a


@pytest.mark.parametrize('model', ['circular', 'three_circle'])
def test_reference_validation_through_dropin(model):
    """crowddynamics/examples/tests/test_validation.py:13-38 with the drop-in
    classes and the reference's trees (examples/validation.py:41-44, 88-96)."""
    say('validation movement', model)
    sim = MultiAgentSimulation(agents=Agents(model, _validation_agents(model, [((0, 0), 0.0)])),
                               cell_capacity=4096)
    sim.logic = L.Reset(sim) << L.Integrator(sim) << L.Adjusting(sim) << L.Orientation(sim)
    start = sim.agents.array.copy()
    sim.exit_condition = lambda s: s.data['iterations'] == 1000
    sim.run()
    dist = np.hypot(*(sim.agents.array[0]['position'] - start[0]['position']))
    assert dist >= 10.0 or np.isclose(dist, 10.0)
    say('validation interaction', model)
    sim = MultiAgentSimulation(agents=Agents(model, _validation_agents(
        model, [((0, 0), 0.0), ((5, 0), math.pi)])), cell_capacity=4096)
    sim.logic = L.Reset(sim) << (L.Integrator(sim) << (
        L.Fluctuation(sim), L.Adjusting(sim) << L.Orientation(sim),
        L.AgentAgentInteractions(sim)))
    start = sim.agents.array.copy()
    sim.exit_condition = lambda s: s.data['iterations'] == 1000
    sim.run()
    dist = np.hypot(*(sim.agents.array[0]['position'] - start[0]['position']))
    assert dist >= 8.0 or np.isclose(dist, 8.0)
