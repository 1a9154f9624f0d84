"""Parity of the HIP path (through the C ABI) with the CPU oracle.

Bars: cell indices, block lists and pair sets bit-exact; fp64 forces and
torques within 1e-12 of the per-agent force scale (sum of the magnitudes of
the summed terms: the two sides sum the same terms in different orders);
per-agent operators that keep the reference's order bit-for-bit up to libm
ulps (1e-14 relative).
"""
import numpy as np
import pytest
import torch

from crowddynamics_amd import core
from crowddynamics_amd import examples as E
from crowddynamics_amd.agents import empty_agents, inertia_rot, shoulders_host
from crowddynamics_amd.device import (DeviceAgents, Workspace, cell_capacity_for,
                                      cell_slots_for, min_body_radius)
from crowddynamics_amd.stepper import Stepper
from oracle import oracle as O
from tests.helpers import assert_vec_close, max_rel

pytestmark = pytest.mark.gpu

F64_FORCE_RTOL = 1e-12
OP_RTOL = 1e-14


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope='module')
def hallway():
    return E.hallway()


@pytest.fixture(scope='module')
def room():
    return E.room()


@pytest.fixture(scope='module')
def corridor_small():
    return E.corridor(n=3000, width=80.0, height=40.0, n_obstacles=40)


def _ws_for(arr, cell, slots=None):
    p = arr['position']
    span = p.max(0) - p.min(0)
    model = 0 if 'position_ls' not in arr.dtype.names else 1
    if slots is None:
        slots = cell_slots_for(cell, min_body_radius(arr))
    return Workspace(len(arr), cell_capacity_for(span[0], span[1], cell), model,
                     cell_slots=slots)


def _perturb(arr, seed, scale=0.3):
    """Random velocities so the social force is active in every direction."""
    rng = np.random.default_rng(seed)
    arr = arr.copy()
    arr['velocity'] = rng.normal(0, scale, arr['velocity'].shape) + arr['velocity']
    return arr


# ---------------------------------------------------------------- binning --

def _check_binning(arr, cell):
    say('binning n=%d cell=%g' % (len(arr), cell))
    dev = DeviceAgents(arr)
    ws = _ws_for(arr, cell)
    core.add_to_cells(dev, cell, ws)
    got = ws.block_list()
    ref = O.add_to_cells(arr['position'], cell)
    names = ('points_indices', 'cells_count', 'cells_offset', 'shape', 'origin', 'cell_of')
    for name, g, r in zip(names, got, ref):
        if name in ('shape',):
            assert tuple(g) == tuple(r), name
        elif name == 'origin':
            assert np.array_equal(np.asarray(g), np.asarray(r)), name
        else:
            assert np.array_equal(np.asarray(g), np.asarray(r)), name
    assert ws.state().status == 0


def test_binning_hallway(hallway):
    _check_binning(hallway.agents, 3.6)
    _check_binning(hallway.agents, 0.6)


def test_binning_room(room):
    _check_binning(room.agents, 0.6)


def test_binning_edge_cases():
    rng = np.random.default_rng(5)
    # single agent
    a = empty_agents(1)
    a['position'] = (3.0, -2.0)
    _check_binning(a, 0.6)
    # points exactly on cell boundaries, negative coordinates, repeated points
    a = empty_agents(64)
    k = np.arange(64)
    a['position'][:, 0] = -5.0 + (k % 8) * 0.5
    a['position'][:, 1] = -1.0 + (k // 8) * 0.25
    a['position'][7] = a['position'][3]
    _check_binning(a, 0.5)
    # all agents in one cell
    a = empty_agents(40)
    a['position'] = rng.uniform(10.0, 10.3, (40, 2))
    _check_binning(a, 3.6)
    # a line of agents (ny = 1)
    a = empty_agents(100)
    a['position'][:, 0] = rng.uniform(0, 50, 100)
    a['position'][:, 1] = 7.0
    _check_binning(a, 0.6)


# ------------------------------------------------------------ pair forces --

def _check_pairs(arr, cell, precision='f64', rtol=F64_FORCE_RTOL, slots=None):
    say('pairs n=%d cell=%g %s model=%s' % (len(arr), cell, precision,
                                            'tc' if 'position_ls' in arr.dtype.names else 'c'))
    arr = arr.copy()
    arr['force'] = 0.0
    tc = 'position_ls' in arr.dtype.names
    if tc:
        arr['torque'] = 0.0
    ref = arr.copy()
    fs, ts = O.agent_agent_scale(ref, cell)
    dev = DeviceAgents(arr, precision=precision)
    ws = _ws_for(arr, cell, slots)
    core.agent_agent_block_list(dev, cell, ws)
    out = arr.copy()
    dev.download(out)
    status = ws.state().status
    assert status & ~4 == 0           # no grid overflow / non-finite
    assert (status & 4 != 0) == (slots is not None and slots < 4)
    assert_vec_close(out['force'], ref['force'], fs, rtol, 'force')
    if tc:
        assert_vec_close(out['torque'], ref['torque'], ts, rtol, 'torque')
    return max_rel(out['force'], ref['force'], fs)


def test_pairs_hallway(hallway):
    _check_pairs(_perturb(hallway.agents, 1), 3.6)
    _check_pairs(_perturb(hallway.agents, 2), 0.6)


def test_pairs_room(room):
    _check_pairs(_perturb(room.agents, 3), 0.6)


def test_pairs_slot_overflow_exact_path():
    """Cells holding more agents than slots take the exact slow path: same
    forces (and the status bit reports it)."""
    rng = np.random.default_rng(13)
    a = empty_agents(300)
    a['radius'] = rng.uniform(0.22, 0.28, 300)
    a['mass'] = 80.0
    a['position'] = rng.uniform(0, 6, (300, 2))
    a['velocity'] = rng.normal(0, 1, (300, 2))
    _check_pairs(a, 1.2, slots=2)


def test_pairs_large_cells_global_path():
    """Cells of 10 m holding ~170 agents each: tiles exceed the LDS capacity,
    3 x 3 regions around inner cells still do, and the centre cell's region
    does not (served from memory).  Same forces on every path."""
    rng = np.random.default_rng(17)
    n = 1500
    a = empty_agents(n)
    a['radius'] = rng.uniform(0.22, 0.28, n)
    a['mass'] = rng.normal(80, 8, n)
    a['position'] = rng.uniform(0, 29.9, (n, 2))
    a['velocity'] = rng.normal(0, 1, (n, 2))
    _check_pairs(a, 10.0)


def test_pairs_contact_heavy():
    """Overlapping agents (contact + friction + damping terms active)."""
    rng = np.random.default_rng(11)
    n = 400
    a = empty_agents(n)
    a['radius'] = rng.uniform(0.22, 0.28, n)
    a['mass'] = rng.normal(80, 8, n)
    a['position'] = rng.uniform(0, 8, (n, 2))
    a['velocity'] = rng.normal(0, 1, (n, 2))
    a['target_velocity'] = 1.2
    _check_pairs(a, 0.6)
    _check_pairs(a, 3.6)


def test_pairs_three_circle(corridor_small):
    _check_pairs(_perturb(corridor_small.agents, 4), 0.6)


def test_pairs_three_circle_contact():
    rng = np.random.default_rng(12)
    n = 300
    a = empty_agents(n, 'three_circle')
    a['radius'] = 0.25
    a['r_t'] = 0.15
    a['r_s'] = 0.1
    a['r_ts'] = 0.15
    a['mass'] = rng.normal(80, 8, n)
    a['position'] = rng.uniform(0, 7, (n, 2))
    a['velocity'] = rng.normal(0, 1, (n, 2))
    a['orientation'] = rng.uniform(-np.pi, np.pi, n)
    shoulders_host(a)
    _check_pairs(a, 0.6)


def test_pairs_fp32_error_bound(room):
    """fp32 mode: forces within 2e-4 of the fp64 oracle's force scale."""
    r = _check_pairs(_perturb(room.agents, 3), 0.6, precision='f32', rtol=2e-4)
    say('fp32 max scaled error %.3e' % r)


# ------------------------------------------------------ per-agent operators --

def _roundtrip(arr, fn_dev, fn_ref, fields, rtol=OP_RTOL, precision='f64'):
    ref = arr.copy()
    fn_ref(ref)
    dev = DeviceAgents(arr, precision=precision)
    fn_dev(dev)
    out = arr.copy()
    dev.download(out)
    for f in fields:
        g, r = out[f], ref[f]
        scale = np.abs(r) if r.ndim == 1 else np.hypot(*r.T)
        assert_vec_close(g, r, scale, rtol, f)


def test_walls(hallway, room, corridor_small):
    for scn in (hallway, room, corridor_small):
        say('walls', scn.name)
        arr = scn.agents.copy()
        # push a few agents into the walls
        arr['position'][:20, 1] = scn.bounds[1] + 0.1
        if 'position_ls' in arr.dtype.names:
            shoulders_host(arr)
        seg_dev = core.segments_to_device(scn.segments)
        _roundtrip(arr, lambda d: core.agent_obstacle(d, seg_dev),
                   lambda r: O.agent_obstacle(r, scn.segments),
                   ['force'] + (['torque'] if 'torque' in arr.dtype.names else []))


def test_adjust_orient(corridor_small, room):
    say('adjust circular')
    _roundtrip(_perturb(room.agents, 7), core.force_adjust_agents, O.force_adjust_agents,
               ['force'])
    say('adjust three-circle')
    arr = _perturb(corridor_small.agents, 8)
    arr['orientation'] = np.random.default_rng(3).uniform(-np.pi, np.pi, len(arr))
    _roundtrip(arr, lambda d: (core.orient_towards_target_direction(d),
                               core.force_adjust_agents(d), core.torque_adjust_agents(d)),
               lambda r: (O.orient_towards_target_direction(r), O.force_adjust_agents(r),
                          O.torque_adjust_agents(r)),
               ['force', 'torque', 'target_orientation'])


def test_navigation_modes(room):
    for mode in (0, 1):
        say('navigation mode', mode)
        arr = room.agents.copy()
        maps_np = room.nav['directions'][None] if mode == 0 else room.nav['potential'][None]
        maps = torch.as_tensor(np.ascontiguousarray(maps_np), device='cuda')
        # some agents outside the grid and some without a target
        arr['position'][:10] = (-3.0, 50.0)
        arr['target'][10:20] = -1
        _roundtrip(arr, lambda d: core.navigation(d, maps, 0.0, 0.0, 0.1, mode),
                   lambda r: O.navigation(r, maps_np, 0.0, 0.0, 0.1, mode),
                   ['target_direction'], rtol=0.0 if mode == 0 else OP_RTOL)


def test_fluctuation(corridor_small, hallway):
    for scn in (hallway, corridor_small):
        say('fluctuation', scn.name)
        arr = scn.agents.copy()
        _roundtrip(arr, lambda d: core.force_fluctuation(d, 17, 42),
                   lambda r: O.fluctuation(r, 17, 42),
                   ['force'] + (['torque'] if 'torque' in arr.dtype.names else []),
                   rtol=1e-13)


def test_integrator(hallway, corridor_small):
    for scn in (hallway, corridor_small):
        say('integrator', scn.name)
        arr = _perturb(scn.agents, 9)
        rng = np.random.default_rng(4)
        arr['force'] = rng.normal(0, 50, arr['force'].shape)
        arr['force_prev'] = rng.normal(0, 50, arr['force'].shape)
        if 'torque' in arr.dtype.names:
            arr['torque'] = rng.normal(0, 5, len(arr))
            arr['torque_prev'] = rng.normal(0, 5, len(arr))
            arr['angular_velocity'] = rng.normal(0, 1, len(arr))
        ref = arr.copy()
        dt_ref = O.velocity_verlet_integrator(ref, 0.001, 0.01)
        dev = DeviceAgents(arr)
        ws = Workspace(len(arr), 16, dev.model, cell_slots=1)
        core.velocity_verlet_integrator(dev, 0.001, 0.01, ws)
        st = ws.state()
        assert st.dt == dt_ref
        out = arr.copy()
        dev.download(out)
        for f in ('position', 'velocity', 'force_prev') + (
                ('orientation', 'angular_velocity', 'position_ls', 'position_rs')
                if 'torque' in arr.dtype.names else ()):
            r = ref[f]
            scale = np.abs(r) if r.ndim == 1 else np.hypot(*r.T)
            assert_vec_close(out[f], r, scale, 1e-15, f)


# -------------------------------------------------------------- full step --

def _oracle_scenario(scn, fluct=True, nav=True):
    nv = scn.nav
    return O.Scenario(cell_size=scn.cell_size, dt_min=scn.dt_min, dt_max=scn.dt_max,
                      seed=scn.seed, segments=scn.segments,
                      nav_maps=nv['maps'] if (nv and nav) else None,
                      nav_mode=nv['mode'] if nv else 0,
                      nav_origin=(nv['minx'], nv['miny']) if nv else (0.0, 0.0),
                      nav_step=nv['step'] if nv else 0.1, fluct=fluct)


def _stepper(scn, precision='f64', fluct=True):
    from crowddynamics_amd import _lib
    from crowddynamics_amd.stepper import ALL_NODES
    flags = ALL_NODES if fluct else ALL_NODES & ~_lib.CD_USE_FLUCT
    return Stepper(scn.agents.copy(), segments=scn.segments, nav=scn.nav,
                   cell_size=scn.cell_size, dt_min=scn.dt_min, dt_max=scn.dt_max,
                   seed=scn.seed, flags=flags, precision=precision, bounds=scn.bounds)


def test_step_one(hallway, room):
    """One fused step against the oracle: the hallway, and the room (its
    navigation lookup) with and without fluctuation."""
    for scn, fluct in ((hallway, True), (room, True), (room, False)):
        say('fused one step', scn.name, 'fluctuation' if fluct else 'no fluctuation')
        ref = scn.agents.copy()
        dts = O.run(ref, _oracle_scenario(scn, fluct=fluct), 1)
        st = _stepper(scn, fluct=fluct)
        st.step(1)
        out = st.download(scn.agents.copy())
        s = st.check()
        assert s.dt == dts[0]
        assert s.step == 1
        # one step moves x by v*dt: compare against the step's displacement scale
        for f in ('position', 'velocity', 'force_prev', 'target_direction'):
            r = ref[f]
            scale = np.hypot(*r.T) + 1.0
            assert_vec_close(out[f], r, scale, 1e-13, f)
        assert np.all(out['force'] == 0.0)


def _one_step_parity(scn, st, k, rtol, what, max_rtol=None, frac=1.0, force_rtol=None):
    """Compare one GPU step and one oracle step from the GPU's current state
    (step index k feeds the fluctuation counter).  With `max_rtol`, at least
    `frac` of the agents must meet `rtol` and all must meet `max_rtol`."""
    state = st.download(scn.agents.copy())
    ref = state.copy()
    O.run(ref, _oracle_scenario(scn), 1, first_step=k)
    st.step(1)
    out = st.download(scn.agents.copy())
    for f in ('position', 'velocity', 'force_prev'):
        r = ref[f]
        scale = np.hypot(*r.T) + 1.0
        tol = force_rtol if (f == 'force_prev' and force_rtol) else rtol
        if max_rtol is None:
            assert_vec_close(out[f], r, scale, tol, '%s step %d %s' % (what, k, f))
        else:
            rel = np.hypot(*(out[f] - r).T) / scale
            say('  %s step %d %s: frac within %.0e = %.5f, max %.3e'
                % (what, k, f, tol, np.mean(rel <= tol), rel.max()))
            assert np.mean(rel <= tol) >= frac
            assert rel.max() <= max_rtol


def test_step_many_hallway(hallway):
    """1000 steps of the dense hallway.  Trajectories of this contact-heavy
    crowd are chaotic (two correct summation orders of the oracle itself
    separate by centimetres within 100 steps), so parity is checked per step
    along the GPU trajectory, plus the dt sequence and the bulk flow."""
    say('fused 1000 steps hallway, one-step parity every 250 steps')
    ref = hallway.agents.copy()
    dts = O.run(ref, _oracle_scenario(hallway), 1000)
    st = _stepper(hallway)
    done = 0
    for k in (0, 250, 500, 750):
        st.step(k - done)
        done = k
        _one_step_parity(hallway, st, k, 1e-12, 'hallway')
        done += 1
    st.step(1000 - done)
    out = st.download(hallway.agents.copy())
    s = st.check()
    assert s.step == 1000
    assert abs(s.time_tot - dts.sum()) < 1e-9
    # bulk flow: mean displacement along +x agrees to 1 %
    dx_gpu = (out['position'][:, 0] - hallway.agents['position'][:, 0]).mean()
    dx_ref = (ref['position'][:, 0] - hallway.agents['position'][:, 0]).mean()
    say('mean x displacement gpu %.4f oracle %.4f' % (dx_gpu, dx_ref))
    assert abs(dx_gpu - dx_ref) < 0.01 * abs(dx_ref)


def test_step_graph_replay_matches_eager(hallway):
    say('graph replay vs eager')
    a = _stepper(hallway)
    a.step(20)
    out_a = a.download(hallway.agents.copy())
    b = _stepper(hallway)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        g = b.capture(5)
        for _ in range(4):
            g.launch()
        s.synchronize()
    out_b = b.download(hallway.agents.copy())
    assert np.array_equal(out_a['position'], out_b['position'])
    assert np.array_equal(out_a['velocity'], out_b['velocity'])
    assert b.state().step == 20


def test_step_deterministic(room):
    say('determinism room 20 steps')
    outs = []
    for _ in range(2):
        st = _stepper(room)
        st.step(20)
        outs.append(st.download(room.agents.copy()))
    assert np.array_equal(outs[0]['position'], outs[1]['position'])
    assert np.array_equal(outs[0]['velocity'], outs[1]['velocity'])


def test_step_room_fp64_and_fp32(room):
    """20 steps of the 20k room: fp64 trajectory within 1e-9 m of the oracle;
    fp32 mode checked per step (one fp32 step vs one fp64 oracle step from the
    same state: 1e-6 relative) and by the median trajectory deviation."""
    say('room 20 steps fp64 / fp32 vs oracle')
    ref = room.agents.copy()
    O.run(ref, _oracle_scenario(room), 20)
    st = _stepper(room, precision='f64')
    st.step(20)
    out = st.download(room.agents.copy())
    st.check()
    err = np.abs(out['position'] - ref['position']).max()
    say('  f64 max |dx| after 20 steps: %.3e m' % err)
    assert err < 1e-9
    st = _stepper(room, precision='f32')
    st.step(10)
    # fp32 forces sit within 2e-4 of the fp64 force scale; the h < 3 sight
    # gate and the contact onset are discontinuous, so a rounding flip moves
    # an agent's step by up to the gated term: bound 99.9 % tightly, all loosely
    # (the step's total force sums terms up to 2 kN in fp32: 1e-4 relative to
    # |F| + 1 N for 99.9 % of the agents)
    _one_step_parity(room, st, 10, 1e-6, 'fp32', max_rtol=1e-2, frac=0.999, force_rtol=1e-4)
    st.step(9)
    out = st.download(room.agents.copy())
    st.check()
    dev = np.hypot(*(out['position'] - ref['position']).T)
    say('  f32 after 20 steps: median |dx| %.3e m, max %.3e m' % (np.median(dev), dev.max()))
    assert np.median(dev) < 1e-5


def test_step_processing_order(room, corridor_small, hallway):
    """The agent-parallel force kernel serving agents in cell order (rebuilt
    every step, or every fourth) equals index order bit for bit: room fp64
    and fp32 (10 steps), three-circle corridor fp64 (5 steps), and the
    hallway with 2 slots per cell, so that overfull cells take the exact slow
    path (5 steps)."""
    import warnings
    from crowddynamics_amd.stepper import ALL_NODES
    for scn, prec, steps, slots in ((room, 'f64', 10, None), (room, 'f32', 10, None),
                                    (corridor_small, 'f64', 5, None), (hallway, 'f64', 5, 2)):
        outs = []
        for every in (0, 1, 4):
            say('%s %s %d steps, order_every %d, slots %s' % (scn.name, prec, steps, every, slots))
            st = Stepper(scn.agents.copy(), segments=scn.segments, nav=scn.nav,
                         cell_size=scn.cell_size, dt_min=scn.dt_min, dt_max=scn.dt_max,
                         seed=scn.seed, precision=prec, bounds=scn.bounds, flags=ALL_NODES,
                         order_every=every, cell_slots=slots)
            st.step(steps)
            with warnings.catch_warnings(record=True) as caught:
                warnings.simplefilter('always')
                st.check()
            if slots is not None:  # the slow path was taken
                assert any('slots' in str(w.message) for w in caught)
            outs.append((st.download(scn.agents.copy()), st.state().time_tot))
        for out, tt in outs[1:]:
            assert tt == outs[0][1]
            for f in ('position', 'velocity', 'force_prev', 'target_direction'):
                assert np.array_equal(out[f], outs[0][0][f]), (scn.name, prec, f)


def test_step_wide_lane_groups(hallway, corridor_small):
    """2, 32 and 64 lanes per agent (32 agents per wave, one agent per half
    wave or per wave) give the 4-lane step bit for bit: hallway (circular)
    and corridor (three-circle, fp64 and fp32), 5 fused steps."""
    for scn, prec in ((hallway, 'f64'), (corridor_small, 'f64'), (corridor_small, 'f32')):
        outs = []
        for lanes in (4, 2, 32, 64):
            say('%s %s 5 steps, %d lanes per agent' % (scn.name, prec, lanes))
            st = Stepper(scn.agents.copy(), segments=scn.segments, nav=scn.nav,
                         cell_size=scn.cell_size, dt_min=scn.dt_min, dt_max=scn.dt_max,
                         seed=scn.seed, bounds=scn.bounds, lanes=lanes, precision=prec)
            st.step(5)
            st.check()
            outs.append(st.download(scn.agents.copy()))
        for out in outs[1:]:
            for f in ('position', 'velocity', 'force_prev'):
                assert np.array_equal(out[f], outs[0][f]), (scn.name, prec, f)


@pytest.mark.parametrize('slots', [None, 1024])
def test_step_fp32_pair_sums_fold(slots):
    """fp32 pair terms gather in 64-bit integers per lane and join the
    128-bit sums every 2048 pairs.  5000 overlapping agents in a 3 x 3 m
    square on 1 m cells: an agent of the centre cell has all 5000 in its
    stencil and within sight, so at 2 lanes per agent its lanes add about
    2500 pairs each (one intermediate fold); at 64 lanes about 80 (none).
    The fused step must agree bit for bit, through the cells' slot records
    (1024 slots, about 560 agents a cell) and through the overfull-cell path
    (the default 16 slots)."""
    import warnings
    rng = np.random.default_rng(5)
    n = 5000
    arr = E._attributes(rng, n)
    arr['position'] = rng.uniform(0.0, 3.0, (n, 2))
    arr['velocity'] = rng.normal(0.0, 0.5, (n, 2))
    arr['target_direction'] = (1.0, 0.0)
    arr['target'] = 0
    outs = []
    for lanes in (2, 64):
        say('dense crowd fp32, slots %s, %d lanes per agent' % (slots, lanes))
        st = Stepper(arr.copy(), cell_size=1.0, dt_min=0.001, dt_max=0.01, seed=3,
                     bounds=(0.0, 0.0, 3.0, 3.0), lanes=lanes, precision='f32',
                     cell_slots=slots)
        st.step(1)
        with warnings.catch_warnings(record=True) as caught:
            warnings.simplefilter('always')
            st.check()
        assert (len(caught) > 0) == (slots is None)  # the overfull-cell path ran
        outs.append(st.download(arr.copy()))
    for f in ('position', 'velocity', 'force_prev'):
        assert np.all(np.isfinite(outs[0][f])), f
        assert np.array_equal(outs[1][f], outs[0][f]), f


def test_step_corridor_fp32(corridor_small):
    """The fp32 mode of three-circle agents against the fp64 oracle, one
    step at a time along the GPU trajectory (steps 0, 4 and 8 of the
    3000-agent corridor): velocity, force and torque in fp32 arithmetic,
    position, orientation and torque accumulation in fp64."""
    st = _stepper(corridor_small, precision='f32')
    done = 0
    for k in (0, 4, 8):
        st.step(k - done)
        done = k + 1
        state = st.download(corridor_small.agents.copy())
        ref = state.copy()
        O.run(ref, _oracle_scenario(corridor_small), 1, first_step=k)
        st.step(1)
        out = st.download(corridor_small.agents.copy())
        for f in ('position', 'velocity', 'force_prev'):
            scale = np.hypot(*ref[f].T) + 1.0
            rel = np.hypot(*(out[f] - ref[f]).T) / scale
            say('  fp32 corridor step %d %s: frac within 1e-4 %.5f, within 1e-6 %.5f, max %.3e'
                % (k, f, np.mean(rel <= 1e-4), np.mean(rel <= 1e-6), rel.max()))
            assert np.mean(rel <= 1e-4) >= 0.999
            assert rel.max() <= (5e-2 if f == 'force_prev' else 1e-3)
        for f in ('torque_prev', 'angular_velocity', 'orientation'):
            scale = np.abs(ref[f]) + 1.0
            rel = np.abs(out[f] - ref[f]) / scale
            say('  fp32 corridor step %d %s: frac within 1e-4 %.5f, max %.3e'
                % (k, f, np.mean(rel <= 1e-4), rel.max()))
            assert np.mean(rel <= 1e-4) >= 0.99
    st.check()


def _pairs_off_stepper(scn, tile=None):
    from crowddynamics_amd import _lib
    from crowddynamics_amd.stepper import ALL_NODES
    return Stepper(scn.agents.copy(), segments=scn.segments, nav=scn.nav,
                   cell_size=scn.cell_size, dt_min=scn.dt_min, dt_max=scn.dt_max,
                   seed=scn.seed, flags=ALL_NODES & ~_lib.CD_USE_AGENT_AGENT,
                   bounds=scn.bounds, tile=tile)


def test_step_without_pairs(hallway):
    """AgentAgentInteractions off: the force kernel still adds the walls and
    integrates every agent.  Five steps against the oracle's tree without
    the node, with the agent-parallel kernel and with the tiled kernel's
    path for a step without binning (the two agree bit for bit)."""
    say('fused 5 steps without agent-agent pairs')
    nv = hallway.nav
    sc = O.Scenario(cell_size=hallway.cell_size, dt_min=hallway.dt_min,
                    dt_max=hallway.dt_max, seed=hallway.seed, segments=hallway.segments,
                    nav_maps=nv['maps'] if nv else None, nav_mode=nv['mode'] if nv else 0,
                    nav_origin=(nv['minx'], nv['miny']) if nv else (0.0, 0.0),
                    nav_step=nv['step'] if nv else 0.1, agent_agent=False)
    ref = hallway.agents.copy()
    dts = O.run(ref, sc, 5)
    outs = []
    for tile in (None, (2, 2)):
        st = _pairs_off_stepper(hallway, tile)
        st.step(5)
        s = st.check()
        assert s.step == 5
        assert abs(s.time_tot - float(np.sum(dts))) < 1e-12
        outs.append(st.download(hallway.agents.copy()))
    for f in ('position', 'velocity', 'force_prev'):
        assert np.array_equal(outs[0][f], outs[1][f]), f
        r = ref[f]
        assert_vec_close(outs[0][f], r, np.hypot(*r.T) + 1.0, 1e-12, 'pairs off ' + f)
    assert np.all(outs[0]['force'] == 0.0)


@pytest.mark.parametrize('order_every', [0, 1])
def test_step_grid_overflow_reported(hallway, order_every):
    """A cell capacity far below the grid: the step's binning is invalid, the
    force kernel adds only the walls and integrates, and the overflow is
    reported (GridOverflow) instead of touching cells beyond the capacity.
    In cell order the order is not rebuilt from the invalid binning, so every
    agent is still integrated exactly once: the run equals index order."""
    from crowddynamics_amd.stepper import GridOverflow
    say('fused step with a grid beyond the cell capacity, order_every %d' % order_every)
    outs = []
    for every in sorted({0, order_every}):
        st = Stepper(hallway.agents.copy(), segments=hallway.segments,
                     cell_size=hallway.cell_size, dt_min=hallway.dt_min,
                     dt_max=hallway.dt_max, seed=hallway.seed, cell_capacity=4,
                     bounds=hallway.bounds, order_every=every)
        st.step(2)
        with pytest.raises(GridOverflow):
            st.check()
        out = st.download(hallway.agents.copy())
        assert np.all(np.isfinite(out['position']))
        assert st.state().step == 2
        outs.append(out)
    for f in ('position', 'velocity'):
        assert np.array_equal(outs[-1][f], outs[0][f])


def test_step_stale_box_reported(hallway):
    """Positions edited on the host between steps without prepare(): an agent
    now lies outside the grid of the bounding box the previous step reduced.
    The step's first kernel cannot bin it, the force kernel skips the pairs
    (walls and integration still run) and the condition is reported as
    GridOverflow.  After prepare() the run continues cleanly."""
    from crowddynamics_amd.stepper import GridOverflow
    say('fused step after a host edit outside the reduced bounding box')
    st = Stepper(hallway.agents.copy(), segments=hallway.segments,
                 cell_size=hallway.cell_size, dt_min=hallway.dt_min, dt_max=hallway.dt_max,
                 seed=hallway.seed, bounds=hallway.bounds)
    st.step(1)
    st.check()
    arr = st.download(hallway.agents.copy())
    arr['position'][5] = arr['position'][5] + np.array([40.0, 0.0])
    st.agents.upload(arr, ['position'])
    st.step(1)
    with pytest.raises(GridOverflow):
        st.check()
    out = st.download(hallway.agents.copy())
    assert np.all(np.isfinite(out['position']))
    assert st.state().step == 2
    out['position'][5] = out['position'][5] - np.array([40.0, 0.0])
    st.agents.upload(out, ['position'])
    st.ws.clear_status()
    st.prepare(2)
    st.step(1)
    st.check()


@pytest.mark.parametrize('tile', [(1, 1), (3, 2)])
def test_step_tiled_equals_direct(hallway, tile):
    """Ten fused steps with the tiled force kernel (which integrates the
    agents of each tile it serves) equal the agent-parallel kernel bit for
    bit."""
    say('fused 10 steps, tiled %s vs agent-parallel' % (tile,))
    outs = []
    for t in (None, tile):
        st = Stepper(hallway.agents.copy(), segments=hallway.segments, nav=hallway.nav,
                     cell_size=hallway.cell_size, dt_min=hallway.dt_min,
                     dt_max=hallway.dt_max, seed=hallway.seed, bounds=hallway.bounds, tile=t)
        st.step(10)
        st.check()
        outs.append(st.download(hallway.agents.copy()))
    for f in ('position', 'velocity', 'force_prev', 'target_direction'):
        assert np.array_equal(outs[0][f], outs[1][f]), f


def test_step_wall_contacts(hallway, corridor_small):
    """Agents pushed into the walls: one fused step (whose force kernel skips
    the segments out of an agent's reach by their bounding boxes) against
    the oracle's step, circular and three-circle."""
    for scn in (hallway, corridor_small):
        say('fused step with wall contacts', scn.name)
        arr = scn.agents.copy()
        arr['position'][:20, 1] = scn.bounds[1] + 0.1
        if 'position_ls' in arr.dtype.names:
            shoulders_host(arr)
        ref = arr.copy()
        O.run(ref, _oracle_scenario(scn, fluct=False), 1)
        from crowddynamics_amd import _lib
        from crowddynamics_amd.stepper import ALL_NODES
        st = Stepper(arr.copy(), segments=scn.segments, nav=scn.nav, cell_size=scn.cell_size,
                     dt_min=scn.dt_min, dt_max=scn.dt_max, seed=scn.seed,
                     flags=ALL_NODES & ~_lib.CD_USE_FLUCT, bounds=scn.bounds)
        st.step(1)
        st.check()
        out = st.download(arr.copy())
        # the pushed agents are in contact: their step is dominated by the wall force
        moved = np.hypot(*(ref['force_prev'][:20]).T)
        assert np.all(moved > 1.0)
        for f in ('position', 'velocity', 'force_prev'):
            r = ref[f]
            assert_vec_close(out[f], r, np.hypot(*r.T) + 1.0, 1e-12, '%s %s' % (scn.name, f))


def _many_walls(scn, n_extra, seed):
    """`scn` with n_extra short segments (0.6 m, random) inside its bounds
    after its own, and agents 0-29 centred 5 cm off the middle of the first
    30 of them (a centre on a segment has no contact normal)."""
    import dataclasses
    rng = np.random.default_rng(seed)
    x0, y0, x1, y1 = scn.bounds
    c = rng.uniform((x0 + 0.5, y0 + 0.5), (x1 - 0.5, y1 - 0.5), (n_extra, 2))
    a = rng.uniform(0.0, np.pi, n_extra)
    h = 0.3 * np.stack([np.cos(a), np.sin(a)], 1)
    extra = np.zeros(n_extra, dtype=core.obstacle_type_linear)
    extra['p0'] = c - h
    extra['p1'] = c + h
    own = np.zeros(len(scn.segments), dtype=core.obstacle_type_linear)
    own['p0'] = scn.segments['p0']
    own['p1'] = scn.segments['p1']
    arr = scn.agents.copy()
    arr['position'][:30] = c[:30] + 0.05 * np.stack([-np.sin(a[:30]), np.cos(a[:30])], 1)
    if 'position_ls' in arr.dtype.names:
        shoulders_host(arr)
    return dataclasses.replace(scn, agents=arr, segments=np.concatenate([own, extra]))


@pytest.mark.parametrize('which,n_extra', [('hallway', 300), ('corridor', 360)])
def test_step_many_walls(hallway, corridor_small, which, n_extra):
    """More wall segments than the force kernel stages in LDS at once (tiles
    of 256): 302 segments with circular agents, 404 with three-circle agents,
    30 agents in contact with segments.  One step against the oracle; five
    steps with the segments in the given and in reversed order (the wall
    terms join exact sums), agent-parallel and (hallway) tiled, bit-identical."""
    from crowddynamics_amd import _lib
    from crowddynamics_amd.stepper import ALL_NODES
    scn = _many_walls(hallway if which == 'hallway' else corridor_small, n_extra, 5)
    say('%s with %d segments: one step against the oracle' % (which, len(scn.segments)))
    ref = scn.agents.copy()
    O.run(ref, _oracle_scenario(scn, fluct=False), 1)
    st = Stepper(scn.agents.copy(), segments=scn.segments, nav=scn.nav, cell_size=scn.cell_size,
                 dt_min=scn.dt_min, dt_max=scn.dt_max, seed=scn.seed,
                 flags=ALL_NODES & ~_lib.CD_USE_FLUCT, bounds=scn.bounds)
    st.step(1)
    st.check()
    out = st.download(scn.agents.copy())
    assert np.all(np.hypot(*(ref['force_prev'][:30]).T) > 1.0)  # in contact
    for f in ('position', 'velocity', 'force_prev'):
        r = ref[f]
        assert_vec_close(out[f], r, np.hypot(*r.T) + 1.0, 1e-12, '%s %s' % (which, f))
    outs = []
    # the segments binned into wall cells (the default at this many
    # segments: 1 m and 0.4 m cells) and scanned whole (wall_grid None)
    runs = ((False, None, 'auto'), (True, None, 'auto'), (False, None, None),
            (True, None, 0.4)) + (((True, (3, 2), 'auto'),) if which == 'hallway' else ())
    for rev, tile, grid in runs:
        say('%s: 5 steps, segments reversed %s, tile %s, wall grid %s' % (which, rev, tile, grid))
        segs = scn.segments[::-1].copy() if rev else scn.segments
        st = Stepper(scn.agents.copy(), segments=segs, nav=scn.nav, cell_size=scn.cell_size,
                     dt_min=scn.dt_min, dt_max=scn.dt_max, seed=scn.seed, bounds=scn.bounds,
                     tile=tile, wall_grid=grid)
        assert (st.wall_cells is None) == (grid is None)
        st.step(5)
        st.check()
        outs.append(st.download(scn.agents.copy()))
    for out in outs[1:]:
        for f in ('position', 'velocity', 'force_prev', 'target_direction'):
            assert np.array_equal(out[f], outs[0][f]), (which, f)
