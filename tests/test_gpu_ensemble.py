"""Batched ensemble (cd_batch_*): every scenario of a batch evolves exactly
as its standalone run.

Each scenario keeps its own grid, walls, dt, elapsed time and noise stream,
and its agents' pair sums run over the same cells in the same order as in
the standalone step, so the bar is bit-for-bit equality of the agent
records and of dt / time_tot, in fp64 and in the fp32 mode.
"""
import numpy as np
import pytest
import torch

from crowddynamics_amd import examples as E
from crowddynamics_amd.ensemble import Ensemble
from crowddynamics_amd.stepper import Stepper

pytestmark = pytest.mark.gpu

FIELDS = ('position', 'velocity', 'force', 'force_prev', 'target_direction')
TC_FIELDS = FIELDS + ('orientation', 'angular_velocity', 'torque', 'torque_prev',
                      'position_ls', 'position_rs')


def say(*a):
    print(*a, flush=True)


def _small_ensemble(S=7, seed=3):
    scns = E.ensemble(n_scenarios=S, seed=seed, n_min=20, n_max=700, rho=(0.5, 3.0))
    # an empty scenario and a single agent between the others
    empty = E.Scenario('empty', scns[0].agents[:0].copy(), scns[0].segments, 3.6,
                       0.001, 0.01, 10, seed=99, bounds=(0.0, 0.0, 5.0, 5.0))
    one = E.Scenario('one', scns[1].agents[:1].copy(), scns[1].segments, 3.6,
                     0.001, 0.01, 10, seed=98, bounds=scns[1].bounds)
    return scns[:3] + [empty] + scns[3:5] + [one] + scns[5:]


def _standalone(scn, steps, precision, cell_size, stream):
    with torch.cuda.stream(stream):
        st = Stepper(scn.agents.copy(), segments=scn.segments, nav=scn.nav,
                     cell_size=cell_size, dt_min=scn.dt_min, dt_max=scn.dt_max,
                     seed=scn.seed, precision=precision, bounds=scn.bounds)
        st.step(steps)
        state = st.check()
        out = st.download(scn.agents.copy())
    return out, state


def _compare(outs, refs, fields):
    for k, (o, r) in enumerate(zip(outs, refs)):
        for f in fields:
            if f in o.dtype.names:
                assert np.array_equal(o[f], r[f]), (k, f, np.max(np.abs(o[f] - r[f])))


@pytest.mark.parametrize('precision', ['f64', 'f32'])
def test_ensemble_equals_standalone(precision):
    scns = _small_ensemble()
    steps = 40
    stream = torch.cuda.Stream()
    say('ensemble of', [s.n for s in scns], precision)
    with torch.cuda.stream(stream):
        ens = Ensemble(scns, precision=precision)
        ens.step(steps)
        st = ens.check()
        outs = ens.download_scenarios()
        states = ens.states()
    assert st.step == steps
    refs = []
    for k, s in enumerate(scns):
        if s.n == 0:
            refs.append(s.agents.copy())
            assert states[k].time_tot == pytest.approx(steps * s.dt_max)
            continue
        out, state = _standalone(s, steps, precision, ens.cell_size, stream)
        refs.append(out)
        assert states[k].dt == state.dt and states[k].time_tot == state.time_tot, k
        assert tuple(states[k].shape) == tuple(state.shape), k
    _compare(outs, refs, FIELDS)


@pytest.mark.parametrize('cell_size', [3.6, 8.0])
def test_ensemble_graph_replay_and_tiles(cell_size):
    """Graph replays equal eager batched steps, and the force kernel changes
    nothing, bit for bit: the agent-parallel kernel with 4 to 64 lanes per
    agent, in index order or in cell order (rebuilt every step or every third),
    and LDS tiles of 1 x 1 up to 16 x 16 cells (at 8 m cells the large tiles exceed the LDS
    capacity and fall back to 3 x 3 regions)."""
    scns = _small_ensemble(S=5, seed=11)
    stream = torch.cuda.Stream()
    res = []
    for tile, lanes, graph, every in (((0, 0), 4, False, 0), ((0, 0), 4, True, 0),
                                      ((0, 0), 8, False, 0), ((0, 0), 16, True, 0),
                                      ((0, 0), 4, False, 1), ((0, 0), 8, True, 3),
                                      ((0, 0), 32, False, 0), ((0, 0), 64, True, 2),
                                      ((1, 1), 4, False, 0), ((3, 5), 4, True, 0),
                                      ((16, 16), 4, False, 0)):
        say('tile', tile, 'lanes', lanes, 'graph', graph, 'order_every', every, 'cell', cell_size)
        with torch.cuda.stream(stream):
            ens = Ensemble(scns, precision='f32', tile=tile, lanes=lanes, cell_size=cell_size,
                           order_every=every)
            if graph:
                g = ens.capture(10)
                for _ in range(3):
                    g.launch()
                    stream.synchronize()
                del g
            else:
                ens.step(30)
            ens.check()
            res.append((ens.download_scenarios(), [(s.dt, s.time_tot) for s in ens.states()]))
    for outs, st in res[1:]:
        _compare(outs, res[0][0], FIELDS)
        assert st == res[0][1]


@pytest.mark.parametrize('precision', ['f64', 'f32'])
def test_ensemble_three_circle(precision):
    scns = [E.corridor(n=n, seed=k, width=25.0 + 5 * k, height=12.0, n_obstacles=4)
            for k, n in enumerate((120, 40, 260))]
    for k, s in enumerate(scns):
        s.seed = 5 + k
    steps = 25
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        ens = Ensemble(scns, precision=precision)
        ens.step(steps)
        ens.check()
        outs = ens.download_scenarios()
        states = ens.states()
    refs = []
    for k, s in enumerate(scns):
        out, state = _standalone(s, steps, precision, ens.cell_size, stream)
        refs.append(out)
        assert states[k].time_tot == state.time_tot
    _compare(outs, refs, TC_FIELDS)


def test_autotuned_cell_sizes_are_physics_neutral():
    """Candidate cell sizes of the autotuner give the same step to rounding
    (fp64: positions after 5 steps within 1e-12 relative)."""
    from crowddynamics_amd.autotune import max_body_radius, neutral_cell_sizes, tune_ensemble
    scns = _small_ensemble(S=4, seed=21)
    sizes = neutral_cell_sizes(max_body_radius(scns))
    assert sizes[0] >= 3.0 + 2 * max_body_radius(scns)
    stream = torch.cuda.Stream()
    outs = []
    for cs in (sizes[0], sizes[-1]):
        with torch.cuda.stream(stream):
            ens = Ensemble(scns, precision='f64', cell_size=cs)
            ens.step(5)
            ens.check()
            outs.append(ens.download_scenarios())
    for a, b in zip(*outs):
        if len(a):
            scale = np.max(np.abs(a['position'])) + 1.0
            assert np.max(np.abs(a['position'] - b['position'])) <= 1e-12 * scale
            assert np.max(np.abs(a['velocity'] - b['velocity'])) <= 1e-9
    with torch.cuda.stream(stream):
        best = tune_ensemble(scns, steps=3, warmup=1, cell_sizes=sizes[:2],
                             kernels=((0, 0, 8), (0, 0, 4, 5), (2, 2, 4)))
    say('autotune', best['cell_size'], best['tile'], best['lanes'], best['order_every'],
        best['ms_per_step'])
    assert best['cell_size'] in sizes[:2] and best['tile'] in ((0, 0), (2, 2))
    assert best['order_every'] in (0, 5)
    assert len(best['table']) == 6


@pytest.mark.parametrize('precision', ['f64', 'f32'])
def test_batch_of_one_room_equals_standalone(precision):
    """A batch of one 5000-agent room (navigation and fluctuation on) equals
    its standalone run bit for bit over 20 steps: the batched step bins and
    adds the own terms on the same wave, the standalone step on two waves
    per 64 agents."""
    scn = E.room(n=5000, width=100.0, height=50.0)
    steps = 20
    stream = torch.cuda.Stream()
    say('batch of one room,', scn.n, 'agents,', precision)
    with torch.cuda.stream(stream):
        ens = Ensemble([scn], precision=precision, cell_size=scn.cell_size)
        ens.step(steps)
        ens.check()
        outs = ens.download_scenarios()
        states = ens.states()
    out, state = _standalone(scn, steps, precision, ens.cell_size, stream)
    assert states[0].dt == state.dt and states[0].time_tot == state.time_tot
    _compare(outs, [out], FIELDS)
