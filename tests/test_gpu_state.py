"""Device state nodes, herding neighbours and snapshots (csrc/cd_state.hip)
against the reference's own routine (matplotlib's Path.contains_points,
which InsideDomain and TargetReached call) and the oracle's restatement of
collective_motion.py (pinned in tests/test_collective_host.py)."""
import numpy as np
import pytest
import torch

from crowddynamics_amd import examples as E
from crowddynamics_amd import state as S
from crowddynamics_amd.agents import empty_agents
from crowddynamics_amd.device import DeviceAgents, Workspace, cell_capacity_for
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _polygon_points(rng, verts, n):
    """Random points around the polygon, plus its vertices, edge midpoints
    and points on the edges' lines (the crossings test's tie cases)."""
    lo, hi = verts.min(0) - 1.0, verts.max(0) + 1.0
    pts = [rng.uniform(lo, hi, (n, 2)), verts]
    nxt = np.roll(verts, -1, 0)
    for t in (0.0, 0.25, 0.5, 1.0):
        pts.append(verts + t * (nxt - verts))
    pts.append(np.stack([rng.uniform(lo[0], hi[0], 200), np.repeat(verts[:, 1], 200 // len(verts) + 1)[:200]], 1))
    return np.concatenate(pts)


@pytest.mark.parametrize('shape', ['rectangle', 'concave'])
def test_inside_domain_matches_contains_points(shape):
    from matplotlib.path import Path
    rng = np.random.default_rng(3)
    if shape == 'rectangle':
        verts = np.array([(0.0, 0.0), (50.0, 0.0), (50.0, 5.0), (0.0, 5.0), (0.0, 0.0)])
    else:
        verts = np.array([(0.0, 0.0), (10.0, 0.0), (10.0, 8.0), (6.0, 3.0), (3.0, 8.0),
                          (0.0, 8.0)])
    pts = _polygon_points(rng, verts, 20000)
    arr = empty_agents(len(pts))
    arr['position'] = pts
    arr['active'] = rng.uniform(size=len(pts)) < 0.5
    before = arr['active'].copy()
    dev = DeviceAgents(arr)
    counter = torch.zeros(1, dtype=torch.int64, device='cuda')
    S.inside_domain(dev, verts, counter)
    out = arr.copy()
    dev.download(out)
    want = Path(verts).contains_points(pts)
    say(shape, 'inside', int(want.sum()), 'of', len(pts))
    assert np.array_equal(out['active'], want)
    assert int(counter.item()) == int(np.sum(before ^ want))


def test_target_reached_accumulates():
    from matplotlib.path import Path
    rng = np.random.default_rng(4)
    target = np.array([(8.0, 1.0), (10.0, 1.0), (10.0, 4.0), (8.0, 4.0)])
    arr = empty_agents(5000)
    arr['position'] = rng.uniform((0, 0), (12, 5), (5000, 2))
    dev = DeviceAgents(arr)
    reached = torch.zeros(5000, dtype=torch.uint8, device='cuda')
    count = torch.zeros(1, dtype=torch.int64, device='cuda')
    ever = np.zeros(5000, dtype=bool)
    for step in range(3):
        S.target_reached(dev, target, reached, count)
        ever |= Path(target).contains_points(arr['position'])
        assert np.array_equal(reached.cpu().numpy().astype(bool), ever)
        assert int(count.item()) == int(ever.sum())
        arr['position'][:, 0] += 1.0
        dev.upload(arr, fields=('position',))


def test_nearest_neighbors_and_herding_equal_oracle():
    rng = np.random.default_rng(7)
    n = 1500
    arr = empty_agents(n)
    arr['position'] = rng.uniform((0, 0), (60, 40), (n, 2))
    arr['velocity'] = rng.normal(0, 1, (n, 2))
    arr['velocity'][::17] = 0.0
    segs = np.array([(10.0, 5.0, 30.0, 12.0), (40.0, 0.0, 40.0, 30.0), (5.0, 30.0, 55.0, 31.0)])
    sight, k = 5.0, 5
    dev = DeviceAgents(arr)
    ws = Workspace(n, cell_capacity_for(60, 40, sight), 0, cell_slots=32)
    nb = S.find_nearest_neighbors(dev, sight, k, segs, ws)
    want = O.nearest_neighbors(arr['position'], sight, k, segs)
    got = nb.cpu().numpy()
    assert np.array_equal(got, want)
    say('neighbours: %d of %d slots filled' % ((got >= 0).sum(), got.size))
    follower = rng.uniform(size=n) < 0.6
    d, has = S.herding_interaction(dev, follower, nb, 0.15, 0.45 * np.pi)
    wd, wh = O.herding_interaction(follower, arr['position'], arr['velocity'], want, 0.15,
                                   0.45 * np.pi)
    assert np.array_equal(has.cpu().numpy(), wh)
    np.testing.assert_allclose(d.cpu().numpy(), wd, rtol=0, atol=1e-14)
    say('herding: %d followers with a direction' % wh.sum())


def test_snapshots_write_the_recorded_states(tmp_path):
    from crowddynamics_amd.stepper import Stepper
    scn = E.hallway(n=60, width=10.0, height=5.0)
    st = Stepper(scn.agents.copy(), segments=scn.segments, cell_size=scn.cell_size,
                 dt_min=scn.dt_min, dt_max=scn.dt_max, seed=scn.seed, bounds=scn.bounds)
    snap = S.Snapshots(st.agents, scn.agents, str(tmp_path), capacity=8)
    states = []
    for _ in range(3):
        st.step(1)
        snap.record()
        states.append(st.download(scn.agents.copy()))
    path = snap.dump()
    saved = np.load(path)
    assert saved.shape == (3, scn.n) and saved.dtype == scn.agents.dtype
    for k in range(3):
        for f in ('position', 'velocity', 'force_prev', 'active'):
            assert np.array_equal(saved[k][f], states[k][f]), (k, f)
    assert snap.dump() is None


def test_hallway_tree_inside_domain_on_device():
    """The Hallway tree with InsideDomain (run after the fused step):
    agents moved outside the domain turn inactive and data['inactive']
    counts them."""
    sim = E.Hallway(size=40)
    sim.update()
    arr = sim.agents.array
    arr['position'][:3] = (-5.0, 2.0)
    sim.update()
    assert sim.data['inactive'] == 3
    assert not sim.agents.array['active'][:3].any()
    assert sim.agents.array['active'][3:].all()
