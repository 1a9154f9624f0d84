"""The device navigation field (cd_distance_map, cd_direction_map,
cd_fill_missing, cd_obstacle_handling) against the oracle
(oracle/cd_field.c, itself pinned in tests/test_field_host.py).

The device relaxes the upwind update to its fixed point and the oracle
marches it in order of distance; both apply the same update to the same
values, so the maps agree to rounding, and the direction maps (differences
and the library hypot) bit for bit from equal distance maps."""
import numpy as np
import pytest
import torch

from crowddynamics_amd import examples as E
from crowddynamics_amd import field as F
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _close(got, want, rtol=1e-12):
    got = got.cpu().numpy() if isinstance(got, torch.Tensor) else got
    assert np.array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    scale = np.max(np.abs(want[ok])) if ok.any() else 1.0
    err = np.max(np.abs(got[ok] - want[ok])) if ok.any() else 0.0
    assert err <= rtol * scale, (err, scale)
    return err / scale


def _room_geometry(width=200.0, height=100.0, exit_width=1.2):
    scn = E.room(n=10, width=width, height=height, exit_width=exit_width)
    walls = np.concatenate([scn.segments['p0'], scn.segments['p1']], 1)
    return walls, np.array(scn.extra['exits'], dtype=np.float64)


@pytest.mark.parametrize('case', ['hallway', 'room', 'maze'])
def test_distance_map_equals_oracle(case):
    rng = np.random.default_rng(2)
    mask = None
    if case == 'hallway':
        mg = F.meshgrid(0.1, 0.0, 0.0, 50.0, 5.0)
        phi = F.target_phi(mg, [(50.0, 0.0, 50.0, 5.0)])
    elif case == 'room':
        mg = F.meshgrid(0.1, 0.0, 0.0, 200.0, 100.0)
        walls, exits = _room_geometry()
        phi = F.target_phi(mg, exits)
        mask = F.grown_mask(mg, walls, 0.5)
    else:
        mg = F.meshgrid(0.25, 0.0, 0.0, 60.0, 40.0)
        phi = F.target_phi(mg, [(5.0, 5.0, 5.0, 5.0), (50.0, 30.0, 55.0, 35.0)])
        rows, cols = mg.shape
        mask = rng.uniform(size=(rows, cols)) < 0.25
        mask[phi > 0] = False
    say(case, 'grid', mg.shape)
    fb = F.FieldBuilder(mg)
    got = fb.distance_map(phi, mask)
    torch.cuda.synchronize()
    want = O.distance_map(phi, mg.step, mask)
    rel = _close(got, want)
    say(case, 'max |device - oracle| / max |oracle| = %.3e' % rel)
    ex, ey = fb.direction_map(got)
    wx, wy = O.direction_map(got.cpu().numpy())
    np.testing.assert_array_equal(ex.cpu().numpy(), wx)
    np.testing.assert_array_equal(ey.cpu().numpy(), wy)


def test_fill_and_obstacle_handling_equal_oracle():
    mg = F.meshgrid(0.1, 0.0, 0.0, 60.0, 30.0)
    walls, exits = _room_geometry(60.0, 30.0)
    fb = F.FieldBuilder(mg)
    grown = F.grown_mask(mg, walls, 0.5)
    dt_ = fb.distance_map(F.target_phi(mg, exits), grown)
    tx, ty = fb.direction_map(dt_)
    bare = np.zeros(mg.shape, dtype=bool)
    F.draw_segments(bare, mg, walls, True)
    fill = grown & ~bare
    reach = 7
    fx, fy = fb.fill_missing(tx, ty, fill, reach)
    # the nearest node with a direction, by brute force on the host
    sx, sy = tx.cpu().numpy(), ty.cpu().numpy()
    gx, gy = fx.cpu().numpy(), fy.cpu().numpy()
    have = ~(np.isnan(sx) | np.isnan(sy))
    rows, cols = mg.shape
    todo = np.argwhere(fill & ~have)
    assert len(todo) > 100
    for r, c in todo[:: max(1, len(todo) // 400)]:
        r0, r1 = max(0, r - reach), min(rows - 1, r + reach)
        c0, c1 = max(0, c - reach), min(cols - 1, c + reach)
        best = None
        for rr in range(r0, r1 + 1):
            for cc in range(c0, c1 + 1):
                if have[rr, cc]:
                    d2 = (rr - r) ** 2 + (cc - c) ** 2
                    if best is None or d2 < best[0]:
                        best = (d2, rr, cc)
        want = (sx[best[1], best[2]], sy[best[1], best[2]]) if best else (np.nan, np.nan)
        np.testing.assert_array_equal((gx[r, c], gy[r, c]), want)
    untouched = ~(fill & ~have)
    np.testing.assert_array_equal(gx[untouched], sx[untouched])
    # obstacle handling
    do = fb.distance_map(F.target_phi(mg, walls))
    ox, oy = fb.direction_map(do)
    bx, by = fx.clone(), fy.clone()
    fb.obstacle_handling(do, ox, oy, bx, by, 0.5, 0.3)
    wx, wy = O.obstacle_handling(do.cpu().numpy(), ox.cpu().numpy(), oy.cpu().numpy(),
                                 gx, gy, 0.5, 0.3)
    for g, w in ((bx, wx), (by, wy)):
        g = g.cpu().numpy()
        assert np.array_equal(np.isnan(g), np.isnan(w))
        ok = ~np.isnan(w)
        assert np.max(np.abs(g[ok] - w[ok])) <= 4e-16


def test_navigation_to_target_room_and_a_step():
    """The room's field built on the device drives the fused step: its
    nearest-node direction maps (reference mode) against the oracle step
    reading the same maps."""
    from crowddynamics_amd.stepper import Stepper
    from tests.helpers import assert_vec_close
    walls, exits = _room_geometry()
    mg = F.meshgrid(0.1, 0.0, 0.0, 200.0, 100.0)
    dmap, (ex, ey) = F.navigation_to_target(mg, exits, walls, radius=0.5, strength=0.3)
    ex, ey = ex.cpu().numpy(), ey.cpu().numpy()
    inside = np.isfinite(ex)
    say('room field: %d of %d nodes with a direction' % (inside.sum(), inside.size))
    assert inside.mean() > 0.98
    maps = np.stack([ex, ey], -1)[None]
    nav = dict(maps=np.ascontiguousarray(maps), mode=0, minx=mg.minx, miny=mg.miny,
               step=mg.step)
    scn = E.room()
    scn = E.Scenario(scn.name, scn.agents, scn.segments, scn.cell_size, scn.dt_min, scn.dt_max,
                     scn.steps, seed=scn.seed, nav=nav, bounds=scn.bounds)
    ref = scn.agents.copy()
    sc = O.Scenario(cell_size=scn.cell_size, dt_min=scn.dt_min, dt_max=scn.dt_max,
                    seed=scn.seed, segments=scn.segments, nav_maps=nav['maps'], nav_mode=0,
                    nav_origin=(mg.minx, mg.miny), nav_step=mg.step)
    O.run(ref, sc, 2)
    st = Stepper(scn.agents.copy(), segments=scn.segments, nav=nav, cell_size=scn.cell_size,
                 dt_min=scn.dt_min, dt_max=scn.dt_max, seed=scn.seed, bounds=scn.bounds)
    st.step(2)
    st.check()
    out = st.download(scn.agents.copy())
    for f in ('position', 'velocity', 'target_direction'):
        r = ref[f]
        assert_vec_close(out[f], r, np.hypot(*r.T) + 1.0, 1e-12, f)
