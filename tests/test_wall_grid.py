"""The host-side wall grid (cd_wall_grid_build, no device call): for any
point, every segment whose bounding box lies within an agent's reach is
listed in the point's cell, and a point outside the grid is out of reach of
every segment.  The force kernel then visits the same segments as a scan of
all of them would (the rest contribute exactly zero)."""
import ctypes as C

import numpy as np
import pytest

from crowddynamics_amd import _lib
from crowddynamics_amd import examples as E


def _grid(segs, reach, cell):
    p = _lib.CDStepParams()
    L = _lib.lib()
    ptr = C.c_void_p(segs.ctypes.data)
    need = L.cd_wall_grid_build(ptr, len(segs), reach, cell, None, 0, C.byref(p))
    assert need > 0
    out = np.zeros(need, dtype=np.int32)
    assert L.cd_wall_grid_build(ptr, len(segs), reach, cell, C.c_void_p(out.ctypes.data), need,
                                C.byref(p)) == need
    ncell = p.wall_nx * p.wall_ny
    return p, out[:ncell + 1], out[ncell + 1:]


@pytest.mark.parametrize('cell', [0.3, 1.0, 2.5])
def test_wall_grid_lists_every_segment_in_reach(cell):
    scn = E.corridor(n=200, width=80.0, height=40.0, n_obstacles=40)
    segs = np.ascontiguousarray(np.concatenate([scn.segments['p0'], scn.segments['p1']], 1))
    reach = 0.25
    p, off, ids = _grid(segs, reach, cell)
    assert off[0] == 0 and np.all(np.diff(off) >= 0) and off[-1] == len(ids)
    assert ids.min() >= 0 and ids.max() < len(segs)
    for c in range(len(off) - 1):  # each cell lists its segments once, in index order
        assert np.all(np.diff(ids[off[c]:off[c + 1]]) > 0)
    rng = np.random.default_rng(1)
    pts = rng.uniform((-2.0, -2.0), (82.0, 42.0), (20000, 2))
    # points on cell boundaries too
    pts[:2000] = np.stack([p.wall_minx + cell * rng.integers(0, p.wall_nx, 2000),
                           p.wall_miny + cell * rng.integers(0, p.wall_ny, 2000)], 1)
    lo = np.minimum(segs[:, 0:2], segs[:, 2:4])
    hi = np.maximum(segs[:, 0:2], segs[:, 2:4])
    d = np.maximum(np.maximum(lo[None] - pts[:, None], pts[:, None] - hi[None]), 0.0)
    near = (d ** 2).sum(-1) <= (reach + 1e-6) ** 2
    qx = (pts[:, 0] - p.wall_minx) / p.wall_cell
    qy = (pts[:, 1] - p.wall_miny) / p.wall_cell
    inside = (qx >= 0) & (qy >= 0) & (qx < p.wall_nx) & (qy < p.wall_ny)
    for k in range(len(pts)):
        want = np.flatnonzero(near[k])
        if not inside[k]:
            assert len(want) == 0
            continue
        c = int(qx[k]) * p.wall_ny + int(qy[k])
        assert np.all(np.isin(want, ids[off[c]:off[c + 1]]))


def test_wall_grid_arguments():
    L = _lib.lib()
    p = _lib.CDStepParams()
    segs = np.zeros((1, 4))
    ptr = C.c_void_p(segs.ctypes.data)
    assert L.cd_wall_grid_build(ptr, 0, 0.25, 1.0, None, 0, C.byref(p)) == _lib.CD_ERR_ARG
    assert L.cd_wall_grid_build(ptr, 1, 0.25, 0.0, None, 0, C.byref(p)) == _lib.CD_ERR_ARG
    segs[0] = (0.0, 0.0, np.nan, 1.0)
    assert L.cd_wall_grid_build(ptr, 1, 0.25, 1.0, None, 0, C.byref(p)) == _lib.CD_ERR_ARG
    segs[0] = (0.0, 0.0, 1e9, 1.0)  # far too many cells
    assert L.cd_wall_grid_build(ptr, 1, 0.25, 1e-3, None, 0, C.byref(p)) == _lib.CD_ERR_ARG
    segs[0] = (0.0, 0.0, 3.0, 0.0)  # a capacity too small: sizes only
    need = L.cd_wall_grid_build(ptr, 1, 0.25, 1.0, None, 0, C.byref(p))
    out = np.full(need, -7, dtype=np.int32)
    assert L.cd_wall_grid_build(ptr, 1, 0.25, 1.0, C.c_void_p(out.ctypes.data), need - 1,
                                C.byref(p)) == need
    assert np.all(out == -7)
