"""Navigation fields built on the device.

The reference builds, per target, a direction map over a node grid of the
domain (crowddynamics/simulation/field.py:131-164, navigation_to_target):

1. shortest_path (core/steering/quickest_path.py:183-197): the obstacles
   grown by `radius` are excluded, the distance to the target is solved by
   fast marching (distance_map, :54-115), turned into directions
   (direction_map, :125-149), and the excluded nodes between the grown and
   the bare obstacles take the direction of the nearest node that has one
   (fill_missing, :152-166);
2. direction_map_obstacles (core/steering/obstacle_handling.py:104-110): the
   distance to the obstacles and its directions;
3. obstacle_handling (obstacle_handling.py:15-74): within `radius` of the
   obstacles the directions bend away from them.

Geometry is rasterised on the host exactly where the reference rasterises
it (draw_geom, core/geometry.py:105-130): line segments by Bresenham's line
between their truncated node indices (skimage.draw.line); grown obstacles
as the nodes within `radius` of a segment (shapely's buffer, whose round
caps are polygons of 16 segments per quarter circle, differs only within
a node of the caps).  Every grid computation then runs in libcdamd.so on
the device: the Eikonal solve (cd_distance_map), the direction maps, the
fill and the blending.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .device import require_gpu, stream_handle
from .exceptions import InvalidValue


@dataclass
class MeshGrid:
    """meshgrid(step, minx, miny, maxx, maxy) (quickest_path.py:22-49):
    nodes at minx + step * col, miny + step * row with inclusive ends."""
    step: float
    minx: float
    miny: float
    maxx: float
    maxy: float

    @property
    def shape(self):
        cols = len(np.arange(self.minx, self.maxx + self.step, self.step))
        rows = len(np.arange(self.miny, self.maxy + self.step, self.step))
        return rows, cols

    def indicer(self, position):
        """Node index (x index, y index) of positions, truncated toward zero."""
        shifted = np.asarray(position, dtype=np.float64) - np.array((self.minx, self.miny))
        return (shifted / self.step).astype(np.int64)


def meshgrid(step, minx, miny, maxx, maxy):
    return MeshGrid(float(step), float(minx), float(miny), float(maxx), float(maxy))


def line_nodes(c0, r0, c1, r1):
    """Nodes of Bresenham's line from node (c0, r0) to (c1, r1), both ends
    included: the major axis advances one node per step and the minor axis
    follows the error term (ties step the minor axis)."""
    dc, dr = abs(c1 - c0), abs(r1 - r0)
    sc = 1 if c1 >= c0 else -1
    sr = 1 if r1 >= r0 else -1
    steep = dr > dc
    if steep:
        dc, dr = dr, dc
    n = dc + 1
    k = np.arange(n, dtype=np.int64)
    # minor-axis offset after k major steps: floor((2 dr k + dc) / (2 dc))
    minor = (2 * dr * k + dc) // (2 * dc) if dc > 0 else np.zeros(n, dtype=np.int64)
    if steep:
        return c0 + sc * minor, r0 + sr * k
    return c0 + sc * k, r0 + sr * minor


def draw_segments(grid, mgrid, segments, value):
    """Draws segments ([W, 4]: x0, y0, x1, y1) onto grid [rows, cols]."""
    rows, cols = grid.shape
    for x0, y0, x1, y1 in np.asarray(segments, dtype=np.float64).reshape(-1, 4):
        (c0, r0), (c1, r1) = mgrid.indicer([(x0, y0), (x1, y1)])
        cc, rr = line_nodes(int(c0), int(r0), int(c1), int(r1))
        ok = (rr >= 0) & (rr < rows) & (cc >= 0) & (cc < cols)
        grid[rr[ok], cc[ok]] = value
    return grid


def grown_mask(mgrid, segments, radius):
    """Nodes within `radius` of any segment (the obstacles' buffer)."""
    rows, cols = mgrid.shape
    out = np.zeros((rows, cols), dtype=bool)
    segs = np.asarray(segments, dtype=np.float64).reshape(-1, 4)
    for x0, y0, x1, y1 in segs:
        lo_c = max(0, int(np.floor((min(x0, x1) - radius - mgrid.minx) / mgrid.step)) - 1)
        hi_c = min(cols - 1, int(np.ceil((max(x0, x1) + radius - mgrid.minx) / mgrid.step)) + 1)
        lo_r = max(0, int(np.floor((min(y0, y1) - radius - mgrid.miny) / mgrid.step)) - 1)
        hi_r = min(rows - 1, int(np.ceil((max(y0, y1) + radius - mgrid.miny) / mgrid.step)) + 1)
        if lo_c > hi_c or lo_r > hi_r:
            continue
        cc, rr = np.meshgrid(np.arange(lo_c, hi_c + 1), np.arange(lo_r, hi_r + 1))
        px = mgrid.minx + mgrid.step * cc
        py = mgrid.miny + mgrid.step * rr
        dx, dy = x1 - x0, y1 - y0
        l2 = dx * dx + dy * dy
        t = np.zeros_like(px) if l2 == 0 else np.clip(((px - x0) * dx + (py - y0) * dy) / l2,
                                                      0.0, 1.0)
        d = np.hypot(px - (x0 + t * dx), py - (y0 + t * dy))
        out[rr, cc] |= d <= radius
    return out


class FieldBuilder:
    """Device buffers and calls of the field construction for one grid."""

    def __init__(self, mgrid, device='cuda'):
        require_gpu()
        self.mgrid = mgrid
        self.rows, self.cols = mgrid.shape
        self.device = torch.device(device)
        L = _lib.lib()
        nbytes = L.cd_field_work_bytes(self.rows, self.cols)
        if nbytes < 0:
            raise InvalidValue('grid of %d x %d nodes is not supported' % (self.rows, self.cols))
        self.work = torch.empty(int(nbytes), dtype=torch.uint8, device=self.device)

    def _f64(self, a):
        return torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64), device=self.device)

    def distance_map(self, phi, mask=None):
        """cd_distance_map: signed distance to the zero contour of phi."""
        phi_d = phi if isinstance(phi, torch.Tensor) else self._f64(phi)
        mask_d = None
        if mask is not None:
            mask_d = torch.as_tensor(np.ascontiguousarray(mask, dtype=np.uint8),
                                     device=self.device)
        out = torch.empty((self.rows, self.cols), dtype=torch.float64, device=self.device)
        _lib.check(_lib.lib().cd_distance_map(
            C.c_void_p(phi_d.data_ptr()),
            C.c_void_p(mask_d.data_ptr()) if mask_d is not None else None,
            self.rows, self.cols, self.mgrid.step, C.c_void_p(out.data_ptr()),
            C.c_void_p(self.work.data_ptr()), self.work.numel(), stream_handle()),
            'distance_map')
        return out

    def direction_map(self, dmap):
        ex = torch.empty_like(dmap)
        ey = torch.empty_like(dmap)
        _lib.check(_lib.lib().cd_direction_map(C.c_void_p(dmap.data_ptr()), self.rows,
                                               self.cols, C.c_void_p(ex.data_ptr()),
                                               C.c_void_p(ey.data_ptr()), stream_handle()),
                   'direction_map')
        return ex, ey

    def fill_missing(self, ex, ey, fill, reach):
        fill_d = torch.as_tensor(np.ascontiguousarray(fill, dtype=np.uint8), device=self.device)
        ox = torch.empty_like(ex)
        oy = torch.empty_like(ey)
        _lib.check(_lib.lib().cd_fill_missing(
            C.c_void_p(ex.data_ptr()), C.c_void_p(ey.data_ptr()), C.c_void_p(fill_d.data_ptr()),
            self.rows, self.cols, int(reach), C.c_void_p(ox.data_ptr()),
            C.c_void_p(oy.data_ptr()), stream_handle()), 'fill_missing')
        return ox, oy

    def obstacle_handling(self, dobs, ox, oy, tx, ty, radius, strength):
        """In place on (tx, ty)."""
        _lib.check(_lib.lib().cd_obstacle_handling(
            C.c_void_p(dobs.data_ptr()), C.c_void_p(ox.data_ptr()), C.c_void_p(oy.data_ptr()),
            C.c_void_p(tx.data_ptr()), C.c_void_p(ty.data_ptr()), dobs.numel(), float(radius),
            float(strength), stream_handle()), 'obstacle_handling')
        return tx, ty


def target_phi(mgrid, targets):
    """phi of distance_map: +1 on the target's nodes, -1 elsewhere."""
    phi = np.full(mgrid.shape, -1.0)
    draw_segments(phi, mgrid, targets, 1.0)
    return phi


def navigation_to_target(mgrid, targets, obstacles=None, radius=0.5, strength=0.3,
                         device='cuda'):
    """Field.navigation_to_target (simulation/field.py:156-164) on the
    device for one target (segments [W, 4]) and optional obstacle segments:
    returns (distance map of the target, (ex, ey) direction map), device
    tensors [rows, cols]."""
    fb = FieldBuilder(mgrid, device)
    phi = target_phi(mgrid, targets)
    have_obs = obstacles is not None and len(np.asarray(obstacles).reshape(-1, 4))
    grown = grown_mask(mgrid, obstacles, radius) if have_obs else None
    dmap_t = fb.distance_map(phi, grown)
    tx, ty = fb.direction_map(dmap_t)
    if not have_obs:
        return dmap_t, (tx, ty)
    bare = np.zeros(mgrid.shape, dtype=bool)
    draw_segments(bare, mgrid, obstacles, True)
    # every node without a direction (the grown band, and the nodes whose
    # gradient reads it) except the obstacles' own takes the direction of
    # the nearest node that has one (fill_missing fills the mask of the
    # direction map minus the obstacles)
    reach = int(np.ceil(radius / mgrid.step)) + 3
    tx, ty = fb.fill_missing(tx, ty, ~bare, reach)
    dmap_o = fb.distance_map(target_phi(mgrid, obstacles))
    ox, oy = fb.direction_map(dmap_o)
    fb.obstacle_handling(dmap_o, ox, oy, tx, ty, radius, strength)
    return dmap_t, (tx, ty)


def navigation_maps(mgrid, targets, obstacles=None, radius=0.5, strength=0.3, mode=0,
                    device='cuda'):
    """The `nav` dictionary of a Stepper (one map per target) from fields
    built on the device: mode 0, the direction maps [T, rows, cols, 2]
    (the reference's nearest-node lookup); mode 1, the potentials
    [T, rows, cols] = distance to the target (bilinear-gradient lookup).
    `targets` is a list of segment arrays, one per target."""
    maps = []
    for tgt in targets:
        dmap, (ex, ey) = navigation_to_target(mgrid, tgt, obstacles, radius, strength, device)
        if mode == 0:
            maps.append(torch.stack([ex, ey], -1).cpu().numpy())
        else:
            maps.append((-dmap).cpu().numpy())
    return dict(maps=np.ascontiguousarray(np.stack(maps)), mode=int(mode), minx=mgrid.minx,
                miny=mgrid.miny, step=mgrid.step)
